"""Fused oversampled-grid FFT (gnufft_plan_set_fft FUSED) against the rocFFT
path and the oracle.

The fused path replaces shift -> FFT -> shift -> crop -> scale ->
deapodization -> coil sum (adjoint) and coil weight -> deapodization -> pad
-> shift -> FFT (forward) by one pruned pass per axis. Each case runs the
full adjoint (stops FFT and DEAPODIZATION, with and without sensitivities)
and forward on one plan with both FFT kinds; both must equal the oracle
(rel-L2 <= 1e-5) and each other. The grid lengths cover even and odd
lengths, radix 2/3/4/5/8 mixes, the lengths with compile-time stage tables
(96, 128, 160, 192, 256 and eight lengths with a prime factor 7 .. 19),
lines above 511 points (large LDS tiles),
anisotropic 3-D grids, the prime radices 7, 11, 13, 17 and 19, and a
length with a factor 23 (rocFFT only: set_fft(FUSED) is refused).
"""

import numpy as np
import pytest

from gpunufft_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-5

# (image dims, osf) -> grid lengths
CASES = [
    ((32, 32), 1.5),        # 48 = 8 * 2 * 3
    ((30, 30), 1.5),        # 45 = 3 * 3 * 5 (odd)
    ((50, 40), 1.5),        # 75 x 60
    ((64, 64), 2.0),        # 128 (compile-time stages)
    ((64, 48), 1.5),        # 96 x 72 (96: compile-time stages)
    ((80, 80), 2.0),        # 160 (compile-time stages, radix 5)
    ((128, 96), 2.0),       # 256 x 192 (compile-time stages, 16-line tiles)
    ((256, 160), 2.0),      # 512 x 320 (large tiles along y: 1024 threads)
    ((320, 384), 2.0),      # 640 x 768 (large tiles, radix 5 / 3)
    ((20, 24, 16), 1.5),    # 30 x 36 x 24
    ((16, 16, 16), 2.0),    # 32^3
    ((25, 16, 10), 2.0),    # 50 x 32 x 20
    ((28, 28), 1.5),        # 42 = 2 * 3 * 7 (radix 7)
    ((44, 26), 2.0),        # 88 x 52: radix 11, 13
    ((68, 76), 2.0),        # 136 x 152: radix 17, 19
    ((56, 52, 34), 2.0),    # 112 x 104 x 68: radix 7, 13, 17 in 3-D
    ((152, 152), 2.0),      # 304 = 16 * 19 (the serving mix's 3-D length, here 2-D)
    ((340, 100), 2.0),      # 680 = 8 * 5 * 17 along x (the x passes)
    ((100, 340), 2.0),      # 680 along y: a large (1024-thread) tile with a prime stage
    ((104, 152), 2.0),      # 208 x 304: compile-time stages with radix 13 / 19
    ((152, 104), 2.0),      # 304 x 208
    ((114, 136), 2.0),      # 228 x 272: compile-time radix 19 / 17
    ((112, 84), 2.0),       # 224 x 168: compile-time radix 7
    ((78, 102), 2.0),       # 156 x 204: compile-time radix 13 / 17
    ((160, 40), 2.0),       # 320 x 80: compile-time 8 * 8 * 5
    ((96, 120), 1.5),       # 144 x 180: compile-time four stages 4 * 4 * 3 * 3, 4 * 5 * 3 * 3
    ((144, 96), 1.5),       # 216 x 144: 8 * 3 * 3 * 3 along x
    ((120, 144), 2.0),      # 240 x 288: 4 * 4 * 5 * 3, 8 * 4 * 3 * 3 (16-line tiles)
    ((20, 24, 108), 2.0),   # 40 x 48 x 216: a four-stage z pass
]


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a ROCm device"
    from gpunufft_amd import capi
    capi.load()
    return torch, capi


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _traj(dims, seed):
    rng = np.random.default_rng(seed)
    d = len(dims)
    return rng.uniform(-0.5, 0.5, (d, 700 if d == 2 else 1500)).astype(np.float32)


@pytest.mark.parametrize("dims,osf", CASES, ids=[f"{'x'.join(map(str, d))}_osf{o}" for d, o in CASES])
@pytest.mark.parametrize("with_sens", [False, True], ids=["nosens", "sens"])
def test_fused_equals_rocfft_and_oracle(oracle, env, dims, osf, with_sens):
    torch, capi = env
    C = 3
    traj = _traj(dims, 1)
    dens = W.ramp_density(traj, len(dims) - 1)
    sens = W.gaussian_sens(dims, C, seed=4) if with_sens else None
    rng = np.random.default_rng(7)
    kd = W.complex_normal(rng, (C, traj.shape[1]))
    ishape = tuple(dims[::-1])
    img = W.complex_normal(rng, ((1,) if with_sens else (C,)) + ishape)
    op = oracle.Operator(traj, dims, 3, 8, osf, dens=dens, sens=sens, order="stable")
    p = capi.Plan(dims, 3, 8, osf)
    p.set_trajectory(_dev(torch, traj), _dev(torch, dens))
    if with_sens:
        p.set_sens(_dev(torch, sens), C)
    assert p.fft() == capi.FFT_FUSED
    print("case", dims, osf, "sens" if with_sens else "nosens", flush=True)
    results = {}
    for kind in (capi.FFT_FUSED, capi.FFT_ROCFFT):
        p.set_fft(kind)
        a_img = torch.empty(((1,) if with_sens else (C,)) + ishape, dtype=torch.complex64,
                            device="cuda")
        a_fft = torch.empty((C,) + ishape, dtype=torch.complex64, device="cuda")
        f_out = torch.empty((C, traj.shape[1]), dtype=torch.complex64, device="cuda")
        p.adjoint(_dev(torch, kd), a_img, C, stop=capi.DEAPODIZATION)
        p.adjoint(_dev(torch, kd), a_fft, C, stop=capi.FFT)
        p.forward(_dev(torch, img), f_out, C)
        torch.cuda.synchronize()
        results[kind] = (a_img.cpu().numpy(), a_fft.cpu().numpy(), f_out.cpu().numpy())
    ref = (op.adjoint(kd), op.adjoint(kd, stop=oracle.FFT), op.forward(img, C))
    for kind, got in results.items():
        for name, g, r in zip(("adjoint", "adjoint_fft", "forward"), got, ref):
            err = oracle.rel_l2(g, r)
            assert err < TOL, (kind, name, err)
    for j, name in enumerate(("adjoint", "adjoint_fft", "forward")):
        assert oracle.rel_l2(results[capi.FFT_FUSED][j], results[capi.FFT_ROCFFT][j]) < TOL, name


def test_unsupported_length_uses_rocfft(oracle, env):
    torch, capi = env
    dims, osf = (46, 46), 2.0  # 92 = 4 * 23
    p = capi.Plan(dims, 3, 8, osf)
    assert p.fft() == capi.FFT_ROCFFT
    with pytest.raises(capi.GnufftError) as e:
        p.set_fft(capi.FFT_FUSED)
    assert e.value.code == capi.ERR_UNSUPPORTED
    traj = _traj(dims, 2)
    p.set_trajectory(_dev(torch, traj))
    rng = np.random.default_rng(3)
    kd = W.complex_normal(rng, (2, traj.shape[1]))
    out = torch.empty((2, 46, 46), dtype=torch.complex64, device="cuda")
    p.adjoint(_dev(torch, kd), out, 2)
    torch.cuda.synchronize()
    op = oracle.Operator(traj, dims, 3, 8, osf, order="stable")
    assert oracle.rel_l2(out.cpu().numpy(), op.adjoint(kd)) < TOL


def test_rocfft_plan_pairs_are_reused(oracle, env):
    """A destroyed plan's rocFFT plans go back to a process-wide pool and
    serve the next plan of the same geometry and coil batch; two live plans
    of one geometry each hold their own pair. Every result equals the
    oracle, for both directions, across reuse and with both plans live."""
    torch, capi = env
    dims, osf, C = (46, 46), 2.0, 2  # 92 = 4 * 23: rocFFT
    traj = _traj(dims, 5)
    rng = np.random.default_rng(11)
    kd = W.complex_normal(rng, (C, traj.shape[1]))
    img = W.complex_normal(rng, (C, 46, 46))
    op = oracle.Operator(traj, dims, 3, 8, osf, order="stable")
    ref_a, ref_f = op.adjoint(kd), op.forward(img, C)

    def make():
        p = capi.Plan(dims, 3, 8, osf)
        assert p.fft() == capi.FFT_ROCFFT
        p.set_trajectory(_dev(torch, traj))
        return p

    def check(p):
        a = torch.empty((C, 46, 46), dtype=torch.complex64, device="cuda")
        f = torch.empty((C, traj.shape[1]), dtype=torch.complex64, device="cuda")
        p.adjoint(_dev(torch, kd), a, C)
        p.forward(_dev(torch, img), f, C)
        torch.cuda.synchronize()
        assert oracle.rel_l2(a.cpu().numpy(), ref_a) < TOL
        assert oracle.rel_l2(f.cpu().numpy(), ref_f) < TOL

    for _ in range(3):  # create, use, destroy: later rounds take pooled plans
        p = make()
        check(p)
        p.close()
    p, q = make(), make()  # both live: one pooled pair, one new
    check(p)
    check(q)
    check(p)
    p.close()
    q.close()


def test_fused_headline_geometry_matches_rocfft(env):
    """koosh 128^3, osf 1.5 (192^3 grid), 8 coils with sensitivities: the
    fused and rocFFT paths agree on the full adjoint and forward."""
    torch, capi = env
    dims = (128, 128, 128)
    traj = W.koosh_3d(2048, 256)  # fewer spokes than the full workload
    C = 8
    sens = W.gaussian_sens(dims, C, seed=0)
    p = capi.Plan(dims, 3, 8, 1.5)
    p.set_trajectory(_dev(torch, traj), _dev(torch, W.ramp_density(traj, 2)))
    p.set_sens(_dev(torch, sens), C)
    rng = np.random.default_rng(5)
    kd = _dev(torch, W.complex_normal(rng, (C, traj.shape[1])))
    img = _dev(torch, W.complex_normal(rng, (1,) + tuple(dims[::-1])))
    outs = {}
    for kind in (capi.FFT_FUSED, capi.FFT_ROCFFT):
        p.set_fft(kind)
        a = torch.empty((1,) + tuple(dims[::-1]), dtype=torch.complex64, device="cuda")
        f = torch.empty((C, traj.shape[1]), dtype=torch.complex64, device="cuda")
        p.adjoint(kd, a, C)
        p.forward(img, f, C)
        torch.cuda.synchronize()
        outs[kind] = (a.cpu().numpy(), f.cpu().numpy())

    def rel(x, y):
        return float(np.linalg.norm((x - y).ravel()) / np.linalg.norm(y.ravel()))

    assert rel(outs[capi.FFT_FUSED][0], outs[capi.FFT_ROCFFT][0]) < TOL
    assert rel(outs[capi.FFT_FUSED][1], outs[capi.FFT_ROCFFT][1]) < TOL


def test_fused_many_coils_long_lines(oracle, env):
    """16 coils on 512-point lines: the x passes split the coils into equal
    groups (15 lines of 513 words fill a tile); adjoint with the coil
    combine accumulating across groups, and forward."""
    torch, capi = env
    dims, osf, C = (256, 256), 2.0, 16
    traj = _traj(dims, 3)
    dens = W.ramp_density(traj, 1)
    sens = W.gaussian_sens(dims, C, seed=5)
    rng = np.random.default_rng(9)
    kd = W.complex_normal(rng, (C, traj.shape[1]))
    img = W.complex_normal(rng, (1,) + tuple(dims[::-1]))
    op = oracle.Operator(traj, dims, 3, 8, osf, dens=dens, sens=sens, order="stable")
    p = capi.Plan(dims, 3, 8, osf)
    p.set_trajectory(_dev(torch, traj), _dev(torch, dens))
    p.set_sens(_dev(torch, sens), C)
    assert p.fft() == capi.FFT_FUSED
    a = torch.empty((1,) + tuple(dims[::-1]), dtype=torch.complex64, device="cuda")
    f = torch.empty((C, traj.shape[1]), dtype=torch.complex64, device="cuda")
    p.adjoint(_dev(torch, kd), a, C)
    p.forward(_dev(torch, img), f, C)
    torch.cuda.synchronize()
    assert oracle.rel_l2(a.cpu().numpy(), op.adjoint(kd)) < TOL
    assert oracle.rel_l2(f.cpu().numpy(), op.forward(img)) < TOL
