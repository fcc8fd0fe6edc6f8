"""Batches of independent plans (gnufft_batch_*): every plan's batched
adjoint and forward equal its own single-plan adjoint and forward and the
oracle (rel-L2 <= 1e-5), for uneven image sizes, trajectory lengths, coil
counts, with and without sensitivities, in 2-D and 3-D, with the fused FFT
(one launch per pass for all plans), with rocFFT (one plan per grid size),
with both in one batch, and with plans whose grid lengths need the fused
FFT's prime stages, in 2-D and 3-D (the batched FFT kernels' prime-stage
variants)."""

import numpy as np
import pytest

from gpunufft_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a ROCm device"
    from gpunufft_amd import capi
    capi.load()
    return torch, capi


def _dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _ops_2d():
    ops = W.uneven_batch_2d(n_ops=7, seed0=500)
    # shrink to test size: fewer samples, smaller images, keep the variety
    out = []
    for j, o in enumerate(ops):
        N = [48, 64, 40, 96, 64, 32, 80][j]
        traj = W.spiral_2d(4 + j, 300 + 97 * j, n_turns=N / 10.0)
        if j == 3:  # heavy centre: sectors split over several chunks
            rng = np.random.default_rng(9)
            traj = np.concatenate([traj, rng.uniform(-0.02, 0.02, (2, 5000)).astype(np.float32)], 1)
        C = [1, 3, 4, 6, 2, 5, 8][j]
        sens = W.gaussian_sens((N, N), C, seed=j) if j % 2 == 0 else None
        out.append(dict(traj=traj, img=(N, N), C=C, sens=sens, dens=W.ramp_density(traj), sw=8))
    return out


def _ops_3d():
    out = []
    for j, (N, C) in enumerate([(16, 2), (24, 3), (20, 1)]):
        traj = W.koosh_3d(96 + 32 * j, 40)
        sens = W.gaussian_sens((N, N, N), C, seed=j) if j != 1 else None
        out.append(dict(traj=traj, img=(N, N, N), C=C, sens=sens, dens=W.ramp_density(traj, 2), sw=8))
    return out


def _ops_2d_primes():
    """Grid lengths with the fused FFT's prime stages (112 = 16 * 7,
    104 = 8 * 13) beside smooth ones: the batch runs the prime-stage
    variants of its FFT kernels for all its fused plans."""
    out = []
    for j, (N, C) in enumerate([(56, 3), (64, 2), (52, 4), (48, 1)]):
        traj = W.spiral_2d(6 + j, 400 + 50 * j, n_turns=N / 10.0)
        sens = W.gaussian_sens((N, N), C, seed=j) if j % 2 == 0 else None
        out.append(dict(traj=traj, img=(N, N), C=C, sens=sens, dens=W.ramp_density(traj), sw=8))
    return out


def _ops_3d_primes():
    """3-D grids of 42 = 2 * 3 * 7 and 39 = 3 * 13 points beside 24 (osf 1.5)."""
    out = []
    for j, (N, C) in enumerate([(28, 2), (16, 1), (26, 2)]):
        traj = W.koosh_3d(96 + 32 * j, 40)
        sens = W.gaussian_sens((N, N, N), C, seed=j) if j != 1 else None
        out.append(dict(traj=traj, img=(N, N, N), C=C, sens=sens, dens=W.ramp_density(traj, 2), sw=8))
    return out


@pytest.mark.parametrize("fft", ["fused", "rocfft", "mixed"])
@pytest.mark.parametrize("make", [_ops_2d, _ops_3d, _ops_2d_primes, _ops_3d_primes],
                         ids=["2d", "3d", "2d_primes", "3d_primes"])
def test_batch_equals_single_plans(oracle, env, make, fft):
    torch, capi = env
    specs = make()
    rng = np.random.default_rng(1)
    plans, kd, img, shapes = [], [], [], []
    for sp in specs:
        p = capi.Plan(sp["img"], 3, sp["sw"], 2.0 if len(sp["img"]) == 2 else 1.5)
        p.set_trajectory(_dev(torch, sp["traj"]), _dev(torch, sp["dens"]))
        if sp["sens"] is not None:
            p.set_sens(_dev(torch, sp["sens"]), sp["C"])
        fused = fft == "fused" or (fft == "mixed" and len(plans) % 2 == 0)
        p.set_fft(capi.FFT_FUSED if fused else capi.FFT_ROCFFT)
        plans.append(p)
        M = sp["traj"].shape[1]
        ishape = tuple(sp["img"][::-1])
        kd.append(_dev(torch, W.complex_normal(rng, (sp["C"], M))))
        img.append(_dev(torch, W.complex_normal(rng, ((1,) if sp["sens"] is not None else (sp["C"],)) + ishape)))
        shapes.append(((1,) if sp["sens"] is not None else (sp["C"],)) + ishape)
    print("batch of", len(plans), [s["img"] for s in specs], flush=True)
    b = capi.Batch(plans, [s["C"] for s in specs])
    a_out = [torch.empty(sh, dtype=torch.complex64, device="cuda") for sh in shapes]
    f_out = [torch.empty(k.shape, dtype=torch.complex64, device="cuda") for k in kd]
    b.adjoint(kd, a_out)
    b.forward(img, f_out)
    # twice: descriptors are re-uploaded per call
    b.adjoint(kd, a_out)
    torch.cuda.synchronize()
    for j, (p, sp) in enumerate(zip(plans, specs)):
        ref_a = torch.empty(shapes[j], dtype=torch.complex64, device="cuda")
        p.adjoint(kd[j], ref_a, sp["C"])
        ref_f = torch.empty(kd[j].shape, dtype=torch.complex64, device="cuda")
        p.forward(img[j], ref_f, sp["C"])
        torch.cuda.synchronize()
        assert oracle.rel_l2(a_out[j].cpu().numpy(), ref_a.cpu().numpy()) < TOL, ("adjoint", j)
        assert oracle.rel_l2(f_out[j].cpu().numpy(), ref_f.cpu().numpy()) < TOL, ("forward", j)
    # one plan against the oracle as well
    j = len(specs) - 1
    sp = specs[j]
    op = oracle.Operator(sp["traj"], sp["img"], 3, sp["sw"], 2.0 if len(sp["img"]) == 2 else 1.5,
                         dens=sp["dens"], sens=sp["sens"], order="stable")
    assert oracle.rel_l2(a_out[j].cpu().numpy(), op.adjoint(kd[j].cpu().numpy())) < TOL
    assert oracle.rel_l2(f_out[j].cpu().numpy(), op.forward(img[j].cpu().numpy())) < TOL


def test_uneven_batch_full_size_autotuned(oracle, env):
    """The benchmark's batch at its real shape: all 64 operators of the
    uneven 2-D workload (N 96-384, 4-16 coils, spirals of 8-48 x
    1024-4096 samples, sensitivities and density), sector widths chosen by
    the autotuner from {4, 8, 16} at full-GPU occupancy, with one operator
    forced to 4 and one to 8 so the batch mixes all three widths. Every
    operator's batched adjoint and forward equal its single-plan results
    (rel-L2 <= 1e-5); three operators are also checked against the oracle:
    an N = 384 operator with 16 coils, and the sw 4 and sw 8 ones."""
    torch, capi = env
    from gpunufft_amd import autotune
    ops = W.uneven_batch_2d()
    big = next(j for j, o in enumerate(ops) if o["img_dims"][0] == 384 and o["coils"] == 16)
    forced = {1 if big != 1 else 2: 4, 3 if big != 3 else 4: 8}
    plans, sws, kd, img = [], [], [], []
    for j, o in enumerate(ops):
        t_d = _dev(torch, o["traj"])
        d_d = _dev(torch, o["dens"])
        sw = forced.get(j) or autotune.best_sector_width(o["img_dims"], o["kw"], o["osf"], t_d,
                                                         d_d, o["coils"], (4, 8, 16))
        sws.append(sw)
        p = capi.Plan(o["img_dims"], o["kw"], sw, o["osf"])
        p.set_trajectory(t_d, d_d)
        p.set_sens(_dev(torch, o["sens"]), o["coils"])
        plans.append(p)
        rng = np.random.default_rng(o["seed"])
        N = o["img_dims"][0]
        kd.append(_dev(torch, W.complex_normal(rng, (o["coils"], o["traj"].shape[1]))))
        img.append(_dev(torch, W.complex_normal(rng, (1, N, N))))
    print("widths", sws, flush=True)
    assert set(sws) == {4, 8, 16} or set(sws) >= {4, 8}
    b = capi.Batch(plans, [o["coils"] for o in ops])
    a_out = [torch.empty((1,) + o["img_dims"], dtype=torch.complex64, device="cuda") for o in ops]
    f_out = [torch.empty(k.shape, dtype=torch.complex64, device="cuda") for k in kd]
    b.adjoint(kd, a_out)
    b.forward(img, f_out)
    torch.cuda.synchronize()
    for j, (p, o) in enumerate(zip(plans, ops)):
        ref_a = torch.empty_like(a_out[j])
        p.adjoint(kd[j], ref_a, o["coils"])
        ref_f = torch.empty_like(f_out[j])
        p.forward(img[j], ref_f, o["coils"])
        torch.cuda.synchronize()
        assert oracle.rel_l2(a_out[j].cpu().numpy(), ref_a.cpu().numpy()) < TOL, ("adjoint", j, sws[j])
        assert oracle.rel_l2(f_out[j].cpu().numpy(), ref_f.cpu().numpy()) < TOL, ("forward", j, sws[j])
    for j in [big] + sorted(forced):
        o = ops[j]
        print("oracle check", j, o["img_dims"], o["coils"], sws[j], flush=True)
        op = oracle.Operator(o["traj"], o["img_dims"], o["kw"], sws[j], o["osf"], dens=o["dens"],
                             sens=o["sens"], order="stable")
        assert oracle.rel_l2(a_out[j].cpu().numpy(), op.adjoint(kd[j].cpu().numpy())) < TOL, j
        assert oracle.rel_l2(f_out[j].cpu().numpy(), op.forward(img[j].cpu().numpy())) < TOL, j
    b.close()


def test_batch_rejects_mixed_dimensions(env):
    torch, capi = env
    p2 = capi.Plan((16, 16), 3, 8, 2.0)
    p2.set_trajectory(_dev(torch, W.radial_2d(4, 16)))
    p3 = capi.Plan((8, 8, 8), 3, 8, 2.0)
    p3.set_trajectory(_dev(torch, W.koosh_3d(8, 8)))
    with pytest.raises(capi.GnufftError):
        capi.Batch([p2, p3], [1, 1])


def test_autotuner_picks_a_candidate_width():
    """The sector-width autotuner returns one of its candidates, for the
    operator alone and at full-GPU occupancy, and caches per shape."""
    import torch
    from gpunufft_amd import autotune
    t = W.spiral_2d(8, 1024, n_turns=8)
    traj = torch.from_numpy(t).cuda()
    dens = torch.from_numpy(W.ramp_density(t)).cuda()
    for fill in (0, 4096):
        sw = autotune.best_sector_width((96, 96), 3, 2.0, traj, dens, 4, widths=(8, 16),
                                        fill_blocks=fill)
        assert sw in (8, 16)
        assert autotune.best_sector_width((96, 96), 3, 2.0, traj, dens, 4, widths=(8, 16),
                                          fill_blocks=fill) == sw
