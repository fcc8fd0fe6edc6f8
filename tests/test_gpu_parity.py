"""HIP path (through the C ABI) against the oracle on identical seeded inputs.

Tolerances: integer tables bit-exact; lookup table bit-exact; floating-point
outputs rel-L2 <= 1e-5 (atomics reorder float sums)."""

import math

import numpy as np
import pytest

from gpunufft_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a ROCm device"
    return torch


@pytest.fixture(scope="module")
def capi():
    from gpunufft_amd import capi as C
    C.load()
    return C


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.cpu().numpy()


CASES = [
    # name, traj, img_dims, osf, kw, sw
    ("radial2d_64", lambda: W.radial_2d(64, 128), (64, 64), 2.0, 3, 8),
    ("golden2d_aniso", lambda: W.radial_2d(40, 96, golden=True), (48, 32), 2.0, 3, 8),
    ("rand2d_odd_grid", lambda: np.random.default_rng(5).uniform(-0.5, 0.5, (2, 3000)).astype(np.float32),
     (37, 29), 1.5, 3, 8),
    ("rand2d_kw5_sw5", lambda: np.random.default_rng(6).uniform(-0.5, 0.5, (2, 2000)).astype(np.float32),
     (40, 40), 1.25, 5, 5),
    ("koosh3d_small", lambda: W.koosh_3d(256, 40), (24, 24, 24), 1.5, 3, 8),
    ("rand3d_aniso_odd", lambda: np.random.default_rng(7).uniform(-0.5, 0.5, (3, 4000)).astype(np.float32),
     (20, 18, 11), 1.5, 3, 8),
    ("rand3d_kw4_sw6", lambda: np.random.default_rng(8).uniform(-0.5, 0.5, (3, 1500)).astype(np.float32),
     (16, 16, 16), 2.0, 4, 6),
    # the sector widths the autotuner chooses from (4, 8, 16): P = 6 and 18
    ("spiral2d_sw16", lambda: W.spiral_2d(12, 1500, n_turns=9.6), (96, 96), 2.0, 3, 16),
    ("spiral2d_sw4", lambda: W.spiral_2d(8, 1200, n_turns=6.4), (64, 64), 2.0, 3, 4),
    ("koosh3d_sw16", lambda: W.koosh_3d(400, 48), (32, 32, 32), 1.5, 3, 16),
    ("rand3d_sw4", lambda: np.random.default_rng(9).uniform(-0.5, 0.5, (3, 3000)).astype(np.float32),
     (20, 20, 20), 1.5, 3, 4),
]
SW_CASES = CASES[7:]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_precompute_tables_bit_exact(oracle, torch_dev, capi, case):
    name, tf, img, osf, kw, sw = case
    traj = tf()
    dens = W.ramp_density(traj)
    print("case", name, flush=True)
    pl = capi.Plan(img, kw, sw, osf)
    pl.set_trajectory(dev(torch_dev, traj), dev(torch_dev, dens))
    t = pl.tables()
    ref = oracle.precompute(traj, img, osf, kw, sw, dens)
    stable = oracle.precompute(traj, img, osf, kw, sw, dens, order="stable")
    assert np.array_equal(pl.assigned_sectors(), ref["sectors"])
    assert np.array_equal(t["sector_data_count"], ref["sector_data_count"])
    assert np.array_equal(t["sector_centers"], ref["sector_centers"])
    # the getters report the reference's std::sort permutation, bit-exact
    assert np.array_equal(t["data_indices"], ref["data_indices"])
    assert np.array_equal(t["traj_sorted"], ref["traj_sorted"])
    assert np.array_equal(t["dens_sorted"], ref["dens_sorted"])
    # sectors of more than 16 samples are where std::sort leaves the stable order
    sdc = ref["sector_data_count"]
    if np.max(np.diff(sdc)) > 16:
        assert not np.array_equal(ref["data_indices"], stable["data_indices"])


@pytest.mark.parametrize("lookup", [0, 1])
@pytest.mark.parametrize("img,osf,kw", [((64, 64), 2.0, 3), ((32, 32, 32), 1.5, 3),
                                        ((30, 20), 1.25, 5), ((16, 16, 16), 1.0, 4)])
def test_lut_and_deapo(oracle, torch_dev, capi, img, osf, kw, lookup):
    pl = capi.Plan(img, kw, 8, osf, lookup)
    pl.set_trajectory(dev(torch_dev, np.zeros((len(img), 1), np.float32)))
    t = pl.tables()
    assert np.array_equal(t["lut"], oracle.load_lut(osf, kw, lookup))
    ref = oracle.compute_deapo(img, kw, osf, lookup)
    got = t["deapo"].reshape(ref.shape)
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 2e-5


def _gridding_case(oracle, torch_dev, capi, traj, img, osf, kw, sw, C, lookup, seed):
    rng = np.random.default_rng(seed)
    pl = capi.Plan(img, kw, sw, osf, lookup)
    pl.set_trajectory(dev(torch_dev, traj))
    pre = oracle.precompute(traj, img, osf, kw, sw, lookup=lookup, order="stable")
    gi = pre["info"]
    lut = oracle.load_lut(osf, kw, lookup)
    M = traj.shape[1]
    data = W.complex_normal(rng, (C, M))
    g_ref = oracle.spread(gi, lookup, lut, pre["traj_sorted"], data, pre["sector_data_count"],
                          pre["sector_centers"])
    grid = torch_dev.empty((C,) + g_ref.shape[1:], dtype=torch_dev.complex64, device="cuda")
    pl.spread_sorted(dev(torch_dev, data), grid, C)
    assert oracle.rel_l2(host(grid), g_ref) < TOL
    gin = W.complex_normal(rng, g_ref.shape)
    s_ref = oracle.interp(gi, lookup, lut, pre["traj_sorted"], gin, pre["sector_data_count"],
                          pre["sector_centers"])
    out = torch_dev.empty((C, M), dtype=torch_dev.complex64, device="cuda")
    pl.interp_sorted(dev(torch_dev, gin), out, C)
    assert oracle.rel_l2(host(out), s_ref) < TOL


@pytest.mark.parametrize("lookup", [0, 1])
@pytest.mark.parametrize("C", [1, 3, 8, 11])
@pytest.mark.parametrize("case", [CASES[0], CASES[2], CASES[3], CASES[4], CASES[5], CASES[6]] + SW_CASES,
                         ids=lambda c: c[0])
def test_spread_interp_kernels(oracle, torch_dev, capi, case, C, lookup):
    name, tf, img, osf, kw, sw = case
    print("case", name, C, lookup, flush=True)
    _gridding_case(oracle, torch_dev, capi, tf(), img, osf, kw, sw, C, lookup, seed=C)


def _kw1_tap_samples(traj, img, osf, kw, sw):
    """Samples with kw + 1 taps along some axis (a position exactly half-way
    between grid points): the tap range of the sample's own sector, in the
    reference's float32 arithmetic (CUDA/inc/gpuNUFFT_utils.hpp:89-104,
    194-199; E3 sector, CUDA/inc/precomp_utils.hpp:34-48)."""
    P = sw + 2 * (kw // 2)
    off, r = P // 2, np.float32(kw / 2.0)
    wide = np.zeros(traj.shape[1], bool)
    for a in range(traj.shape[0]):
        G = int(math.floor(img[a] * osf))
        Gf = np.float32(G)
        pos = traj[a].astype(np.float32)
        v = (pos * Gf).astype(np.float64) + 0.5 * float(Gf)
        sec = np.clip(np.floor(np.round(v) / float(np.float32(sw))).astype(int), 0,
                      int(math.ceil(G / sw)) - 1)
        c = (sec * sw + sw // 2).astype(np.float32)
        u = (pos * Gf + np.float32(0.5) * Gf - c + np.float32(off)).astype(np.float32)
        wide |= (np.floor(u + r) - np.ceil(u - r) + 1) > kw
    return int(wide.sum())


@pytest.mark.parametrize("lookup", [0, 1])
@pytest.mark.parametrize("C", [1, 4, 5])
@pytest.mark.parametrize("case", [
    ("radial2d_osf15", lambda: W.radial_2d(64, 128), (64, 64), 1.5, 3, 8),
    ("axis3d_osf15", lambda: W.koosh_3d(6, 96), (32, 32, 32), 1.5, 3, 8)], ids=lambda c: c[0])
def test_half_grid_positions(oracle, torch_dev, capi, case, C, lookup):
    """Samples exactly half-way between grid points take kw + 1 taps along
    that axis (the last with weight zero in nearest mode); the kernels run
    such waves with four taps per axis. Spokes along the axes at osf 1.5
    land on such positions every fourth sample: the case holds them (checked
    first), and spread and interp equal the oracle (rel-L2 <= 1e-5)."""
    name, tf, img, osf, kw, sw = case
    traj = tf()
    if name.startswith("axis3d"):
        # spokes along x, y and z through the centre, plus the Fibonacci set
        r = np.linspace(-0.5, 0.5, 96, endpoint=False).astype(np.float32)
        z = np.zeros_like(r)
        axes = np.concatenate([np.stack([r, z, z]), np.stack([z, r, z]), np.stack([z, z, r])], 1)
        traj = np.concatenate([axes, traj], 1).astype(np.float32)
    assert _kw1_tap_samples(traj, img, osf, kw, sw) >= 16
    print("case", name, C, lookup, flush=True)
    _gridding_case(oracle, torch_dev, capi, traj, img, osf, kw, sw, C, lookup, seed=C + 3)


@pytest.mark.parametrize("n_heavy", [6000, 9000])
def test_heavy_sector_chunks(oracle, torch_dev, capi, n_heavy):
    """Thousands of samples in one sector (split over many workgroups): the
    central sector holds 6007 or 9010 samples, 6 or 9 chunks of 1024, so the
    3-D interp's runs of up to four chunks end full and partial (4 + 2,
    4 + 4 + 1)."""
    rng = np.random.default_rng(1)
    traj = np.concatenate([rng.uniform(-0.01, 0.01, (3, n_heavy)),
                           rng.uniform(-0.5, 0.5, (3, 500))], axis=1).astype(np.float32)
    print("case heavy", n_heavy, flush=True)
    _gridding_case(oracle, torch_dev, capi, traj, (16, 16, 16), 2.0, 3, 8, 8, 0, seed=2)


def _op_pair(oracle, torch_dev, capi, traj, img, osf, kw, sw, C, lookup, dens, sens):
    pl = capi.Plan(img, kw, sw, osf, lookup)
    pl.set_trajectory(dev(torch_dev, traj), None if dens is None else dev(torch_dev, dens))
    if sens is not None:
        pl.set_sens(dev(torch_dev, sens), C)
    op = oracle.Operator(traj, img, kw, sw, osf, dens=dens, sens=sens, lookup=lookup,
                         order="stable")
    return pl, op


@pytest.mark.parametrize("lookup", [0, 1])
@pytest.mark.parametrize("use_dens,use_sens", [(False, False), (True, False), (True, True)])
@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[2], CASES[4], CASES[5]] + SW_CASES,
                         ids=lambda c: c[0])
def test_operator_all_stages(oracle, torch_dev, capi, case, use_dens, use_sens, lookup):
    name, tf, img, osf, kw, sw = case
    print("case", name, use_dens, use_sens, lookup, flush=True)
    traj = tf()
    M = traj.shape[1]
    C = 4
    rng = np.random.default_rng(42)
    dens = W.ramp_density(traj) if use_dens else None
    sens = W.gaussian_sens(img, C, seed=3) if use_sens else None
    pl, op = _op_pair(oracle, torch_dev, capi, traj, img, osf, kw, sw, C, lookup, dens, sens)
    gi = op.gi
    kd = W.complex_normal(rng, (C, M))
    ishape = oracle.img_shape(gi)
    gshape = oracle.grid_shape(gi)
    k_dev = dev(torch_dev, kd)
    for stop in (oracle.CONVOLUTION, oracle.FFT, oracle.DEAPODIZATION):
        ref = op.adjoint(kd, stop=stop)
        if stop == oracle.CONVOLUTION:
            shape = (C,) + gshape
        elif stop == oracle.DEAPODIZATION and use_sens:
            shape = (1,) + ishape
        else:
            shape = (C,) + ishape
        if stop == oracle.FFT and use_sens:
            ref = op.adjoint(kd, stop=stop)  # per coil
        out = torch_dev.empty(shape, dtype=torch_dev.complex64, device="cuda")
        pl.adjoint(k_dev, out, C, stop)
        assert oracle.rel_l2(host(out), ref) < TOL, ("adjoint", stop)
    x = W.complex_normal(rng, ((1,) if use_sens else (C,)) + ishape)
    ref = op.forward(x)
    out = torch_dev.empty((C, M), dtype=torch_dev.complex64, device="cuda")
    pl.forward(dev(torch_dev, x), out, C)
    assert oracle.rel_l2(host(out), ref) < TOL, "forward"


@pytest.mark.parametrize("case", [CASES[0], CASES[4], CASES[7]], ids=lambda c: c[0])
def test_workgroups_per_cu_same_result(oracle, torch_dev, capi, case):
    """The workgroups-per-CU target changes the coil block of the
    kernel-width-3 kernels (4, 2 or 1 coils), not the result: the
    sub-grid sums are exact integers and the interp sums run in a fixed
    order; only the float L1 behind each chunk's fixed-point scale is summed
    over other lanes, so outputs agree with the default's to rel-L2 1e-6.
    Out-of-range targets are refused, and the autotuner keeps one of its
    candidates."""
    from gpunufft_amd import autotune

    name, tf, img, osf, kw, sw = case
    traj = tf()
    M = traj.shape[1]
    C = 5
    rng = np.random.default_rng(3)
    kd = dev(torch_dev, W.complex_normal(rng, (C, M)))
    pl = capi.Plan(img, kw, sw, osf, 0)
    pl.set_trajectory(dev(torch_dev, traj), dev(torch_dev, W.ramp_density(traj)))
    info = pl.info()
    ishape = tuple(int(x) for x in info.img[:len(img)][::-1])
    x = dev(torch_dev, W.complex_normal(rng, (C,) + ishape))

    def run():
        img_out = torch_dev.empty((C,) + ishape, dtype=torch_dev.complex64, device="cuda")
        k_out = torch_dev.empty((C, M), dtype=torch_dev.complex64, device="cuda")
        pl.adjoint(kd, img_out, C, oracle.DEAPODIZATION)
        pl.forward(x, k_out, C)
        return host(img_out), host(k_out)

    ref = run()
    for n in (1, 2, 3, 4):
        pl.set_workgroups_per_cu(n)
        got = run()
        assert oracle.rel_l2(got[0], ref[0]) < 1e-6 and oracle.rel_l2(got[1], ref[1]) < 1e-6, (name, n)
    for bad in (-1, 9):
        with pytest.raises(capi.GnufftError):
            pl.set_workgroups_per_cu(bad)
    assert autotune.best_workgroups_per_cu(pl, C) in (2, 3, 4)
    assert oracle.rel_l2(run()[0], ref[0]) < 1e-6


# 3-D grids of 96 points along z (compile-time FFT kernels for the z
# pass): adjoint FFT and DEAPODIZATION stops with density and coil maps,
# partial and full coil blocks, sector widths 8 and 16, an anisotropic grid
FOLD_FFT_CASES = [
    ("koosh3d_64_sw8", lambda: W.koosh_3d(900, 64), (64, 64, 64), 1.5, 3, 8),
    ("koosh3d_64_sw16", lambda: W.koosh_3d(900, 64), (64, 64, 64), 1.5, 3, 16),
    ("rand3d_aniso_z96", lambda: np.random.default_rng(11).uniform(-0.5, 0.5, (3, 30000)).astype(np.float32),
     (40, 48, 64), 1.5, 3, 8),
]


@pytest.mark.parametrize("C", [3, 8])
@pytest.mark.parametrize("case", FOLD_FFT_CASES, ids=lambda c: c[0])
def test_adjoint_fixed_z_pass(oracle, torch_dev, capi, case, C):
    name, tf, img, osf, kw, sw = case
    print("case", name, C, flush=True)
    traj = tf()
    M = traj.shape[1]
    rng = np.random.default_rng(C)
    dens = W.ramp_density(traj)
    sens = W.gaussian_sens(img, C, seed=4)
    kd = W.complex_normal(rng, (C, M))
    k_dev = dev(torch_dev, kd)
    for use_sens in (False, True):
        pl, op = _op_pair(oracle, torch_dev, capi, traj, img, osf, kw, sw, C, 0, dens,
                          sens if use_sens else None)
        ishape = oracle.img_shape(op.gi)
        for stop in (oracle.FFT, oracle.DEAPODIZATION):
            ref = op.adjoint(kd, stop=stop)
            per_coil = not (use_sens and stop == oracle.DEAPODIZATION)
            out = torch_dev.empty(((C,) if per_coil else (1,)) + ishape, dtype=torch_dev.complex64,
                                  device="cuda")
            pl.adjoint(k_dev, out, C, stop)
            assert oracle.rel_l2(host(out), ref) < TOL, (name, C, use_sens, stop)


@pytest.mark.parametrize("case", [CASES[1], CASES[4], CASES[7]], ids=lambda c: c[0])
def test_load_precomputed_std_order(oracle, torch_dev, capi, case):
    """Precomputed tables in the reference's std::sort order (as a
    MATLAB-precomputed operator supplies them,
    gpuNUFFT_operator_factory.cpp:576-626) load, and the loaded operator's
    adjoint and forward match the oracle's (rel-L2 <= 1e-5); the getters
    return the loaded tables unchanged."""
    name, tf, img, osf, kw, sw = case
    traj = tf()
    dens = W.ramp_density(traj)
    C = 3
    pre = oracle.precompute(traj, img, osf, kw, sw, dens, order="std")
    pl = capi.Plan(img, kw, sw, osf)
    pl.load_precomputed(pre["traj_sorted"], pre["data_indices"], pre["sector_data_count"],
                        pre["sector_centers"], pre["dens_sorted"])
    t = pl.tables()
    assert np.array_equal(t["data_indices"], pre["data_indices"])
    assert np.array_equal(t["traj_sorted"], pre["traj_sorted"])
    op = oracle.Operator(traj, img, kw, sw, osf, dens=dens, order="std")
    rng = np.random.default_rng(3)
    kd = W.complex_normal(rng, (C, traj.shape[1]))
    ishape = oracle.img_shape(op.gi)
    out = torch_dev.empty((C,) + ishape, dtype=torch_dev.complex64, device="cuda")
    pl.adjoint(dev(torch_dev, kd), out, C)
    assert oracle.rel_l2(host(out), op.adjoint(kd)) < TOL
    x = W.complex_normal(rng, (C,) + ishape)
    kout = torch_dev.empty((C, traj.shape[1]), dtype=torch_dev.complex64, device="cuda")
    pl.forward(dev(torch_dev, x), kout, C)
    assert oracle.rel_l2(host(kout), op.forward(x)) < TOL
    # malformed tables are refused before any upload
    bad = pre["data_indices"].copy()
    bad[0] = traj.shape[1]
    with pytest.raises(capi.GnufftError) as ei:
        pl.load_precomputed(pre["traj_sorted"], bad, pre["sector_data_count"], pre["sector_centers"])
    assert ei.value.code == capi.ERR_INVALID_ARGUMENT


def test_forward_density_estimation_osf1(oracle, torch_dev, capi):
    traj = W.radial_2d(16, 32)
    pl, op = _op_pair(oracle, torch_dev, capi, traj, (32, 32), 1.0, 3, 8, 2, 0, None, None)
    x = W.complex_normal(np.random.default_rng(0), (2, 32, 32))
    ref = op.forward(x, stop=oracle.DENSITY_ESTIMATION)
    out = torch_dev.empty(ref.shape, dtype=torch_dev.complex64, device="cuda")
    pl.forward(dev(torch_dev, x), out, 2, oracle.DENSITY_ESTIMATION)
    assert oracle.rel_l2(host(out), ref) < TOL


def test_reference_fixtures_on_gpu(oracle, fixtures, torch_dev, capi):
    f = fixtures["forward_2d_32_4"]
    c = np.array(f["coords_flat"], np.float32).reshape(2, 3)
    img = np.full((f["coils"], 32, 32), 1 + 1j, np.complex64)
    for lk, key in [(0, "expect_nearest"), (1, "expect_texture")]:
        pl = capi.Plan(f["img"], f["kw"], f["sw"], f["osf"], lk)
        pl.set_trajectory(dev(torch_dev, c))
        out = torch_dev.empty((f["coils"], 3), dtype=torch_dev.complex64, device="cuda")
        pl.forward(dev(torch_dev, img), out, f["coils"])
        y = host(out)
        assert np.all(np.abs(y[:, 0].real - f[key]) < f["eps"])
        assert np.all(np.abs(y[:, 0] - y[0, 0]) < 1e-5)
    f = fixtures["deapo_2d_30_b1"]
    C = f["coils"]
    pl = capi.Plan(f["img"], f["kw"], f["sw"], f["osf"], 0)
    pl.set_trajectory(dev(torch_dev, np.zeros((2, 1), np.float32)),
                      dev(torch_dev, np.ones(1, np.float32)))
    pl.set_sens(dev(torch_dev, np.full((C, 32, 32), complex(*f["sens"]), np.complex64)), C)
    out = torch_dev.empty((1, 32, 32), dtype=torch_dev.complex64, device="cuda")
    pl.adjoint(dev(torch_dev, np.full((C, 1), complex(*f["data"]), np.complex64)), out, C)
    assert abs(host(out)[0, 0, 0].real - f["expect"]) < f["eps"]


@pytest.mark.parametrize("key", ["gpu_conv", "gpu_fft_stop"])
def test_reference_gpu_known_answers(oracle, fixtures, torch_dev, capi, key):
    """The reference's GPU convolution / FFT-stage known answers
    (gpuNUFFT_conv_tests.cpp, gpuNUFFT_2D_conv_tests.cpp,
    gpuNUFFT_fft_tests.cpp, transcribed in tests/golden) on the HIP path:
    every value within the tolerance the reference asserts, and within 1e-5
    (relative to the largest grid value) of the oracle at the same point."""
    from test_oracle_fixtures import reference_gpu_fixture_run
    stop = oracle.CONVOLUTION if key == "gpu_conv" else oracle.FFT
    for f in fixtures[key]:
        print("fixture", f["src"], flush=True)
        traj = np.array(f["traj"], np.float32)
        data = np.array([complex(*v) for v in f["data"]], np.complex64)[None]
        pl = capi.Plan(tuple(f["img"]), f["kw"], f["sw"], f["osf"], 0)
        pl.set_trajectory(dev(torch_dev, traj))
        gi = oracle.make_info(tuple(f["img"]), f["osf"], f["kw"], f["sw"])
        shape = (1,) + (oracle.grid_shape(gi) if stop == oracle.CONVOLUTION else oracle.img_shape(gi))
        out = torch_dev.empty(shape, dtype=torch_dev.complex64, device="cuda")
        pl.adjoint(dev(torch_dev, data), out, 1, stop)
        got = host(out)[0]
        scale = float(np.max(np.abs(got)))
        for ck, ref in reference_gpu_fixture_run(oracle, f, stop):
            x, y, z, comp, exp, eps = ck
            v = got[z, y, x] if got.ndim == 3 else got[y, x]
            v = float(v.real if comp == 0 else v.imag)
            assert abs(v - exp) <= eps, (f["src"], ck, v)
            assert abs(v - ref) <= 1e-5 * scale, (f["src"], ck, v, ref)


def _adjointness(torch_dev, pl, img_shape, C, M, use_sens, seed):
    torch = torch_dev
    g = torch.Generator(device="cuda").manual_seed(seed)
    xs = (1,) if use_sens else (C,)
    x = torch.randn(xs + img_shape, dtype=torch.complex64, device="cuda", generator=g)
    y = torch.randn((C, M), dtype=torch.complex64, device="cuda", generator=g)
    Ax = torch.empty((C, M), dtype=torch.complex64, device="cuda")
    AHy = torch.empty(xs + img_shape, dtype=torch.complex64, device="cuda")
    pl.forward(x, Ax, C)
    pl.adjoint(y, AHy, C)
    lhs = torch.vdot(Ax.flatten().to(torch.complex128), y.flatten().to(torch.complex128))
    rhs = torch.vdot(x.flatten().to(torch.complex128), AHy.flatten().to(torch.complex128))
    return abs(complex(lhs - rhs)) / abs(complex(lhs))


@pytest.mark.parametrize("cfg", ["radial2d_256_mc16", "koosh3d_128_mc8"])
def test_full_size_against_oracle(oracle, torch_dev, capi, cfg):
    """Full workload sizes (radial 256^2 with 16 coils, koosh 128^3 with 8
    coils, sensitivities and density) against the oracle operator, run with
    OpenMP over up to 16 host threads (private grids per thread): adjoint
    and forward within rel-L2 1e-5."""
    import os

    w = W.config(cfg)
    torch = torch_dev
    traj, img, C = w["traj"], w["img_dims"], w["coils"]
    M = traj.shape[1]
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    print("config", cfg, "oracle threads", threads, flush=True)
    sens = W.gaussian_sens(img, C, seed=w["seed"])
    kd = W.complex_normal(np.random.default_rng(w["seed"]), (C, M))
    pl = capi.Plan(img, w["kw"], w["sw"], w["osf"], 0)
    pl.set_trajectory(dev(torch, traj), dev(torch, w["dens"]))
    pl.set_sens(dev(torch, sens), C)
    op = oracle.Operator(traj, img, w["kw"], w["sw"], w["osf"], dens=w["dens"], sens=sens,
                         order="stable", threads=threads)
    x_ref = op.adjoint(kd)
    k_ref = op.forward(x_ref)
    x = torch.empty(x_ref.shape, dtype=torch.complex64, device="cuda")
    pl.adjoint(dev(torch, kd), x, C)
    k = torch.empty((C, M), dtype=torch.complex64, device="cuda")
    pl.forward(dev(torch, x_ref), k, C)
    ea, ef = oracle.rel_l2(host(x), x_ref), oracle.rel_l2(host(k), k_ref)
    print("rel-L2 adjoint %.3g forward %.3g" % (ea, ef), flush=True)
    assert ea < 1e-5 and ef < 1e-5, (ea, ef)


@pytest.mark.parametrize("cfg", ["radial2d_256_mc16", "koosh3d_128_mc8"])
def test_full_size_properties(torch_dev, capi, cfg):
    """Full workload sizes (radial 256^2 with 16 coils, koosh 128^3 with 8
    coils): adjointness and linearity (size-independent checks)."""
    w = W.config(cfg)
    print("config", cfg, flush=True)
    torch = torch_dev
    traj, img, C = w["traj"], w["img_dims"], w["coils"]
    M = traj.shape[1]
    pl = capi.Plan(img, w["kw"], w["sw"], w["osf"], 0)
    pl.set_trajectory(dev(torch, traj), dev(torch, w["dens"]))
    ishape = tuple(img[::-1])
    err = _adjointness(torch, pl, ishape, C, M, False, 1)
    assert err < 1e-5, err
    sens = W.gaussian_sens(img, C, seed=w["seed"])
    pl.set_sens(dev(torch, sens), C)
    err = _adjointness(torch, pl, ishape, C, M, True, 2)
    assert err < 1e-5, err
    # linearity of the adjoint
    g = torch.Generator(device="cuda").manual_seed(3)
    y1 = torch.randn((C, M), dtype=torch.complex64, device="cuda", generator=g)
    y2 = torch.randn((C, M), dtype=torch.complex64, device="cuda", generator=g)
    a, b = 0.7 - 0.2j, -1.3 + 0.5j
    o1 = torch.empty((1,) + ishape, dtype=torch.complex64, device="cuda")
    o2, o3 = torch.empty_like(o1), torch.empty_like(o1)
    pl.adjoint(y1, o1, C)
    pl.adjoint(y2, o2, C)
    pl.adjoint((a * y1 + b * y2).contiguous(), o3, C)
    lin = torch.linalg.vector_norm(o3 - (a * o1 + b * o2)) / torch.linalg.vector_norm(o3)
    assert float(lin) < 1e-5
    torch.cuda.synchronize()
