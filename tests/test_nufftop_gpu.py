"""Python drop-in `gpuNUFFT.NUFFTOp` (pybind11 over the C++ gpuNUFFT:: API)
against the reference's own Python test and the oracle.

test_reference_* restate python/test_nufftOp.py:18-136 of the reference
(same image, trajectory, coil maps, arguments; the adjoint identity
<A x, y> = <x, A^H y> at rtol 1e-5). The oracle comparisons use the linear
(texture-emulating) lookup, the factory default the Python module selects;
floating point rel-L2 <= 1e-5, integer tables bit-exact.
"""

import numpy as np
import pytest

from gpunufft_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def NUFFTOp():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a ROCm device"
    from gpunufft_amd.gpuNUFFT import NUFFTOp as op
    return op


class RefSetup:
    """Inputs of python/test_nufftOp.py:31-88 (plots removed)."""

    def __init__(self):
        self.img_size = [64, 128]
        x, y = np.meshgrid(np.linspace(-1, 1, self.img_size[0]), np.linspace(-1, 1, self.img_size[1]))
        img = (x ** 2 + y ** 2 < 0.5 ** 2).T
        self.img = img.astype(np.complex64)
        n_lines = self.img_size[1]
        ns_per_line = self.img_size[0]
        self.kspace_loc = np.ones([ns_per_line * n_lines, 2])
        readout = np.linspace(-0.5, 0.5, ns_per_line)
        self.kspace_loc[:, 0] = np.tile(readout, n_lines)
        self.kspace_loc[:, 1] = np.tile(np.linspace(-0.5, 0.5, n_lines), (ns_per_line, 1)).T.reshape(-1)
        self.weights = np.ones(self.kspace_loc.shape[0])
        self.n_coils = 2
        x, y = np.meshgrid(np.linspace(0, 1, self.img_size[0]), np.linspace(0, 1, self.img_size[1]))
        c1 = ((1 / (x ** 2 + y ** 2 + 1)).T).astype(np.complex64)
        c2 = np.flip(np.flip(c1, axis=1), axis=0)
        self.multi_img = np.tile(img, (self.n_coils, 1, 1))
        self.coil_maps = np.stack([c1, c2])

    def op(self, NUFFTOp, sens_maps=None):
        return NUFFTOp(np.reshape(self.kspace_loc, self.kspace_loc.shape[::-1], order="F"),
                       self.img_size, self.n_coils, sens_maps, self.weights, 3, 8, 2, True)


def test_reference_multicoil_with_sense(NUFFTOp):
    s = RefSetup()
    operator = s.op(NUFFTOp, s.coil_maps)
    x = operator.op(np.reshape(s.img.T, s.img.size))
    y = np.random.default_rng(0).random(x.shape)
    img_adj = operator.adj_op(x)
    adj_y = np.squeeze(operator.adj_op(y)).T
    assert np.squeeze(img_adj).T.shape == s.img.shape
    np.testing.assert_allclose(np.vdot(s.img, adj_y), np.vdot(x, y), rtol=1e-5)


def test_reference_multicoil_without_sense(NUFFTOp):
    s = RefSetup()
    operator = s.op(NUFFTOp)
    x = operator.op(np.asarray([np.reshape(ch.T, ch.size) for ch in s.multi_img]).T)
    y = np.random.default_rng(1).random(x.shape)
    img_adj = np.asarray([ch.T for ch in np.squeeze(operator.adj_op(x))])
    adj_y = np.asarray([ch.T for ch in np.squeeze(operator.adj_op(y))])
    assert img_adj.shape == s.multi_img.shape
    np.testing.assert_allclose(np.vdot(s.multi_img, adj_y), np.vdot(x, y), rtol=1e-5)


CASES = [
    # traj, img (w, h[, d]), osf, kw, sw, coils, sens, balance
    (lambda: W.radial_2d(48, 96), (48, 40), 2.0, 3, 8, 3, True, True),
    (lambda: W.radial_2d(32, 64, golden=True), (32, 32), 2.0, 3, 8, 2, False, False),
    (lambda: W.koosh_3d(128, 32), (20, 20, 20), 2.0, 3, 8, 2, True, True),
    (lambda: np.random.default_rng(3).uniform(-0.5, 0.5, (3, 3000)).astype(np.float32),
     (16, 12, 10), 1.5, 5, 8, 3, False, True),
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_nufftop_matches_oracle(oracle, NUFFTOp, case):
    tf, img, osf, kw, sw, C, use_sens, balance = CASES[case]
    traj = tf()
    M = traj.shape[1]
    print("case", case, img, C, use_sens, balance, flush=True)
    rng = np.random.default_rng(10 + case)
    dens = W.ramp_density(traj)
    sens = W.gaussian_sens(img, C, seed=4) if use_sens else None
    op = NUFFTOp(traj, np.array(img, np.int32), C, sens, dens, kw, sw, osf, balance)
    ref = oracle.Operator(traj, img, kw, sw, osf, dens=dens, sens=sens,
                          lookup=oracle.LOOKUP_LINEAR, order="stable")
    ishape = oracle.img_shape(ref.gi)
    x = W.complex_normal(rng, ((1,) if use_sens else (C,)) + ishape)
    k = op.op(x)
    assert k.shape == (C, M)
    assert oracle.rel_l2(k, ref.forward(x)) < TOL
    kd = W.complex_normal(rng, (C, M))
    a = op.adj_op(kd)
    depth = img[2] if len(img) == 3 else 1
    assert a.shape == ((depth, img[1], img[0]) if use_sens else (C, depth, img[1], img[0]))
    r = ref.adjoint(kd)
    assert oracle.rel_l2(a.reshape(r.shape), r) < TOL
    g = op.adj_op(kd, True)
    rg = ref.adjoint(kd, stop=oracle.DENSITY_ESTIMATION)
    assert g.shape[0] == C and g.size == rg.size
    assert oracle.rel_l2(g.reshape(rg.shape), rg) < TOL


@pytest.mark.parametrize("balance", [True, False])
def test_nufftop_tables(oracle, NUFFTOp, balance):
    traj = W.radial_2d(40, 80, golden=True)
    img = (40, 36)
    dens = W.ramp_density(traj)
    op = NUFFTOp(traj, np.array(img, np.int32), 1, None, dens, 3, 8, 2, balance)
    t = op.get_tables()
    # the reference's std::sort permutation, bit-exact
    pre = oracle.precompute(traj, img, 2.0, 3, 8, dens, lookup=oracle.LOOKUP_LINEAR, order="std")
    assert t["operator_type"] == (3 if balance else 1)  # BALANCED_TEXTURE / TEXTURE
    assert np.array_equal(t["data_indices"], pre["data_indices"])
    assert np.array_equal(t["sector_data_count"], pre["sector_data_count"])
    assert np.array_equal(t["sector_centers"], pre["sector_centers"])
    assert np.array_equal(t["kspace_traj"], pre["traj_sorted"].reshape(-1))
    lut = oracle.load_lut(2.0, 3, oracle.LOOKUP_LINEAR)
    assert np.array_equal(t["kernel"], np.outer(lut, lut).astype(np.float32).reshape(-1))
    ref_deapo = oracle.compute_deapo(img, 3, 2.0, oracle.LOOKUP_LINEAR)
    assert np.max(np.abs(t["deapo"] - ref_deapo.reshape(-1)) / np.abs(ref_deapo.reshape(-1))) < 2e-5
    if balance:
        order = oracle.processing_order(pre["sector_data_count"])
        assert np.array_equal(t["sector_processing_order"], np.asarray(order, np.uint32).reshape(-1, 2))


def test_nufftop_interpolate_data_osf1(oracle, NUFFTOp):
    traj = W.radial_2d(16, 32)
    img = (32, 32)
    op = NUFFTOp(traj, np.array(img, np.int32), 2, None, np.ones(traj.shape[1], np.float32), 3, 8, 1)
    ref = oracle.Operator(traj, img, 3, 8, 1.0, dens=np.ones(traj.shape[1], np.float32),
                          lookup=oracle.LOOKUP_LINEAR, order="stable")
    x = W.complex_normal(np.random.default_rng(0), (2, 32, 32))
    assert oracle.rel_l2(op.op(x, True), ref.forward(x, stop=oracle.DENSITY_ESTIMATION)) < TOL


def test_nufftop_argument_errors(NUFFTOp):
    traj = W.radial_2d(8, 16)
    with pytest.raises(ValueError):
        NUFFTOp(traj[:1], np.array([16, 16], np.int32), 1)
    op = NUFFTOp(traj, np.array([16, 16], np.int32), 2)
    with pytest.raises(ValueError):
        op.adj_op(np.zeros((1, traj.shape[1]), np.complex64))
