"""gnufft_adjoint_noalloc (the adjoint whose work memory is the library's
fixed per-device area) against gnufft_adjoint, through the C ABI.

The two entry points share their kernels; they differ in where the work grid
and the spread sub-grids live and in how the coils are batched (the fixed
area holds fewer coils at a time, which changes only the float order of the
coil-combine sum). Tolerance: rel-L2 <= 1e-6.

Two calls in a row on one device: a 3-D plan with sensitivities (8 coils,
several coil batches in the area), then a smaller 2-D plan (4 coils) whose
work memory is a prefix of the first call's; plus the grid stop (caller's
buffer is the grid) and the size limit (a plan whose one coil needs more
than the area is refused with GNUFFT_ERR_UNSUPPORTED before any work)."""

import numpy as np
import pytest

from gpunufft_amd import workloads as W

pytestmark = pytest.mark.gpu
TOL = 1e-6


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a ROCm device"
    return torch


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def _plan(torch, capi, traj, img, osf, sens_coils, seed):
    p = capi.Plan(img, 3, 8, osf)
    p.set_trajectory(torch.from_numpy(traj).cuda(),
                     torch.from_numpy(W.ramp_density(traj)).cuda())
    if sens_coils:
        p.set_sens(torch.from_numpy(W.gaussian_sens(img, sens_coils, seed=seed)).cuda(), sens_coils)
    return p


def _compare(torch, capi, p, img, n_coils, M, seed, stop):
    rng = np.random.default_rng(seed)
    kd = torch.from_numpy(W.complex_normal(rng, (n_coils, M))).cuda()
    info = p.info()
    if stop == capi.CONVOLUTION:
        shape = (n_coils, int(info.grid_count))
    else:
        combined = info.n_sens_coils > 0
        shape = ((1 if combined else n_coils),) + tuple(img[::-1])
    ref = torch.zeros(shape, dtype=torch.complex64, device="cuda")
    out = torch.full(shape, 7.0 + 7.0j, dtype=torch.complex64, device="cuda")  # stale contents
    p.adjoint(kd, ref, n_coils, stop=stop)
    p.adjoint_noalloc(kd, out, n_coils, stop=stop)
    torch.cuda.synchronize()
    e = _rel(out.cpu().numpy(), ref.cpu().numpy())
    print("noalloc vs adjoint: stop %d, %d coils, rel-L2 %.2e" % (stop, n_coils, e), flush=True)
    return e


def test_noalloc_matches_adjoint_two_calls(torch_dev):
    from gpunufft_amd import capi
    capi.load()
    cap = int(capi.load().gnufft_noalloc_capacity())
    assert cap == 64 << 20
    # call 1: 3-D koosh 96^3, osf 1.5 (144^3 grid), 8 coils with
    # sensitivities; one coil's work memory (24 MB grid + sub-grids) leaves
    # room for one coil at a time, so the area takes the coils in 8 batches
    traj3 = W.koosh_3d(1500, 96)
    p3 = _plan(torch_dev, capi, traj3, (96, 96, 96), 1.5, 8, 3)
    assert _compare(torch_dev, capi, p3, (96, 96, 96), 8, traj3.shape[1], 11,
                    capi.DEAPODIZATION) <= TOL
    # call 2, smaller input: 2-D radial 96^2, 4 coils, no sensitivities
    traj2 = W.radial_2d(48, 128)
    p2 = _plan(torch_dev, capi, traj2, (96, 96), 2.0, 0, 4)
    assert _compare(torch_dev, capi, p2, (96, 96), 4, traj2.shape[1], 12,
                    capi.DEAPODIZATION) <= TOL
    # grid stop: the caller's buffer is the grid, only the sub-grids use the area
    assert _compare(torch_dev, capi, p2, (96, 96), 4, traj2.shape[1], 13,
                    capi.CONVOLUTION) <= TOL


def test_noalloc_refuses_plans_beyond_the_area(torch_dev):
    from gpunufft_amd import capi
    L = capi.load()
    # 256^3 image at osf 2: one coil's 512^3 work grid alone is 1 GiB
    traj = W.koosh_3d(64, 32)
    p = capi.Plan((256, 256, 256), 3, 8, 2.0)
    p.set_trajectory(torch_dev.from_numpy(traj).cuda())
    kd = torch_dev.zeros((1, traj.shape[1]), dtype=torch_dev.complex64, device="cuda")
    out = torch_dev.zeros((1, 256, 256, 256), dtype=torch_dev.complex64, device="cuda")
    rc = L.gnufft_adjoint_noalloc(p._h, kd.data_ptr(), out.data_ptr(), 1, capi.DEAPODIZATION, None)
    assert rc == capi.ERR_UNSUPPORTED
    assert b"no-allocation area" in L.gnufft_last_error()
