"""Sector-width autotuner.

The sector width trades LDS footprint and per-sector fixed costs (staging,
sub-grid stores, fold) against sector occupancy. The best width depends on
the image size, trajectory density and coil count. For one operator
shape, best_sector_width builds a plan for every candidate width, times the
spread (with its reduction and fold) and the interpolation on device-resident
data with HIP events, and returns the fastest width. Results are cached per
(image size, kernel width, osf, sample count, coil count) so operators of one
shape are tuned once.

A batch runs every operator inside launches that fill the GPU, so what an
operator costs there is its share of a full GPU, not the latency of a lone
launch (which favours small sectors: more workgroups for a small operator).
With fill_blocks set (the default), the timed launches repeat the coil set
until they hold about that many workgroups, and the time is divided by the
repeat count.

Workgroup size and coil blocking follow from the width: the kernels size
their LDS rounds and coil blocks to the padded sector of the chosen width.
For one plan, best_workgroups_per_cu then times the kernel-width-3 gridding
kernels at each workgroups-per-CU target (the coil block they pick: 4, 2
or 1 coils of the padded sector in LDS) and keeps the fastest on the plan.
"""

from __future__ import annotations

from . import capi

_CACHE: dict = {}


def best_sector_width(img_dims, kernel_width, osf, traj, dens, n_coils, widths=(4, 8, 16),
                      lookup=capi.LOOKUP_NEAREST, reps=3, fill_blocks=4096, max_repeat=16):
    """traj (d, M) / dens (M,) float32 device tensors; returns the width of
    `widths` with the shortest spread + interp time per coil set, timed with
    the coil set repeated (at most max_repeat times) until a launch holds
    about fill_blocks workgroups (fill_blocks=0: the operator alone)."""
    import torch

    key = (tuple(img_dims), kernel_width, float(osf), int(traj.shape[1]), int(n_coils),
           tuple(widths), lookup, int(fill_blocks), int(max_repeat))
    if key in _CACHE:
        return _CACHE[key]
    M = traj.shape[1]
    best, best_t = None, None
    for sw in widths:
        try:
            p = capi.Plan(img_dims, kernel_width, sw, osf, lookup)
        except capi.GnufftError:
            continue  # padded sector too large for this width
        p.set_trajectory(traj, dens)
        info = p.info()
        rep = 1
        if fill_blocks:
            per_set = max(1, info.n_chunks) * -(-n_coils // 8)  # chunks x 8-coil blocks
            rep = max(1, min(max_repeat, -(-fill_blocks // per_set)))
        C = n_coils * rep
        data = torch.zeros((C, M), dtype=torch.complex64, device="cuda")
        gshape = (C,) + tuple(int(x) for x in info.grid[:len(img_dims)][::-1])
        grid = torch.empty(gshape, dtype=torch.complex64, device="cuda")
        out = torch.empty((C, M), dtype=torch.complex64, device="cuda")
        p.spread_sorted(data, grid, C)
        p.interp_sorted(grid, out, C)
        ts = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            p.spread_sorted(data, grid, C)
            p.interp_sorted(grid, out, C)
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b))
        t = min(ts) / rep
        if best_t is None or t < best_t:
            best, best_t = sw, t
        p.close()
    if best is None:
        raise capi.GnufftError(capi.OK + 5, "no sector width of %s fits this operator" % (widths,))
    _CACHE[key] = best
    return best


def best_workgroups_per_cu(plan, n_coils, targets=(2, 3, 4), reps=3):
    """plan: a capi.Plan with its trajectory set. Times spread + interp of
    n_coils coils (sorted-order entry points, device-resident zeros) at each
    workgroups-per-CU target, sets the fastest on the plan and returns it.
    The gridded values do not depend on the target."""
    import torch

    info = plan.info()
    M = int(info.n_samples)
    dims = int(info.dim)
    gshape = (n_coils,) + tuple(int(x) for x in info.grid[:dims][::-1])
    data = torch.zeros((n_coils, M), dtype=torch.complex64, device="cuda")
    grid = torch.empty(gshape, dtype=torch.complex64, device="cuda")
    out = torch.empty((n_coils, M), dtype=torch.complex64, device="cuda")
    best, best_t = None, None
    for n in targets:
        plan.set_workgroups_per_cu(n)
        plan.spread_sorted(data, grid, n_coils)
        plan.interp_sorted(grid, out, n_coils)
        ts = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            plan.spread_sorted(data, grid, n_coils)
            plan.interp_sorted(grid, out, n_coils)
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b))
        t = min(ts)
        if best_t is None or t < best_t:
            best, best_t = n, t
    plan.set_workgroups_per_cu(best)
    return best
