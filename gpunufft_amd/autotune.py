"""Sector-width autotuner.

The sector width trades LDS footprint and per-sector fixed costs (staging,
sub-grid stores, fold) against sector occupancy. The best width depends on
the image size, trajectory density and coil count. For one operator
shape, best_sector_width builds a plan for every candidate width, times the
spread (with its reduction and fold) and the interpolation on device-resident
data with HIP events, and returns the fastest width. Results are cached per
(image size, kernel width, osf, sample count, coil count) so operators of one
shape are tuned once.

Workgroup size and coil blocking follow from the width: the kernels size
their LDS rounds and coil blocks to the padded sector of the chosen width.
"""

from __future__ import annotations

from . import capi

_CACHE: dict = {}


def best_sector_width(img_dims, kernel_width, osf, traj, dens, n_coils, widths=(4, 8, 16),
                      lookup=capi.LOOKUP_NEAREST, reps=3):
    """traj (d, M) / dens (M,) float32 device tensors; returns the width of
    `widths` with the shortest spread + interp time."""
    import torch

    key = (tuple(img_dims), kernel_width, float(osf), int(traj.shape[1]), int(n_coils),
           tuple(widths), lookup)
    if key in _CACHE:
        return _CACHE[key]
    M = traj.shape[1]
    data = torch.zeros((n_coils, M), dtype=torch.complex64, device="cuda")
    best, best_t = None, None
    for sw in widths:
        try:
            p = capi.Plan(img_dims, kernel_width, sw, osf, lookup)
        except capi.GnufftError:
            continue  # padded sector too large for this width
        p.set_trajectory(traj, dens)
        g = p.info().grid
        gshape = (n_coils,) + tuple(int(x) for x in g[:len(img_dims)][::-1])
        grid = torch.empty(gshape, dtype=torch.complex64, device="cuda")
        out = torch.empty((n_coils, M), dtype=torch.complex64, device="cuda")
        p.spread_sorted(data, grid, n_coils)
        p.interp_sorted(grid, out, n_coils)
        ts = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            p.spread_sorted(data, grid, n_coils)
            p.interp_sorted(grid, out, n_coils)
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b))
        t = min(ts)
        if best_t is None or t < best_t:
            best, best_t = sw, t
        p.close()
    if best is None:
        raise capi.GnufftError(capi.OK + 5, "no sector width of %s fits this operator" % (widths,))
    _CACHE[key] = best
    return best
