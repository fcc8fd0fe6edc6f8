"""Benchmark: gridded samples/s of the sector NUFFT, adjoint + forward, on one GPU.

Workload (3-D koosh-ball): 128^3 complex64 image, osf 1.5, kernel width 3,
sector width 8, 16384 spokes x 256 samples (Fibonacci-sphere directions,
radii linear in [-0.5, 0.5)), 8 coils with Gaussian-blob sensitivities,
|k|^2 density compensation, complex-normal data (seeded, synthetic).

One step = one full adjoint (k-space [coil][sample] in the caller's order in
HBM -> coil-combined image) plus one full forward (image -> k-space) through
the C ABI, everything resident in HBM. value = 2*M*C / step time.

roofline: the adjoint spread kernel, algorithmic bytes per launch
8*M*C + 4*d*M + 4*M (density) + 8*Ng*C, over its average launch time measured
with HIP events on the launch stream (library stage timer).
cpu_baseline: the oracle restatement (single thread) running the same
adjoint + forward on one of the eight coils of the same workload; its output
doubles as the parity check of the GPU result for that coil.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
WORKLOAD = "koosh3d_128_mc8"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run only this many steps (for rocprofv3), no baseline")
    return ap.parse_args()


def traffic_summary():
    """HBM bytes per spread launch from the committed rocprofv3 counter runs."""
    p = os.path.join(ROOT, "profiles", "spread_koosh3d_128_mc8_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only (this build has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gpunufft_amd import capi
    from gpunufft_amd import workloads as W

    torch.cuda.set_device(0)
    w = W.config(WORKLOAD)
    traj, img, C, dens = w["traj"], w["img_dims"], w["coils"], w["dens"]
    d, M = traj.shape
    rng = np.random.default_rng(w["seed"])
    sens = W.gaussian_sens(img, C, seed=w["seed"])
    kd = W.complex_normal(rng, (C, M))
    ishape = tuple(img[::-1])

    plan = capi.Plan(img, w["kw"], w["sw"], w["osf"])
    t0 = time.perf_counter()
    plan.set_trajectory(torch.from_numpy(traj).cuda(), torch.from_numpy(dens).cuda())
    precompute_s = time.perf_counter() - t0
    plan.set_sens(torch.from_numpy(sens).cuda(), C)
    info = plan.info()
    Ng = int(info.grid_count)

    k_dev = torch.from_numpy(kd).cuda()
    img_dev = torch.empty((1,) + ishape, dtype=torch.complex64, device="cuda")
    k_out = torch.empty((C, M), dtype=torch.complex64, device="cuda")
    stream = torch.cuda.current_stream()

    def step():
        plan.adjoint(k_dev, img_dev, C, stream=stream)
        plan.forward(img_dev, k_out, C, stream=stream)

    if args.profile_steps:
        for _ in range(args.profile_steps):
            step()
        torch.cuda.synchronize()
        print(json.dumps({"profile_steps": args.profile_steps}))
        return

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    plan.set_timing(True)
    plan.timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stages = plan.timing(reset=True)
    plan.set_timing(False)

    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = 2.0 * M * C / ((t1 - t0) / args.steps)

    spread_ms = stages["spread"][0] / max(1, stages["spread"][1])
    interp_ms = stages["interp"][0] / max(1, stages["interp"][1])
    spread_bytes = 8 * M * C + 4 * d * M + 4 * M + 8 * Ng * C
    interp_bytes = 8 * Ng * C + 4 * d * M + 4 * M + 8 * M * C
    achieved = spread_bytes / (spread_ms * 1e-3) / 1e9
    traffic = traffic_summary()
    roofline = {
        "kernel": "spread_kernel<3,4,8> (adjoint gridding)",
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": None if traffic is None else traffic.get("hbm_bytes_per_launch"),
        "algorithmic_bytes": spread_bytes,
        "avg_launch_ms": round(spread_ms, 4),
    }

    out = {
        "metric": "gridded samples/s/GPU adj+fwd at rel-L2<=1e-5 (%HBM roof); p50/p99 CG-SENSE ms",
        "value": round(value, 1),
        "unit": "gridded samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded koosh-ball trajectory, complex-normal data, Gaussian coil maps)",
        "config": {
            "workload": "koosh3d 128^3 osf1.5 kw3 sw8, 16384x256 spokes, 8 coils, sens+dens",
            "image": list(img), "osf": w["osf"], "kernel_width": w["kw"],
            "sector_width": w["sw"], "samples": M, "coils": C, "grid": list(info.grid)[:d],
            "chunks": int(info.n_chunks),
        },
        "roofline": roofline,
        "stages_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in stages.items()},
        "gridding_only": {
            "spread_ms": round(spread_ms, 4), "interp_ms": round(interp_ms, 4),
            "samples_per_s": round(2.0 * M * C / ((spread_ms + interp_ms) * 1e-3), 1),
            "interp_hbm_frac": round(interp_bytes / (interp_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        },
        "precompute_s": round(precompute_s, 4),
    }

    if not args.no_cpu_baseline:
        out["cpu_baseline"], out["parity_rel_l2_coil0"] = cpu_baseline(w, plan, kd, torch)
    print(json.dumps(out))


def cpu_baseline(w, plan, kd, torch):
    """Oracle restatement, one thread, adjoint + forward of coil 0 (no
    sensitivities) of the same workload; also the parity check of that coil."""
    import numpy as np

    from oracle import oracle as O

    os.environ.setdefault("OMP_NUM_THREADS", "1")
    traj, img = w["traj"], w["img_dims"]
    M = traj.shape[1]
    op = O.Operator(traj, img, w["kw"], w["sw"], w["osf"], dens=w["dens"], order="stable")
    y0 = kd[:1]
    t0 = time.perf_counter()
    x_ref = op.adjoint(y0)
    k_ref = op.forward(x_ref)
    t = time.perf_counter() - t0
    # GPU, same single-coil problem (no sensitivities) for the parity figure
    plan.set_sens(None, 0)
    xi = torch.empty(x_ref.shape, dtype=torch.complex64, device="cuda")
    ko = torch.empty((1, M), dtype=torch.complex64, device="cuda")
    plan.adjoint(torch.from_numpy(np.ascontiguousarray(y0)).cuda(), xi, 1)
    plan.forward(xi, ko, 1)
    torch.cuda.synchronize()
    err = max(O.rel_l2(xi.cpu().numpy(), x_ref), O.rel_l2(ko.cpu().numpy(), k_ref))
    base = {
        "value": round(2.0 * M / t, 1),
        "unit": "gridded samples/s",
        "cores": 1,
        "kind": "port",
        "sample": "adjoint+forward of coil 0 of 8 (M=%d samples, no sens), %.1f s" % (M, t),
    }
    return base, err


if __name__ == "__main__":
    main()
