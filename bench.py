"""Benchmark: gridded samples/s of the sector NUFFT, adjoint + forward, on one GPU.

Workload (3-D koosh-ball): 128^3 complex64 image, osf 1.5, kernel width 3,
sector width 8, 16384 spokes x 256 samples (Fibonacci-sphere directions,
radii linear in [-0.5, 0.5)), 8 coils with Gaussian-blob sensitivities,
|k|^2 density compensation, complex-normal data (seeded, synthetic).

One step = one full adjoint (k-space [coil][sample] in the caller's order in
HBM -> coil-combined image) plus one full forward (image -> k-space) through
the C ABI, everything resident in HBM. value = 2*M*C / step time.

roofline: the adjoint gridding stage (spread_lane_kernel, reduce_chunks_kernel,
fold_kernel: samples in HBM -> grid in HBM), algorithmic bytes per launch
8*M*C + 4*d*M + 4*M (density) + 8*Ng*C, over the stage's average time
measured with HIP events on the launch stream (library stage timer).
cpu_baseline: the oracle restatement running the same adjoint + forward
(all 8 coils, sensitivities, density) on the host, OpenMP over all cores
this process may use (and, beside it, on one thread); its outputs double as
the full-size parity check of the GPU outputs.

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
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the batch and serving figures of the default run")
    ap.add_argument("--workload", default=WORKLOAD,
                    choices=[WORKLOAD, "uneven_batch_2d", "serving"],
                    help="headline 3-D workload, or 64 uneven 2-D operators batched vs "
                         "sequential plain launches")
    ap.add_argument("--requests", type=int, default=100, help="serving: number of requests")
    ap.add_argument("--rate", type=float, default=30.0, help="serving: Poisson arrival rate (1/s)")
    ap.add_argument("--slots", type=int, default=4, help="serving: concurrent requests")
    ap.add_argument("--sector-widths", default="4,8,16",
                    help="uneven_batch_2d: comma-separated sector widths the autotuner "
                         "chooses from per operator (one width: no tuning)")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="run only this many steps (for rocprofv3), no baseline")
    ap.add_argument("--profile-copy-mb", type=int, default=0,
                    help="in profile mode, also copy this many MiB device-to-device "
                         "(known-size reference for HBM traffic counters)")
    ap.add_argument("--summarize-counters", nargs="+", metavar="CSV",
                    help="write the committed counter summaries of profiles/ from rocprofv3 "
                         "--pmc counter_collection.csv files of --profile-steps 1 runs "
                         "(no GPU needed)")
    ap.add_argument("--summarize-trace", metavar="CSV",
                    help="write profiles/koosh3d_128_mc8_kernel_trace.md from the "
                         "kernel_stats.csv of a rocprofv3 --kernel-trace --stats run of "
                         "--profile-steps 3 (no GPU needed)")
    ap.add_argument("--trace-workload", default="koosh3d_128_mc8",
                    choices=["koosh3d_128_mc8", "uneven_batch_2d"],
                    help="with --summarize-trace: the workload the trace was taken of "
                         "(uneven_batch_2d: a run of --workload uneven_batch_2d, written to "
                         "profiles/uneven_batch_2d_kernel_trace.md)")
    ap.add_argument("--sources-sha256", default=None,
                    help="with --summarize-counters: the kernel-source hash the profiled run "
                         "printed (default: the hash of the sources here)")
    return ap.parse_args()


# sources whose kernels the traffic counters measured
TRAFFIC_SOURCES = ["gpunufft_amd/csrc/hip/gridding.hip", "gpunufft_amd/csrc/hip/gnufft_internal.hpp"]


def sources_sha256():
    """SHA-256 over the kernel sources a counter summary was taken from."""
    import hashlib
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def traffic_summary():
    """HBM bytes per adjoint gridding launch from the committed rocprofv3
    counter runs, or None when the summary was taken from other kernel
    sources than the ones built here (its sources_sha256 stamp differs)."""
    p = os.path.join(ROOT, "profiles", "spread_koosh3d_128_mc8_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        t = json.load(f)
    if t.get("sources_sha256") != sources_sha256():
        return None
    return t


GRIDDING_KERNELS = ("spread", "reduce", "fold", "interp")


def summarize_counters(paths, sha=None):
    """Per-dispatch counter values of the gridding kernels from rocprofv3
    --pmc runs of `bench.py --profile-steps 1 --profile-copy-mb 1024`.

    HBM bytes: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what
    the same counters read for the 1 GiB device-to-device copy of the same
    run (the largest copy dispatch), so that the copy counts 1 GiB each way.
    Writes profiles/spread_koosh3d_128_mc8_traffic.json (stamped with the
    SHA-256 of the kernel sources) and profiles/gridding_koosh3d_128_mc8_counters.md."""
    import csv
    import collections

    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in paths:
        with open(path) as f:
            for r in csv.DictReader(f):
                name = r["Kernel_Name"]
                base = name[5:] if name.startswith("void ") else name
                key = None
                for k in GRIDDING_KERNELS:
                    if base.startswith("gnufft::%s" % k) and "batch" not in base:
                        key = base.split("(")[0].replace("gnufft::", "")
                if "copyBuffer" in name:
                    key = "copy"
                if key is None:
                    continue
                vals[key][r["Counter_Name"]].append(float(r["Counter_Value"]))
    copy = vals.get("copy", {})
    cal = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        if copy.get(c):
            cal[c] = (1 << 30) / (max(copy[c]) * 1024.0)
    kernels = {}
    for k, cs in sorted(vals.items()):
        if k == "copy":
            continue
        e = {c: sum(v) / len(v) for c, v in cs.items()}
        if "FETCH_SIZE" in e and "FETCH_SIZE" in cal:
            e["fetch_bytes"] = e["FETCH_SIZE"] * 1024.0 * cal["FETCH_SIZE"]
        if "WRITE_SIZE" in e and "WRITE_SIZE" in cal:
            e["write_bytes"] = e["WRITE_SIZE"] * 1024.0 * cal["WRITE_SIZE"]
        kernels[k] = e
    adj = [k for k in kernels if k.split("<")[0] in ("spread_lane_kernel", "spread_kernel",
                                                      "reduce_chunks_kernel", "fold_kernel")]
    hbm = sum(kernels[k].get("fetch_bytes", 0) + kernels[k].get("write_bytes", 0) for k in adj)
    from gpunufft_amd import workloads as W
    w = W.config(WORKLOAD)
    d, M = w["traj"].shape
    C = w["coils"]
    Ng = int(round(w["osf"] * w["img_dims"][0])) ** 3
    alg = 8 * M * C + 4 * d * M + 4 * M + 8 * Ng * C
    out = {
        "workload": "koosh3d 128^3 osf1.5 kw3 sw8, 16384x256 samples, 8 coils",
        "sources_sha256": sha or sources_sha256(),
        "sources": TRAFFIC_SOURCES,
        "commands": [
            "cd /tmp && export TMPDIR=/tmp && timeout -s KILL 90 rocprofv3 --pmc <COUNTERS> "
            "--output-format csv -d $R/gpurun_out/<dir> -o run -- python3 $R/bench.py "
            "--profile-steps 1 --profile-copy-mb 1024   (one run per counter set)",
            "python3 bench.py --summarize-counters $R/gpurun_out/<dir>/run_counter_collection.csv ...",
        ],
        "calibration": {"copy_bytes": 1 << 30, "scale": cal,
                        "note": "FETCH_SIZE / WRITE_SIZE of a 1 GiB device copy mapped to 1 GiB"},
        "kernels": kernels,
        "adjoint_gridding_kernels": adj,
        "hbm_bytes_per_launch": round(hbm) if hbm else None,
        "algorithmic_bytes_per_launch": alg,
        "traffic_over_algorithmic": round(hbm / alg, 3) if hbm else None,
    }
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    if hbm:
        with open(os.path.join(ROOT, "profiles", "spread_koosh3d_128_mc8_traffic.json"), "w") as f:
            json.dump(out, f, indent=1)
    lines = ["# Gridding kernel counters: koosh3d 128^3 osf 1.5 kw 3 sw 8, 16384x256 samples, "
             "8 coils (one adjoint + forward step)", "",
             "Per dispatch, averaged over dispatches; HBM bytes scaled by the 1 GiB copy "
             "(calibration %s). Sources SHA-256 %s." % (json.dumps(cal), out["sources_sha256"]),
             "", "Commands: " + " ; ".join(out["commands"]), ""]
    cols = ["fetch_bytes", "write_bytes", "SQ_INSTS_VALU", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT",
            "SQ_WAIT_INST_LDS", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES", "SQ_ACTIVE_INST_VALU"]
    cols = [c for c in cols if any(c in e for e in kernels.values())]
    lines.append("| kernel | " + " | ".join(cols) + " |")
    lines.append("|---|" + "---|" * len(cols))
    for k, e in kernels.items():
        lines.append("| `%s` | " % k + " | ".join("%.4g" % e[c] if c in e else "-" for c in cols) + " |")
    lines.append("")
    for k, e in kernels.items():
        if "SQ_WAIT_ANY" in e and e.get("SQ_WAVE_CYCLES"):
            lines.append("- `%s`: SQ_WAIT_ANY / SQ_WAVE_CYCLES = %.2f" % (k, e["SQ_WAIT_ANY"] / e["SQ_WAVE_CYCLES"]))
        if "SQ_LDS_BANK_CONFLICT" in e and e.get("SQ_INSTS_LDS"):
            lines.append("- `%s`: SQ_LDS_BANK_CONFLICT / SQ_INSTS_LDS = %.2f cycles per LDS instruction"
                         % (k, e["SQ_LDS_BANK_CONFLICT"] / e["SQ_INSTS_LDS"]))
    if hbm:
        lines.append("- adjoint gridding stage (%s): %.4g HBM bytes per launch = %.2fx the %d "
                     "algorithmic bytes" % (", ".join(adj), hbm, hbm / alg, alg))
    with open(os.path.join(ROOT, "profiles", "gridding_koosh3d_128_mc8_counters.md"), "w") as f:
        f.write("\n".join(lines) + "\n")
    print(json.dumps({"traffic": out["hbm_bytes_per_launch"], "kernels": list(kernels)}))


def summarize_trace(path, sha=None, workload="koosh3d_128_mc8"):
    """Kernel time table of a --profile-steps 3 kernel trace (three steps of
    the headline workload after one plan set-up), or of a run of
    --workload uneven_batch_2d (workload="uneven_batch_2d")."""
    import csv
    rows = list(csv.DictReader(open(path)))
    if workload == "uneven_batch_2d":
        return _summarize_batch_trace(rows, path, sha)
    lines = ["# Kernel trace: koosh3d 128^3 osf 1.5 kw 3 sw 8, 16384x256 samples, 8 coils, "
             "adjoint + forward x3", "",
             "Command (one MI355X, from the repository root R):", "", "```",
             "cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv "
             "-d $R/gpurun_out/kt -o run -- python3 $R/bench.py --profile-steps 3",
             "python3 bench.py --summarize-trace $R/gpurun_out/kt/run_kernel_stats.csv", "```", "",
             "Kernel sources SHA-256 %s (bench.py sources_sha256)." % (sha or sources_sha256()), "",
             "`--profile-steps 3` builds the plan (sector precompute, once) and runs three full "
             "steps (adjoint: spread, chunk reduction, fold, fused FFT passes z, y, x; forward: "
             "fused FFT passes x, y, z, interp). `__amd_rocclr_copyBuffer` is the host-to-device "
             "upload of the synthetic workload.", "",
             "| kernel | calls | total us | avg us | % |", "|---|---|---|---|---|"]
    for r in rows:
        name = r["Name"]
        short = name.replace("(anonymous namespace)::", "").split("(")[0]
        short = short.replace("void ", "").replace("gnufft::", "")
        if "rocprim" in name:
            short = "rocprim (sort/scan of the sector precompute)"
        lines.append("| `%s` | %s | %.1f | %.1f | %.1f |" % (
            short, r["Calls"], float(r["TotalDurationNs"]) / 1e3, float(r["AverageNs"]) / 1e3,
            float(r["Percentage"])))
    with open(os.path.join(ROOT, "profiles", "koosh3d_128_mc8_kernel_trace.md"), "w") as f:
        f.write("\n".join(lines) + "\n")


def _summarize_batch_trace(rows, path, sha):
    """Kernel table of a `--workload uneven_batch_2d` trace: the batched
    launches (kernel names with "batch") and the plans' own launches of the
    plain path, with the batched time per step."""
    def short(name):
        if "rocprim" in name:
            return "rocprim (sort/scan of the sector precompute)"
        s = name.replace("(anonymous namespace)::", "").split("(")[0]
        return s.replace("void ", "").replace("gnufft::", "")
    batch = [r for r in rows if "batch" in r["Name"] and "gnufft" in r["Name"]]
    steps = max((int(r["Calls"]) for r in batch if "spread_batch" in r["Name"]), default=0)
    per_step = sum(float(r["TotalDurationNs"]) for r in batch) / 1e3 / steps if steps else 0.0
    lines = ["# Kernel trace: uneven 2-D batch, 64 operators (N 96-384, 4-16 coils, spirals of "
             "8-48 x 1024-4096 samples, sens + density, osf 2, kw 3, autotuned sector widths)", "",
             "Command (one MI355X, from the repository root R):", "", "```",
             "cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv "
             "-d $R/gpurun_out/ktb -o run -- python3 $R/bench.py --workload uneven_batch_2d",
             "python3 bench.py --summarize-trace $R/gpurun_out/ktb/run_kernel_stats.csv "
             "--trace-workload uneven_batch_2d", "```", "",
             "Kernel sources SHA-256 %s (bench.py sources_sha256)." % (sha or sources_sha256()), "",
             "The run tunes the sector widths, then times the 64 operators one after another "
             "(each plan's own launches) and as one batch (one launch per stage for all plans), "
             "warm-up and timed steps alike. Kernels with `batch` in their name are the batched "
             "launches: %d steps, **%.1f us of batched kernel time per step** (adjoint + forward "
             "of all 64 operators)." % (steps, per_step), "",
             "| kernel | calls | total us | avg us | per batched step us | % |",
             "|---|---|---|---|---|---|"]
    for r in rows:
        b = r in batch
        lines.append("| `%s` | %s | %.1f | %.1f | %s | %.1f |" % (
            short(r["Name"]), r["Calls"], float(r["TotalDurationNs"]) / 1e3,
            float(r["AverageNs"]) / 1e3,
            "%.1f" % (float(r["TotalDurationNs"]) / 1e3 / steps) if b and steps else "-",
            float(r["Percentage"])))
    with open(os.path.join(ROOT, "profiles", "uneven_batch_2d_kernel_trace.md"), "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    args = parse()
    if args.summarize_trace:
        summarize_trace(args.summarize_trace, args.sources_sha256, args.trace_workload)
        return
    if args.summarize_counters:
        summarize_counters(args.summarize_counters, args.sources_sha256)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only (this build has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gpunufft_amd import capi
    from gpunufft_amd import workloads as W

    torch.cuda.set_device(0)
    if args.workload == "uneven_batch_2d":
        print(json.dumps(bench_batch(args)))
        return
    if args.workload == "serving":
        print(json.dumps(bench_serving(args)))
        return
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
        if args.profile_copy_mb:
            a = torch.empty(args.profile_copy_mb << 18, dtype=torch.float32, device="cuda")
            b = torch.empty_like(a)
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"profile_steps": args.profile_steps,
                          "copy_bytes": args.profile_copy_mb << 20,
                          "sources_sha256": sources_sha256()}))
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
        "kernel": "adjoint gridding stage: spread_lane_kernel + reduce_chunks_kernel + "
                  "fold_kernel (3-D, kernel width 3, 4-coil blocks)",
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
        "value_definition": "2*M*C per step / step time (whole adjoint+forward operator incl. "
                            "FFTs); gridding_only.samples_per_s = 2*M*C/(t_spread+t_interp)",
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
        out["cpu_baseline"], out["parity_rel_l2"] = cpu_baseline(w, plan, sens, kd, torch)
    del plan, k_dev, img_dev, k_out
    torch.cuda.empty_cache()
    if not args.headline_only:
        # the metric's other two figures, measured in the same run: the
        # batched-over-plain speed-up of the uneven 2-D batch and the
        # CG-SENSE serving latencies
        b = bench_batch(args)
        out["batch_uneven_2d"] = {k: b[k] for k in (
            "speedup_batched_vs_plain", "ms_per_step", "plain_ms_per_step", "value",
            "batched_vs_plain_rel_l2")}
        out["batch_uneven_2d"]["sector_widths"] = sorted(set(b["config"]["sector_widths"]))
        sv = bench_serving(args)
        out["serving_cg_sense"] = {k: sv[k] for k in (
            "p50_ms", "p99_ms", "mean_ms", "value", "unit", "peak_concurrent", "saturated",
            "one_at_a_time")}
        out["serving_cg_sense"]["config"] = sv["config"]["workload"]
    print(json.dumps(out))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _oracle_flags():
    try:
        with open(os.path.join(ROOT, "oracle", "Makefile")) as f:
            for line in f:
                if line.startswith("CXXFLAGS"):
                    return "g++ " + line.split("=", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, plan, sens, kd, torch):
    """The oracle restatement of the same adjoint + forward (all coils,
    sensitivities, density) on the host cores: OpenMP over every core this
    process may run on (sectors shared out, one private grid per thread,
    summed at the end; threaded FFTs), and one thread as the reference's
    own CPU gridding runs. The multi-core outputs are also the full-size
    parity check of the GPU adjoint and forward on the same inputs."""
    import numpy as np

    from oracle import oracle as O

    traj, img = w["traj"], w["img_dims"]
    M = traj.shape[1]
    C = kd.shape[0]
    cores = len(os.sched_getaffinity(0))
    op_mt = O.Operator(traj, img, w["kw"], w["sw"], w["osf"], dens=w["dens"], sens=sens,
                       order="stable", threads=cores)
    t0 = time.perf_counter()
    x_ref = op_mt.adjoint(kd)
    k_ref = op_mt.forward(x_ref)
    t_mt = time.perf_counter() - t0
    op_1 = O.Operator(traj, img, w["kw"], w["sw"], w["osf"], dens=w["dens"], sens=sens,
                      order="stable")
    t0 = time.perf_counter()
    x_1 = op_1.adjoint(kd)
    k_1 = op_1.forward(x_1)
    t_1 = time.perf_counter() - t0
    x_gpu = torch.empty(x_ref.shape, dtype=torch.complex64, device="cuda")
    plan.adjoint(torch.from_numpy(np.ascontiguousarray(kd)).cuda(), x_gpu, C)
    k_gpu = torch.empty((C, M), dtype=torch.complex64, device="cuda")
    plan.forward(torch.from_numpy(np.ascontiguousarray(x_ref)).cuda(), k_gpu, C)
    torch.cuda.synchronize()
    parity = {"adjoint": O.rel_l2(x_gpu.cpu().numpy(), x_ref),
              "forward": O.rel_l2(k_gpu.cpu().numpy(), k_ref),
              "cpu_1_thread_vs_cpu_%d_threads" % cores: max(O.rel_l2(x_1, x_ref),
                                                              O.rel_l2(k_1, k_ref))}
    return ({
        "value": round(2.0 * M * C / t_mt, 1),
        "unit": "gridded samples/s",
        "cores": cores,
        "kind": "port",
        "sample": "full workload: adjoint+forward of all %d coils (M=%d, sens, density), oracle "
                  "restatement, OpenMP over %d threads (private grids) + scipy.fft workers=%d, "
                  "%.2f s" % (C, M, cores, cores, t_mt),
        "cpu_model": _cpu_model(),
        "compiler_flags": _oracle_flags(),
        "single_thread": {"value": round(2.0 * M * C / t_1, 1), "cores": 1,
                          "seconds": round(t_1, 2)},
    }, parity)


def bench_batch(args):
    """64 independent 2-D operators (uneven sizes, spiral trajectories, 4-16
    coils, sens + density): one step = adjoint + forward of every operator,
    run (a) as 64 sequential plain per-operator calls and (b) as one batch
    (one launch per stage). value = batched gridded samples/s."""
    import numpy as np
    import torch

    from gpunufft_amd import autotune, capi
    from gpunufft_amd import workloads as W

    ops = W.uneven_batch_2d()
    widths = [int(x) for x in args.sector_widths.split(",")]
    t0 = time.perf_counter()
    plans, kd, img, kout, iout, sws = [], [], [], [], [], []
    for j, o in enumerate(ops):
        C = o["coils"]
        N = o["img_dims"][0]
        traj_d = torch.from_numpy(o["traj"]).cuda()
        dens_d = torch.from_numpy(o["dens"]).cuda()
        sw = widths[0] if len(widths) == 1 else autotune.best_sector_width(
            o["img_dims"], o["kw"], o["osf"], traj_d, dens_d, C, widths)
        sws.append(sw)
        p = capi.Plan(o["img_dims"], o["kw"], sw, o["osf"])
        p.set_trajectory(traj_d, dens_d)
        p.set_sens(torch.from_numpy(o["sens"]).cuda(), C)
        plans.append(p)
        rng = np.random.default_rng(o["seed"])
        M = o["traj"].shape[1]
        kd.append(torch.from_numpy(W.complex_normal(rng, (C, M))).cuda())
        img.append(torch.from_numpy(W.complex_normal(rng, (1, N, N))).cuda())
        kout.append(torch.empty((C, M), dtype=torch.complex64, device="cuda"))
        iout.append(torch.empty((1, N, N), dtype=torch.complex64, device="cuda"))
    setup_s = time.perf_counter() - t0
    coils = [o["coils"] for o in ops]
    batch = capi.Batch(plans, coils)
    MC = sum(o["traj"].shape[1] * o["coils"] for o in ops)

    def plain_step():
        for p, k, i, ko, io, C in zip(plans, kd, img, kout, iout, coils):
            p.adjoint(k, io, C)
            p.forward(i, ko, C)

    def batch_step():
        batch.adjoint(kd, iout)
        batch.forward(img, kout)

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / args.steps

    t_plain = timed(plain_step)
    ref_a = [x.clone() for x in iout]
    ref_f = [x.clone() for x in kout]
    t_batch = timed(batch_step)
    err = 0.0
    for a, b_, c, d in zip(iout, ref_a, kout, ref_f):
        for x, y in ((a, b_), (c, d)):
            num = torch.linalg.vector_norm((x - y).flatten()).item()
            den = torch.linalg.vector_norm(y.flatten()).item()
            err = max(err, num / den if den > 0 else num)
    out = {
        "metric": "gridded samples/s adj+fwd, 64 uneven 2-D operators: batched vs plain launches",
        "value": round(2.0 * MC / t_batch, 1),
        "unit": "gridded samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_batch * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded variable-density spirals, complex-normal data, Gaussian coil maps)",
        "config": {"workload": "uneven_batch_2d: 64 operators, N in {96..384}, spirals 8-48 x "
                               "1024-4096, 4-16 coils, osf 2, kw 3, sens+dens",
                   "sample_coil_pairs": MC, "sector_widths": sws},
        "plain_ms_per_step": round(t_plain * 1e3, 4),
        "plain_value": round(2.0 * MC / t_plain, 1),
        "speedup_batched_vs_plain": round(t_plain / t_batch, 3),
        "batched_vs_plain_rel_l2": err,
        "setup_s": round(setup_s, 2),
    }
    batch.close()
    for p in plans:
        p.close()
    return out


def bench_serving(args):
    """CG-SENSE serving mix (70 % 2-D, 30 % 3-D, 10 iterations each, Poisson
    arrivals): continuous batching over `slots` streams with HBM admission
    and graph-replayed iterations, against the same stream served one request
    at a time. value = requests/s of the continuous-batching server."""
    from gpunufft_amd import serve

    reqs = serve.serving_mix(n_requests=args.requests, rate_hz=args.rate, seed=7)
    # first-use costs outside the measurement: every geometry of the mix once
    # (rocFFT compiles kernels for lengths it first sees, code objects load,
    # and the rocFFT plan pairs of destroyed plans stay pooled for later
    # plans of the same geometry, as in a server that has seen them before)
    warm = serve.serving_mix(n_requests=args.requests, rate_hz=args.rate, seed=7, iterations=1)
    for r in warm:
        r.arrival = 0.0
    serve.Server(slots=1).run(warm)
    cont = serve.Server(slots=args.slots).run(reqs)
    seq = serve.Server(slots=1).run(serve.serving_mix(n_requests=args.requests,
                                                      rate_hz=args.rate, seed=7))
    # saturated: every request waiting at t = 0, so the server, not the
    # arrival process, sets the rate (the mix's capacity in requests/s)
    sat_reqs = serve.serving_mix(n_requests=args.requests, rate_hz=args.rate, seed=7)
    for r in sat_reqs:
        r.arrival = 0.0
    sat = serve.Server(slots=args.slots).run(sat_reqs)
    n2 = sum(1 for r in reqs if len(r.img_dims) == 2)
    out = {
        "metric": "CG-SENSE serving: p50/p99 latency (ms) and requests/s, continuous batching",
        "value": round(cont["requests_per_s"], 3),
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": args.requests,
        "warmup": args.requests,
        "ms_per_step": round(cont["makespan_s"] * 1e3 / args.requests, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded radial/spiral/koosh trajectories, Gaussian coil maps, "
                "complex-normal data)",
        "config": {"workload": "serving mix: %d requests (%d 2-D N 128-384 8-32 coils, %d 3-D "
                               "N 96-160 8 coils), osf 1.5/2, kw 3, 10 CG iterations, Poisson "
                               "%.1f/s, %d slots" % (len(reqs), n2, len(reqs) - n2, args.rate,
                                                      args.slots)},
        "p50_ms": round(cont["p50_ms"], 3),
        "p99_ms": round(cont["p99_ms"], 3),
        "mean_ms": round(cont["mean_ms"], 3),
        "peak_concurrent": cont["peak_concurrent"],
        "one_at_a_time": {k: round(v, 3) if isinstance(v, float) else v for k, v in seq.items()},
        "saturated": {"requests_per_s": round(sat["requests_per_s"], 3),
                      "p50_ms": round(sat["p50_ms"], 3), "p99_ms": round(sat["p99_ms"], 3)},
    }
    return out


if __name__ == "__main__":
    main()
