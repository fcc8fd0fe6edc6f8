"""CG-SENSE serving: a device-resident request loop with continuous batching.

Requests (one reconstruction each: trajectory, density weights, coil maps,
k-space data, iteration count) arrive on a Poisson schedule. The loop
admits a waiting request when its arrival time has passed, a slot is free
and its estimated device memory fits in the free HBM (less a reserve). An
admitted request gets its own stream, its own plan (sector precompute on
that stream) and a CGSense solver; its iterations replay a captured HIP
graph. Requests on different streams overlap on the device while the loop
keeps admitting. Completion is detected by polling an event, after which the
request's device memory is released. Each result is the same whether a
request runs alone or in the mix (deterministic kernels and reductions).

The CG algorithm and its regularisation follow the reference demo
(matlab/demo/utils/cg_sense_2d.m:46-82, useMulticoil).
"""

from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi
from . import workloads as W


@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the start of the run
    img_dims: tuple
    traj: np.ndarray
    dens: np.ndarray
    sens: np.ndarray
    kdata: np.ndarray
    osf: float
    kernel_width: int = 3
    sector_width: int = 8
    iterations: int = 10
    alpha: float = 1e-6
    # filled by the server
    start: float = math.nan
    finish: float = math.nan
    result: object = None
    extra: dict = field(default_factory=dict)

    @property
    def coils(self):
        return self.sens.shape[0]

    @property
    def latency(self):
        return self.finish - self.arrival

    def device_bytes(self):
        """Estimate of the device memory the request holds while active."""
        d = len(self.img_dims)
        Ng = int(np.prod([int(n * self.osf) for n in self.img_dims]))
        Ni = int(np.prod(self.img_dims))
        M = self.traj.shape[1]
        C = self.coils
        P = self.sector_width + 2 * (self.kernel_width // 2)
        # grid + FFT work, sub-grid scratch (chunks ~ sectors), sens, data,
        # CG vectors, tables
        sectors = int(np.prod([math.ceil(int(n * self.osf) / self.sector_width)
                               for n in self.img_dims]))
        scratch = (sectors + M // 1024 + 1) * C * P ** d * 8
        return (2 * Ng * C * 8 + scratch + C * Ni * 8 + 2 * C * M * 8 + 5 * Ni * 8 +
                M * (4 * d + 16) + Ni * 4)


def serving_mix(n_requests=40, rate_hz=40.0, seed=7, iterations=10, small=False):
    """Seeded request stream: 70 % 2-D (N in [128, 384], radial or spiral,
    8-32 coils), 30 % 3-D (N in [96, 160], koosh-ball, 8 coils), osf 1.5 or
    2, kernel width 3, Poisson arrivals at rate_hz. small=True shrinks every
    request (for tests)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for i in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        osf = float(rng.choice([1.5, 2.0]))
        if rng.random() < 0.7:
            N = int(rng.integers(128, 385)) // 8 * 8
            C = int(rng.integers(8, 33))
            if small:
                N, C = N // 4, min(C, 6)
            if rng.random() < 0.5:
                traj = W.radial_2d(max(8, N // 2), 2 * N, golden=True)
            else:
                L = int(rng.integers(8, 33))
                traj = W.spiral_2d(L, max(256, 8 * N), n_turns=max(1.0, N / (2.0 * L)))
            dims = (N, N)
        else:
            N = int(rng.integers(96, 161)) // 8 * 8
            C = 8
            if small:
                N, C = N // 4, 3
            traj = W.koosh_3d(max(64, N * N // 4), 2 * N)
            dims = (N, N, N)
        dens = W.ramp_density(traj, len(dims) - 1)
        sens = W.gaussian_sens(dims, C, seed=seed * 1000 + i)
        kdata = W.complex_normal(rng, (C, traj.shape[1]))
        reqs.append(Request(rid=i, arrival=t, img_dims=dims, traj=traj, dens=dens, sens=sens,
                            kdata=kdata, osf=osf, iterations=iterations))
    return reqs


class Server:
    """Continuous-batching CG-SENSE server on the current device."""

    def __init__(self, slots=4, hbm_reserve=0.1, use_graph=True, keep_results=False):
        import torch

        self.torch = torch
        self.slots = slots
        self.hbm_reserve = hbm_reserve
        self.use_graph = use_graph
        self.keep_results = keep_results
        self.streams = [torch.cuda.Stream() for _ in range(slots)]

    def _start(self, req, slot):
        torch = self.torch
        s = self.streams[slot]
        with torch.cuda.stream(s):
            dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to("cuda", non_blocking=True)
            traj, dens, sens, kd = dev(req.traj), dev(req.dens), dev(req.sens), dev(req.kdata)
            plan = capi.Plan(req.img_dims, req.kernel_width, req.sector_width, req.osf, stream=s)
            plan.set_trajectory(traj, dens, stream=s)
            plan.set_sens(sens, req.coils, stream=s)
            cg = capi.CGSense(plan, req.coils, req.alpha)
            x = torch.empty((1,) + tuple(req.img_dims[::-1]), dtype=torch.complex64, device="cuda")
            cg.solve(kd, x, req.iterations, self.use_graph, stream=s)
            ev = torch.cuda.Event()
            ev.record(s)
        return dict(req=req, slot=slot, plan=plan, cg=cg, x=x, ev=ev, keep=(traj, dens, sens, kd))

    def run(self, requests):
        """Serve `requests` (arrival times relative to the start); returns a
        summary dict. Request objects get start / finish / result filled."""
        torch = self.torch
        pending = sorted(requests, key=lambda r: r.arrival)
        active = []
        # finished requests' plans and solvers: releasing device memory
        # synchronises the whole device, so it waits until nothing is in
        # flight (or admission needs the memory)
        graveyard = []

        def bury():
            for a in graveyard:
                a["cg"].close()
                a["plan"].close()
            graveyard.clear()

        free_slots = list(range(self.slots))
        t0 = time.perf_counter()
        total = torch.cuda.mem_get_info()[1]
        peak_active = 0
        while pending or active:
            now = time.perf_counter() - t0
            # completions
            for a in list(active):
                if a["ev"].query():
                    r = a["req"]
                    r.finish = time.perf_counter() - t0
                    if self.keep_results:
                        r.result = a["x"].cpu().numpy()
                    graveyard.append(a)
                    free_slots.append(a["slot"])
                    active.remove(a)
            if graveyard and not active:
                bury()
            # admissions
            while pending and pending[0].arrival <= now and free_slots:
                r = pending[0]
                free = torch.cuda.mem_get_info()[0]
                if r.device_bytes() > free - self.hbm_reserve * total and active:
                    if graveyard:
                        bury()
                        continue
                    break  # wait for memory; a request alone is always admitted
                pending.pop(0)
                slot = free_slots.pop(0)
                r.start = time.perf_counter() - t0
                active.append(self._start(r, slot))
                peak_active = max(peak_active, len(active))
            if not active and pending:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.002))
            elif active:
                time.sleep(0.0001)
        bury()
        makespan = max(r.finish for r in requests)
        lat = np.array([r.latency for r in requests]) * 1e3
        return {
            "requests": len(requests),
            "makespan_s": makespan,
            "requests_per_s": len(requests) / makespan,
            "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(lat.mean()),
            "peak_concurrent": peak_active,
        }
