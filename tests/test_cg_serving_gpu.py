"""Device-resident CG-SENSE (gnufft_cg_*) and the serving loop.

- The CG iterate equals a host CG (the reference demo's loop,
  matlab/demo/utils/cg_sense_2d.m:63-82) run on the oracle operator, within
  rel-L2 1e-4 after 5 iterations (operator rounding of ~1e-6 per product
  accumulates over the iterations).
- Graph replay equals eager execution: the captured iteration is replayed
  in several solves, with eager solves and synchronisations in between, and
  every replay is compared with the eager result (rel-L2 <= 1e-6).
- Serving: every request's result in a concurrent mix equals its result
  when served alone (rel-L2 <= 1e-6), on shrunken requests and on
  full-size ones; at full size the solver's objective also falls with the
  iterations (a property CG fixes, no reference needed at that size).
"""

import numpy as np
import pytest

from gpunufft_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a ROCm device"
    from gpunufft_amd import capi
    capi.load()
    return torch, capi


def _problem(dims=(32, 32), C=3, seed=0):
    traj = W.radial_2d(32, 64, golden=True) if len(dims) == 2 else W.koosh_3d(128, 24)
    dens = W.ramp_density(traj)
    sens = W.gaussian_sens(dims, C, seed=seed)
    kd = W.complex_normal(np.random.default_rng(seed), (C, traj.shape[1]))
    return traj, dens, sens, kd


def _gpu_cg(torch, capi, dims, traj, dens, sens, kd, iters, alpha, use_graph, osf=2.0):
    C = sens.shape[0]
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    p = capi.Plan(dims, 3, 8, osf)
    p.set_trajectory(dev(traj), dev(dens))
    p.set_sens(dev(sens), C)
    cg = capi.CGSense(p, C, alpha)
    x = torch.empty((1,) + tuple(dims[::-1]), dtype=torch.complex64, device="cuda")
    return p, cg, x, dev(kd)


@pytest.mark.parametrize("dims", [(32, 32), (16, 16, 16)], ids=["2d", "3d"])
def test_cg_matches_host_cg_on_oracle(oracle, env, dims):
    torch, capi = env
    traj, dens, sens, kd = _problem(dims)
    alpha, iters = 1e-3, 5
    op = oracle.Operator(traj, dims, 3, 8, 2.0, dens=dens, sens=sens, order="stable")
    y = op.adjoint(kd).astype(np.complex128)
    x = np.zeros_like(y)
    r = y.copy()
    p = r.copy()
    rr = np.vdot(r, r).real
    for _ in range(iters):
        Ap = op.adjoint(op.forward(p.astype(np.complex64))).astype(np.complex128) + alpha * p
        a = rr / np.vdot(p, Ap).real
        x = x + a * p
        rn = r - a * Ap
        b = np.vdot(rn, rn).real / rr
        r = rn
        rr = np.vdot(r, r).real
        p = r + b * p
    pl, cg, xg, kdev = _gpu_cg(torch, capi, dims, traj, dens, sens, kd, iters, alpha, True)
    cg.solve(kdev, xg, iters, use_graph=True)
    torch.cuda.synchronize()
    assert oracle.rel_l2(xg.cpu().numpy(), x.astype(np.complex64)) < 1e-4


def test_graph_replay_equals_eager(oracle, env):
    torch, capi = env
    dims = (48, 40)
    traj, dens, sens, kd = _problem(dims, C=4, seed=3)
    pl, cg, xg, kdev = _gpu_cg(torch, capi, dims, traj, dens, sens, kd, 10, 1e-6, False)
    cg.solve(kdev, xg, 10, use_graph=False)
    torch.cuda.synchronize()
    eager = xg.cpu().numpy()
    for rep in range(3):
        print("replay", rep, flush=True)
        xg.zero_()
        cg.solve(kdev, xg, 10, use_graph=True)
        torch.cuda.synchronize()
        got = xg.cpu().numpy()
        assert oracle.rel_l2(got, eager) < 1e-6, rep
        # eager work between replays
        cg.solve(kdev, xg, 3, use_graph=False)
        torch.cuda.synchronize()


def test_graph_recaptured_after_plan_changes(oracle, env):
    """The captured iteration holds the plan's device pointers: after
    set_sens (new sensitivities) and after set_trajectory (new tables of the
    same length), the next graph solve captures again and equals an eager
    solve of the changed plan (rel-L2 <= 1e-6)."""
    torch, capi = env
    dims = (40, 40)
    traj, dens, sens, kd = _problem(dims, C=3, seed=5)
    pl, cg, xg, kdev = _gpu_cg(torch, capi, dims, traj, dens, sens, kd, 6, 1e-4, True)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    cg.solve(kdev, xg, 6, use_graph=True)
    torch.cuda.synchronize()
    first = xg.cpu().numpy()
    # new sensitivities
    sens2 = W.gaussian_sens(dims, 3, seed=9)
    pl.set_sens(dev(sens2), 3)
    xg.zero_()
    cg.solve(kdev, xg, 6, use_graph=True)
    torch.cuda.synchronize()
    graph2 = xg.cpu().numpy()
    cg.solve(kdev, xg, 6, use_graph=False)
    torch.cuda.synchronize()
    eager2 = xg.cpu().numpy()
    print("sens change: graph vs eager", oracle.rel_l2(graph2, eager2), flush=True)
    assert oracle.rel_l2(graph2, eager2) < 1e-6
    assert oracle.rel_l2(graph2, first) > 1e-3  # the new maps took effect
    # new trajectory of the same length (rotated spokes)
    th = 0.37
    rot = np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]], np.float32)
    traj2 = np.clip(rot @ traj, -0.5, 0.4999).astype(np.float32)
    pl.set_trajectory(dev(traj2), dev(dens))
    xg.zero_()
    cg.solve(kdev, xg, 6, use_graph=True)
    torch.cuda.synchronize()
    graph3 = xg.cpu().numpy()
    cg.solve(kdev, xg, 6, use_graph=False)
    torch.cuda.synchronize()
    eager3 = xg.cpu().numpy()
    print("trajectory change: graph vs eager", oracle.rel_l2(graph3, eager3), flush=True)
    assert oracle.rel_l2(graph3, eager3) < 1e-6
    # a different sample count cannot reuse the solver's vectors
    pl.set_trajectory(dev(traj[:, :-8].copy()), dev(dens[:-8].copy()))
    with pytest.raises(capi.GnufftError) as ei:
        cg.solve(kdev, xg, 2, use_graph=True)
    assert ei.value.code == capi.ERR_STATE


def test_batch_rejects_changed_plan(env):
    torch, capi = env
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    traj = W.radial_2d(16, 32)
    plans = []
    for n in (24, 32):
        p = capi.Plan((n, n), 3, 8, 2.0)
        p.set_trajectory(dev(traj))
        plans.append(p)
    b = capi.Batch(plans, [2, 2])
    kd = [dev(W.complex_normal(np.random.default_rng(0), (2, traj.shape[1]))) for _ in plans]
    out = [torch.empty((2, n, n), dtype=torch.complex64, device="cuda") for n in (24, 32)]
    b.adjoint(kd, out)
    torch.cuda.synchronize()
    plans[1].set_trajectory(dev(traj))
    with pytest.raises(capi.GnufftError) as ei:
        b.adjoint(kd, out)
    assert ei.value.code == capi.ERR_STATE
    b.close()


def test_serving_mix_equals_alone(oracle, env):
    torch, capi = env
    from gpunufft_amd import serve

    reqs = serve.serving_mix(n_requests=6, rate_hz=500.0, seed=11, iterations=4, small=True)
    print("mix", [r.img_dims for r in reqs], flush=True)
    srv = serve.Server(slots=3, keep_results=True)
    summary = srv.run(reqs)
    assert summary["requests"] == 6
    mixed = [r.result for r in reqs]
    alone = serve.Server(slots=1, keep_results=True)
    for r, m in zip(reqs, mixed):
        r.arrival = 0.0
        alone.run([r])
        assert oracle.rel_l2(m, r.result) < 1e-6, r.rid


def _cg_objective(torch, capi, req, plan, x):
    """CG's objective f(x) = 1/2 (|A x|^2 + alpha |x|^2) - Re <A^H y, x> of
    the request's operator A (sensitivities, density), in float64 on the
    host from the device products. CG from x = 0 minimises f over growing
    Krylov spaces, so f(0) = 0 > f(x_1) >= f(x_k) in exact arithmetic."""
    C, M = req.kdata.shape
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    b = torch.empty_like(x)
    plan.adjoint(dev(req.kdata), b, C)
    ax = torch.empty((C, M), dtype=torch.complex64, device="cuda")
    plan.forward(x, ax, C)
    xh = x.cpu().numpy().astype(np.complex128).ravel()
    bh = b.cpu().numpy().astype(np.complex128).ravel()
    axh = ax.cpu().numpy().astype(np.complex128).ravel()
    return 0.5 * (np.vdot(axh, axh).real + req.alpha * np.vdot(xh, xh).real) - np.vdot(bh, xh).real


def test_serving_mix_full_size(oracle, env):
    """Full-size serving requests (2-D N 368 with 28 coils on a 736-point
    grid, N 192 and 224 with 22 and 25 coils, 3-D 152^3 with 8 coils on a
    304^3 grid; osf 2): every result of the concurrent mix equals the
    request served alone (rel-L2 <= 1e-6), and the solver's objective falls
    from 0 at x = 0 to f(x_1) and further to f(x_6) for the largest 2-D and
    the 3-D request."""
    torch, capi = env
    from gpunufft_amd import serve

    reqs = serve.serving_mix(n_requests=4, rate_hz=500.0, seed=1, iterations=6)
    print("full-size mix", [(r.img_dims, r.coils) for r in reqs], flush=True)
    srv = serve.Server(slots=3, keep_results=True)
    assert srv.run(reqs)["requests"] == 4
    mixed = [r.result for r in reqs]
    alone = serve.Server(slots=1, keep_results=True)
    for r, m in zip(reqs, mixed):
        r.arrival = 0.0
        alone.run([r])
        print("alone", r.rid, flush=True)
        assert oracle.rel_l2(m, r.result) < 1e-6, r.rid
    for r in (reqs[0], reqs[3]):
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
        p = capi.Plan(r.img_dims, r.kernel_width, r.sector_width, r.osf)
        p.set_trajectory(dev(r.traj), dev(r.dens))
        p.set_sens(dev(r.sens), r.coils)
        cg = capi.CGSense(p, r.coils, r.alpha)
        kd = dev(r.kdata)
        f = {}
        for k in (1, 6):
            x = torch.empty((1,) + tuple(r.img_dims[::-1]), dtype=torch.complex64, device="cuda")
            cg.solve(kd, x, k, True)
            torch.cuda.synchronize()
            if k == 6:
                assert oracle.rel_l2(r.result, x.cpu().numpy()) < 1e-6, r.rid
            f[k] = _cg_objective(torch, capi, r, p, x)
        print("objective", r.rid, f, flush=True)
        assert f[1] < 0.0 and f[6] < f[1], (r.rid, f)
        cg.close()
