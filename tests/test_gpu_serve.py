"""Serving loop (dgx_server_*): continuous batching of forward solves.

* Each result is bitwise identical whether the request is served alone or in
  the mix: a request's kernels read and write only its own buffers.
* Served results match one uninterrupted dgx_advance of the same request: to
  1e-13 for the batched 2D path, whose residual kernel fuses the register
  update (same fma forms as the direct path's axpby and axpy); to rounding for the fused 3D path, where each slice starts by
  recomputing the face traces from the state instead of taking them from the
  previous stage's epilogue (1e-12 FP64, 1e-6 mixed).
* Batched groups (2D Euler and shallow water of one order, one launch per
  stage for all members) give the per-request path's results.
* Shallow-water and Euler requests share the loop; a shallow-water request
  with the balance term keeps its initial state, as the reference's does.
* Admission respects the memory budget and the slot count.
* A request larger than the budget on its own is rejected at submission.
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import dgx, ic, mesh as meshgen, serve

pytestmark = pytest.mark.gpu


def _euler2d_job(rng, n, p, steps):
    m = meshgen.periodic_quads(n, n, 0.0, 1.0, 0.0, 1.0)
    g = dgx.GPUDomain(m, p, dgx.DGX_EULER)
    x = g.coords()
    g.close()
    base = [1.0, 0.5, 0.0, 1.0 / 1.4 / 0.4 + 0.5 * 0.25]
    U = np.stack([b + ic.fourier_modes(x, (1.0, 1.0), 8, 0.05, rng) for b in base])
    h = 1.0 / n
    dt = 0.2 * h / ((p + 1) ** 2 * 1.6)
    return dict(mesh=m, p=p, eqntype=dgx.DGX_EULER, U0=U, inttype=dgx.DGX_LSRK14, dt=dt, nsteps=steps)


def _shallow_job(rng, n, p, steps, balance=False):
    m = meshgen.periodic_quads(n, n, 0.0, 1.0, 0.0, 1.0)
    x = meshgen.quad_nodes(m, p)
    U = np.stack([b + ic.fourier_modes(x, (1.0, 1.0), 8, 0.05, rng) for b in (1.0, 0.0, 0.0)])
    dt = 0.2 * (1.0 / n) / ((p + 1) ** 2 * 1.2)
    return dict(mesh=m, p=p, eqntype=dgx.DGX_SHALLOW, U0=U, inttype=dgx.DGX_LSRK14, dt=dt, nsteps=steps,
                balance=balance)


def _ns3d_job(rng, n, p, steps, precision=dgx.DGX_FP64):
    m = meshgen.periodic_hexes(n, n, n, (0.0,) * 3, (2 * np.pi,) * 3)
    g = dgx.GPUDomain(m, p, dgx.DGX_NAVIER_STOKES3D, vis=0.01, Pr=0.71)
    x = g.coords()
    g.close()
    U = ic.taylor_green(x[0], x[1], x[2])
    dt = 0.5 * (2 * np.pi / n) / ((p + 1) ** 2 * 12.0)
    return dict(mesh=m, p=p, eqntype=dgx.DGX_NAVIER_STOKES3D, U0=U, inttype=dgx.DGX_LSRK14, dt=dt,
                nsteps=steps, vis=0.01, Pr=0.71, precision=precision)


def _served_alone(job, steps_per_slice=5, batch=True):
    srv = serve.Server(hbm_budget=1 << 34, max_active=1, steps_per_slice=steps_per_slice, batch=batch)
    rid = srv.submit(**job)
    srv.run_until_done(timeout_s=60.0)
    U = srv.result(rid)[0]
    srv.close()
    return U


def _alone(job):
    g = dgx.GPUDomain(job["mesh"], job["p"], job["eqntype"], vis=job.get("vis", 0.0), Pr=job.get("Pr", 0.72),
                      precision=job.get("precision", dgx.DGX_FP64), balance=job.get("balance", True))
    g.upload(job["U0"])
    g.advance(job["inttype"], job["dt"], job["nsteps"])
    g.synchronize()
    U = g.download(dgx.DGX_NODAL)
    g.close()
    return U


def _mix():
    rng = np.random.default_rng(2024)
    jobs = []
    for n, p, steps in [(4, 2, 7), (8, 4, 13), (6, 3, 20), (12, 5, 11), (4, 6, 16), (10, 2, 9),
                        (8, 8, 12), (5, 4, 25)]:
        jobs.append(_euler2d_job(rng, n, p, steps))
    for n, p, steps in [(6, 3, 14), (9, 5, 11), (4, 8, 17)]:
        jobs.append(_shallow_job(rng, n, p, steps))
    jobs.append(_shallow_job(rng, 5, 4, 12, balance=True))
    jobs.append(_ns3d_job(rng, 2, 4, 6))
    jobs.append(_ns3d_job(rng, 2, 7, 5))
    jobs.append(_ns3d_job(rng, 2, 5, 7, precision=dgx.DGX_MIXED))
    return jobs


def test_results_identical_alone_and_in_the_mix():
    jobs = _mix()
    sizes = [serve.domain_bytes(j["mesh"], j["p"], j["eqntype"], j.get("precision", 0)) for j in jobs]
    budget = sum(sorted(sizes)[-3:])                   # only a few requests fit at once
    srv = serve.Server(hbm_budget=budget, max_active=4, steps_per_slice=5)
    rng = np.random.default_rng(7)
    arrivals = serve.poisson_arrivals(len(jobs), rate=200.0, rng=rng)
    print(f"[serve] {len(jobs)} requests, budget {budget / 2**20:.1f} MiB", flush=True)
    ids = serve.serve_timeline(srv, jobs, arrivals, timeout_s=120.0)
    peak_bytes, peak_active = srv.peak()
    assert peak_bytes <= budget and peak_active <= 4
    for rid, job in zip(ids, jobs):
        assert srv.status(rid) == serve.DONE
        U, t = srv.result(rid)
        assert t[0] <= t[1] <= t[2]
        alone = _served_alone(job)
        diff = np.max(np.abs(U - alone))
        print(f"[serve] request {rid}: max |mix - alone| = {diff:.3e}", flush=True)
        assert np.array_equal(U, alone), diff
        direct = _alone(job)
        rel = np.max(np.abs(U - direct)) / np.max(np.abs(direct))
        print(f"[serve] request {rid}: max |mix - direct| / max |direct| = {rel:.3e}", flush=True)
        if job["mesh"].nsd == 2:
            assert rel <= 1e-13, rel
        else:
            tol = 1e-6 if job.get("precision", 0) == dgx.DGX_MIXED else 1e-12
            assert rel <= tol, rel
    srv.close()


def test_batched_groups_match_unbatched_service():
    """Batched 2D requests (shared launches per stage, fused register update)
    give the per-request path's results: requests of two orders and both
    physics, with different step counts, so members leave their groups at
    different slice boundaries."""
    rng = np.random.default_rng(21)
    jobs = [_euler2d_job(rng, 5, 3, 9), _euler2d_job(rng, 7, 3, 23), _shallow_job(rng, 6, 3, 17),
            _shallow_job(rng, 4, 5, 6), _euler2d_job(rng, 6, 5, 14), _shallow_job(rng, 5, 3, 11, balance=True)]
    out = []
    for batch in (True, False):
        srv = serve.Server(hbm_budget=1 << 32, max_active=8, steps_per_slice=4, batch=batch)
        ids = [srv.submit(**j) for j in jobs]
        srv.run_until_done(timeout_s=60.0)
        out.append([srv.result(i)[0] for i in ids])
        srv.close()
    for a, b in zip(*out):
        rel = np.max(np.abs(a - b)) / np.max(np.abs(b))
        print(f"[serve] batched vs per-request: {rel:.3e} (bitwise: {np.array_equal(a, b)})", flush=True)
        assert rel <= 1e-13, rel


@pytest.mark.parametrize("slice_steps", [4, 9])
def test_persistent_slices_match_per_stage_launches(slice_steps):
    """A group's slice as one persistent cooperative kernel (grid-wide barriers
    between stages, alternating trace buffers) gives bitwise the results of
    two launches per stage: both physics, two orders, members leaving at
    different slice boundaries, later arrivals joining running groups."""
    rng = np.random.default_rng(47)
    jobs = [_euler2d_job(rng, 5, 3, 9), _euler2d_job(rng, 7, 3, 23), _shallow_job(rng, 6, 3, 17),
            _shallow_job(rng, 4, 5, 6), _euler2d_job(rng, 6, 5, 14), _shallow_job(rng, 5, 3, 11, balance=True),
            _euler2d_job(rng, 9, 3, 12)]
    out = []
    for persistent in (True, False):
        print(f"[serve] persistent={persistent} slice={slice_steps}", flush=True)
        srv = serve.Server(hbm_budget=1 << 32, max_active=8, steps_per_slice=slice_steps, batch=True,
                           persistent=persistent)
        ids = [srv.submit(**j) for j in jobs[:4]]
        srv.progress()
        ids += [srv.submit(**j) for j in jobs[4:]]
        srv.run_until_done(timeout_s=60.0)
        for i in ids:
            assert srv.status(i) == serve.DONE
        out.append([srv.result(i)[0] for i in ids])
        srv.close()
    for a, b in zip(*out):
        print(f"[serve] persistent vs per-stage: max diff {np.max(np.abs(a - b)):.3e}", flush=True)
        assert np.array_equal(a, b)


def test_member_joining_during_a_slice_steps_from_the_next():
    """A request admitted while its group's slice is in flight joins the group
    at the next slice and still takes exactly its own number of steps."""
    rng = np.random.default_rng(33)
    first, late = _euler2d_job(rng, 16, 4, 30), _euler2d_job(rng, 12, 4, 17)
    srv = serve.Server(hbm_budget=1 << 32, max_active=4, steps_per_slice=10)
    a = srv.submit(**first)
    srv.progress()                                      # admits `first` and enqueues its first slice
    b = srv.submit(**late)
    srv.progress()                                      # `late` joins while that slice may still run
    srv.run_until_done(timeout_s=60.0)
    Ua, Ub = srv.result(a)[0], srv.result(b)[0]
    srv.close()
    assert np.array_equal(Ua, _served_alone(first))
    assert np.array_equal(Ub, _served_alone(late))


def test_graph_replay_matches_direct_launches():
    rng = np.random.default_rng(3)
    job = _euler2d_job(rng, 6, 4, 23)
    out = []
    for graphs in (True, False):
        srv = serve.Server(hbm_budget=1 << 30, max_active=2, steps_per_slice=4, use_graphs=graphs, batch=False)
        rid = srv.submit(**job)
        srv.run_until_done(timeout_s=60.0)
        out.append(srv.result(rid)[0])
        srv.close()
    assert np.array_equal(out[0], out[1])


def test_shallow_balance_keeps_the_initial_state():
    """With the balance term (on by default, as in the reference) the first
    residual is subtracted from every residual, so the initial state is a
    fixed point of the served solve; without it the state evolves."""
    rng = np.random.default_rng(11)
    for balance in (True, False):
        job = _shallow_job(rng, 6, 4, 12, balance=balance)
        U = _served_alone(job, steps_per_slice=4)
        moved = np.max(np.abs(U - job["U0"]))
        if balance:
            g = dgx.GPUDomain(job["mesh"], job["p"], job["eqntype"], balance=False)
            g.upload(job["U0"])
            U0p = g.download()                           # the projected initial state
            g.close()
            assert np.max(np.abs(U - U0p)) < 1e-13, np.max(np.abs(U - U0p))
        else:
            assert moved > 1e-4, moved


def test_oversized_request_rejected():
    rng = np.random.default_rng(5)
    job = _euler2d_job(rng, 8, 4, 3)
    need = serve.domain_bytes(job["mesh"], job["p"], job["eqntype"])
    srv = serve.Server(hbm_budget=need - 1, max_active=2, steps_per_slice=2)
    with pytest.raises(dgx.DgxError):
        srv.submit(**job)
    srv.close()
