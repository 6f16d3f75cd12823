"""The FP32-mixed quad path (DGX_MIXED on affine quads: element tensors in LDS
and their contractions in FP32; state, RK registers, face traces, metric
terms and Riemann fluxes FP64) against the FP64 path, which the oracle pins
to 1e-13 (tests/test_gpu_parity.py, tests/test_gpu_shallow.py).

Stated errors (measured maxima in brackets):
* one residual, relative L2 per conserved field: <= 1e-4 [5.2e-5] (FP32
  contractions of O(1) fluxes whose divergence is a fraction of them);
* the state after 200 LSRK14 steps at CFL ~0.2, relative L2 per field:
  <= 5e-4 [1.5e-4].  The largest are the fields that are only the 0.05
  perturbation about a zero mean (rho v about a free stream along x, the
  shallow-water momenta about rest), measured against their own small norm;
* the same error against the norm of the whole state: <= 1e-5 [3.2e-6].
The serving loop batches mixed requests in groups of their own; a mixed
request's result is bitwise the same alone and in a mix.
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import dgx, ic, mesh as meshgen, serve

pytestmark = pytest.mark.gpu

TOL_RHS_MIXED = 1e-4
TOL_STATE_MIXED = 5e-4
TOL_STATE_MIXED_OF_STATE = 1e-5


def rel_fields(a, b):
    return np.array([np.linalg.norm(a[v] - b[v]) / max(np.linalg.norm(b[v]), 1e-300) for v in range(len(b))])


def rel_of_state(a, b):
    """per field, the error against the norm of the whole state"""
    return np.array([np.linalg.norm(a[v] - b[v]) for v in range(len(b))]) / np.linalg.norm(b)


def _state(x, seed, base):
    """base (rest or free stream) plus 8 seeded Fourier modes of amplitude 0.05 per field"""
    rng = np.random.default_rng(seed)
    return np.stack([b + ic.fourier_modes(x, (10.0, 10.0), 8, 0.05, rng) for b in base])


def _euler_state(x, seed):
    return _state(x, seed, [1.0, 0.5, 0.0, (1.0 / 1.4) / 0.4 + 0.125])


def _shallow_state(x, seed):
    return _state(x, seed, [1.0, 0.0, 0.0])


def _pair(mesh, p, eqn, flux=0, bc_state=None, balance=True):
    kw = dict(flux=flux, bc_state=bc_state, balance=balance)
    return (dgx.GPUDomain(mesh, p, eqn, precision=dgx.DGX_FP64, **kw),
            dgx.GPUDomain(mesh, p, eqn, precision=dgx.DGX_MIXED, **kw))


CASES = [("euler", 2, 0, "periodic"), ("euler", 5, 0, "periodic"), ("euler", 8, 1, "periodic"),
         ("euler", 6, 0, "walls"), ("shallow", 3, 0, "periodic"), ("shallow", 7, 0, "periodic")]


def _mesh(kind):
    if kind == "walls":
        return meshgen.box_quads(6, 5, 0.0, 10.0, 0.0, 10.0, bc=(0, 1, 0, 1))
    return meshgen.periodic_quads(6, 5, 0.0, 10.0, 0.0, 10.0)


@pytest.mark.parametrize("phys,p,flux,kind", CASES)
def test_mixed_residual_and_steps_against_fp64(phys, p, flux, kind):
    mesh = _mesh(kind)
    eqn = dgx.DGX_EULER if phys == "euler" else dgx.DGX_SHALLOW
    bc = [1.0, 0.5, 0.0, (1.0 / 1.4) / 0.4 + 0.125] if kind == "walls" else None
    g64, g32 = _pair(mesh, p, eqn, flux=flux, bc_state=bc, balance=False)
    x = g64.coords()
    U = _euler_state(x, p) if phys == "euler" else _shallow_state(x, p)
    g64.upload(U)
    g32.upload(U)
    assert np.array_equal(g64.download(dgx.DGX_MODAL), g32.download(dgx.DGX_MODAL))
    e_rhs = rel_fields(g32.rhs(dgx.DGX_MODAL), g64.rhs(dgx.DGX_MODAL))
    print(f"[mixed2d] {phys} p={p} flux={flux} {kind}: residual rel L2 per field {e_rhs}", flush=True)
    assert np.all(e_rhs <= TOL_RHS_MIXED), e_rhs
    speed = 1.5 if phys == "euler" else 1.0
    dt = 0.2 * (10.0 / 6) / ((p + 1) ** 2 * speed)
    g64.advance(dgx.DGX_LSRK14, dt, 200)
    g32.advance(dgx.DGX_LSRK14, dt, 200)
    U32, U64 = g32.download(dgx.DGX_MODAL), g64.download(dgx.DGX_MODAL)
    e_state, e_whole = rel_fields(U32, U64), rel_of_state(U32, U64)
    print(f"[mixed2d] {phys} p={p} flux={flux} {kind}: state after 200 LSRK14 steps rel L2 per field {e_state}, "
          f"against the whole state {e_whole}", flush=True)
    assert np.all(e_state <= TOL_STATE_MIXED), e_state
    assert np.all(e_whole <= TOL_STATE_MIXED_OF_STATE), e_whole
    g64.close()
    g32.close()


def _job(rng, n, p, steps, eqn, precision):
    m = meshgen.periodic_quads(n, n, 0.0, 10.0, 0.0, 10.0)
    x = meshgen.quad_nodes(m, p)
    U = _euler_state(x, int(rng.integers(1 << 30))) if eqn == dgx.DGX_EULER else \
        _shallow_state(x, int(rng.integers(1 << 30)))
    speed = 1.5 if eqn == dgx.DGX_EULER else 1.0
    dt = 0.2 * (10.0 / n) / ((p + 1) ** 2 * speed)
    job = dict(mesh=m, p=p, eqntype=eqn, U0=U, inttype=dgx.DGX_LSRK14, dt=dt, nsteps=steps, precision=precision)
    if eqn == dgx.DGX_SHALLOW:
        job["balance"] = False
    return job


def test_mixed_requests_served_alone_and_in_a_mix():
    """Mixed and FP64 requests of the same orders in one server: they batch
    apart, each result is bitwise the same alone and in the mix, and the
    mixed results stay within the stated error of their FP64 twins."""
    rng = np.random.default_rng(77)
    specs = [(6, 3, 12, dgx.DGX_EULER), (9, 5, 9, dgx.DGX_EULER), (5, 3, 15, dgx.DGX_SHALLOW),
             (7, 6, 8, dgx.DGX_SHALLOW)]
    jobs = []
    for n, p, steps, eqn in specs:
        j = _job(rng, n, p, steps, eqn, dgx.DGX_MIXED)
        jobs.append(j)
        jobs.append(dict(j, precision=dgx.DGX_FP64))
    srv = serve.Server(hbm_budget=1 << 32, max_active=8, steps_per_slice=4)
    ids = [srv.submit(**j) for j in jobs]
    srv.run_until_done(timeout_s=60.0)
    mix = [srv.result(i)[0] for i in ids]
    srv.close()
    for k, j in enumerate(jobs):
        alone = serve.Server(hbm_budget=1 << 32, max_active=1, steps_per_slice=4)
        rid = alone.submit(**j)
        alone.run_until_done(timeout_s=60.0)
        U = alone.result(rid)[0]
        alone.close()
        print(f"[mixed2d] request {k} precision {j['precision']}: max |mix - alone| {np.max(np.abs(mix[k] - U))}",
              flush=True)
        assert np.array_equal(mix[k], U)
    for k in range(0, len(jobs), 2):
        e, ew = rel_fields(mix[k], mix[k + 1]), rel_of_state(mix[k], mix[k + 1])
        print(f"[mixed2d] request {k}: mixed vs FP64 rel L2 per field {e}, against the whole state {ew}", flush=True)
        assert np.all(e <= TOL_STATE_MIXED) and np.all(ew <= TOL_STATE_MIXED_OF_STATE), (e, ew)
