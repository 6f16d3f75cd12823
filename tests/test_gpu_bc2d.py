"""Slip-wall and state boundaries of 2D Euler on the GPU (ghost2 in k_rhs2 and
k_rhs2c) against the oracle's restatement of Euler::WallBC
(src/EulerBC.hpp:43-62) and StateBC (src/BC.cpp:17-28).

Bar: relative L2 per field <= 1e-12 for one residual, <= 1e-11 after steps.
"""
import numpy as np
import pytest

from dgm_amd import dgx, mesh as meshgen
from oracle_lib import Oracle
from test_gpu_parity import TOL_RHS, TOL_STEP, projected_state, rel_l2

pytestmark = pytest.mark.gpu

GAMMA = 1.4
FREE = [1.0, 0.4, 0.05, 1.0 / (GAMMA - 1) + 0.5 * (0.4 ** 2 + 0.05 ** 2)]


def _state(x):
    rho = 1.0 + 0.1 * np.sin(2 * x[0]) * np.cos(3 * x[1])
    u = 0.4 + 0.05 * np.cos(x[1])
    v = 0.05 * np.sin(3 * x[0])
    p = 1.0 + 0.05 * np.cos(2 * x[0] + x[1])
    return np.stack([rho, rho * u, rho * v, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)])


@pytest.mark.parametrize("p,flux", [(2, 0), (5, 1), (8, 0)])
def test_rhs_walls_and_state(p, flux):
    print(f"[bc2d] affine p={p} flux={flux}", flush=True)
    m = meshgen.box_quads(5, 3, 0.0, 2.0, 0.0, 1.0, bc=(0, 1, 0, 1))
    orc = Oracle(m, p, 1, gamma=GAMMA, flux=flux)
    orc.set_bc_state(FREE)
    x = orc.coords()
    U = projected_state(orc, _state(x))
    F_ref = orc.rhs(U)
    gpu = dgx.GPUDomain(m, p, dgx.DGX_EULER, gamma=GAMMA, flux=flux, bc_state=FREE + [0.0])
    gpu.upload(U)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"[bc2d] rel L2 {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    orc.set_state(U)
    orc.advance(dgx.DGX_LSRK14, 1e-3, 3)
    gpu.advance(dgx.DGX_LSRK14, 1e-3, 3)
    err = rel_l2(gpu.download(), orc.state())
    assert np.all(err <= TOL_STEP), err
    gpu.close()


@pytest.mark.parametrize("p", [3, 6])
def test_rhs_curved_bump_channel(p):
    """A channel whose bottom wall carries a bump: the curved kernel's walls."""
    print(f"[bc2d] curved p={p}", flush=True)
    m = meshgen.box_quads(6, 3, 0.0, 3.0, 0.0, 1.0, bc=(0, 1, 0, 1))
    bump = lambda x, y: (x, y + 0.1 * (1 - np.cos(2 * np.pi * x / 3.0)) / 2 * (1.0 - y))
    xn = meshgen.mapped_nodes(m, p, bump)
    orc = Oracle(m, p, 1, gamma=GAMMA)
    orc.set_bc_state(FREE)
    orc.set_nodes(xn)
    U = projected_state(orc, _state(xn))
    F_ref = orc.rhs(U)
    gpu = dgx.GPUDomain(m, p, dgx.DGX_EULER, gamma=GAMMA, bc_state=FREE + [0.0], xnodes=xn)
    gpu.upload(U)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"[bc2d] rel L2 {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    gpu.close()


def test_batched_serving_with_walls():
    """Requests with walls join the batched groups like periodic ones."""
    from dgm_amd import serve
    m = meshgen.box_quads(4, 3, 0.0, 2.0, 0.0, 1.0, bc=(0, 1, 0, 1))
    p, dt, n = 3, 1e-3, 20
    orc = Oracle(m, p, 1, gamma=GAMMA)
    orc.set_bc_state(FREE)
    U = projected_state(orc, _state(orc.coords()))
    orc.set_state(U)
    orc.advance(dgx.DGX_LSRK14, dt, n)
    srv = serve.Server(hbm_budget=1 << 34, max_active=2, steps_per_slice=5, batch=True)
    rid = srv.submit(m, p, dgx.DGX_EULER, U, dgx.DGX_LSRK14, dt, n, bc_state=FREE)
    srv.run_until_done(timeout_s=60.0)
    out = srv.result(rid)[0]
    srv.close()
    err = rel_l2(out, orc.state())
    print(f"[bc2d] served rel L2 {err}", flush=True)
    assert np.all(err <= TOL_STEP), err
