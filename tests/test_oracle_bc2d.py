"""Slip-wall and state boundaries of the oracle's 2D Euler residual
(Euler::WallBC, src/EulerBC.hpp:43-62; StateBC, src/BC.cpp:17-28).

Pinned, as in 3D, by what the boundary fixes:
* a uniform flow parallel to the walls, with the free stream imposed on the
  other boundaries, is an exact steady solution (Lax-Friedrichs and Roe);
* the domain integral of the mass residual equals 1/2 of the wall integral of
  m.n: the ghost keeps rho and has no normal momentum, so the Lax-Friedrichs
  mass flux there is 1/2 m.n, and the interior faces cancel;
* a state boundary holding a uniform flow's own values is transparent;
* the wall ghost keeps the internal energy: a state at rest next to a wall is
  steady for any pressure.
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import mesh as meshgen
from oracle_lib import Oracle
from test_oracle_invariants import _integral

GAMMA = 1.4


def _uniform(x, rho, u, v, p):
    one = np.ones_like(x[0])
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho * one, rho * u * one, rho * v * one, E * one])


@pytest.mark.parametrize("flux", [0, 1])
def test_parallel_flow_is_steady(flux):
    m = meshgen.box_quads(3, 2, 0.0, 1.0, 0.0, 0.7, bc=(0, 1, 0, 1))
    orc = Oracle(m, 4, 1, gamma=GAMMA, flux=flux)
    U = _uniform(orc.coords(), 1.2, 0.5, 0.0, 1.0 / GAMMA)
    orc.set_bc_state(U[:, 0, 0])
    F = orc.rhs(U)
    assert np.max(np.abs(F)) < 1e-12, np.max(np.abs(F))


def test_rest_state_next_to_walls_is_steady():
    m = meshgen.box_quads(2, 3, 0.0, 1.0, 0.0, 1.0, bc=(0, 0, 0, 0))
    for flux in (0, 1):
        orc = Oracle(m, 3, 1, gamma=GAMMA, flux=flux)
        U = _uniform(orc.coords(), 0.8, 0.0, 0.0, 2.5)
        assert np.max(np.abs(orc.rhs(U))) < 1e-12


def test_wall_mass_budget():
    p, H = 3, 0.8
    m = meshgen.box_quads(3, 3, 0.0, 1.0, 0.0, H, bc=(0, None, 0, None))
    orc = Oracle(m, p, 1, gamma=GAMMA)
    x = orc.coords()
    rng = np.random.default_rng(4)
    U = _uniform(x, 1.0, 0.3, 0.1, 1.0 / GAMMA)
    for v in range(4):
        a = rng.uniform(-0.05, 0.05, 3)
        U[v] += a[0] * np.sin(2 * np.pi * x[0]) + a[1] * np.cos(2 * np.pi * x[1]) + a[2] * x[1] * (H - x[1] + 0.3)
    orc.set_state(U)
    orc.project_state()
    U = orc.state()
    F = orc.rhs(U)
    lhs = _integral(orc, F)[0]
    # 1/2 of the wall integral of m.n by GLL quadrature of the state's trace
    # (degree p in x, exact with q = p + 2 points): n = -y at y = 0, +y at y = H
    b = orc.basis()
    w = b["w"]
    q = p + 2
    rhs = 0.0
    for e in range(m.ne):
        hx = m.vert[e, 0].max() - m.vert[e, 0].min()
        my = U[2, e].reshape(q, q)                     # node i*q + j at (a_i, b_j)
        if abs(m.vert[e, 1].min()) < 1e-12:
            rhs += -0.5 * np.dot(w, my[:, 0]) * hx / 2.0
        if abs(m.vert[e, 1].max() - H) < 1e-12:
            rhs += 0.5 * np.dot(w, my[:, -1]) * hx / 2.0
    assert abs(rhs) > 1e-4
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(rhs)), (lhs, rhs)


def test_state_boundary_of_a_uniform_flow_matches_periodic():
    m = meshgen.box_quads(2, 2, 0.0, 1.0, 0.0, 1.0, bc=(1, 1, 1, 1))
    per = meshgen.periodic_quads(2, 2, 0.0, 1.0, 0.0, 1.0)
    for flux in (0, 1):
        a, b = Oracle(m, 3, 1, gamma=GAMMA, flux=flux), Oracle(per, 3, 1, gamma=GAMMA, flux=flux)
        U = _uniform(a.coords(), 0.9, 0.2, 0.1, 0.8)
        a.set_bc_state(U[:, 0, 0])
        assert np.max(np.abs(a.rhs(U) - b.rhs(U))) < 1e-12
