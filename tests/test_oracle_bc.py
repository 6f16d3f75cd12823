"""Slip-wall and state boundaries of the oracle's 3D Euler residual
(Euler3d::WallBC, src/Euler3d.cpp:876-909; StateBC, src/BC.cpp:17-28).

The reference's own fixtures with walls (runs/chan) add viscous terms and bulk
forcing outside this path, so these are pinned by what the boundary fixes:
* a uniform flow parallel to a wall, with the free stream imposed on the
  other boundary, is an exact steady solution: F = 0 to rounding;
* the domain integral of the mass residual on a box with two walls equals
  1/2 of the wall integral of m.n: the wall's ghost keeps rho and removes the
  normal momentum, so the Lax-Friedrichs mass flux there is 1/2 m.n, and the
  interior faces cancel;
* a state boundary holding the interior's own trace values, i.e. a
  transparent boundary for a uniform state, gives the same residual as the
  periodic box.
"""
from __future__ import annotations

import numpy as np

from dgm_amd import mesh as meshgen
from oracle_lib import Oracle
from test_oracle_invariants import _integral

GAMMA = 1.4


def _uniform(x, rho, u, v, w, p):
    one = np.ones_like(x[0])
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho * one, rho * u * one, rho * v * one, rho * w * one, E * one])


def test_parallel_flow_is_steady():
    m = meshgen.channel_hexes(3, 2, 3, (0, 0, 0), (1.0, 1.0, 0.7), bc_bottom=0, bc_top=1)
    orc = Oracle(m, 4, 8, gamma=GAMMA)
    U = _uniform(orc.coords(), 1.2, 0.5, -0.2, 0.0, 1.0 / GAMMA)
    orc.set_bc_state(U[:, 0, 0])
    F = orc.rhs(U)
    assert np.max(np.abs(F)) < 1e-12, np.max(np.abs(F))


def test_wall_mass_budget():
    nz, p, H = 3, 3, 0.8
    m = meshgen.channel_hexes(2, 2, nz, (0, 0, 0), (1.0, 1.0, H), bc_bottom=0, bc_top=0)
    orc = Oracle(m, p, 8, gamma=GAMMA)
    x = orc.coords()
    rng = np.random.default_rng(2)
    U = _uniform(x, 1.0, 0.3, 0.1, 0.0, 1.0 / GAMMA)
    for v in range(5):
        a = rng.uniform(-0.05, 0.05, 3)
        U[v] += a[0] * np.sin(2 * np.pi * x[0]) + a[1] * np.cos(2 * np.pi * x[1]) + a[2] * x[2] * (H - x[2] + 0.3)
    orc.set_state(U)
    orc.project_state()
    U = orc.state()
    F = orc.rhs(U)
    lhs = _integral(orc, F)[0]
    # 1/2 of the wall integral of m.n, by Gauss quadrature of the modal state on z = 0 (n = -z) and z = H (n = +z)
    b = orc.basis()
    zg, wg = b["zg"], b["wg"]
    from numpy.polynomial import legendre
    L = p + 1
    phi = lambda l, t: legendre.legval(t, np.eye(L)[l]) * np.sqrt((2 * l + 1) / 2.0)
    Uh = orc.modal()
    J, M = orc.geometry()
    rhs = 0.0
    for e in range(m.ne):
        zc = m.vert[e, 2].mean()
        for side, tz, sgn in ((0, -1.0, -1.0), (5, 1.0, 1.0)):
            on = (side == 0 and abs(m.vert[e, 2].min()) < 1e-12) or (side == 5 and abs(m.vert[e, 2].max() - H) < 1e-12)
            if not on:
                continue
            c = Uh[3, e].reshape(L, L, L)                   # rho w, modes (a, b, c)
            wz = np.array([phi(k, tz) for k in range(L)])
            face = np.einsum("abc,c->ab", c, wz)
            Pa = np.array([[phi(a, g) for g in zg] for a in range(L)])
            vals = Pa.T @ face @ Pa                          # at the Gauss points
            hx = m.vert[e, 0].max() - m.vert[e, 0].min()
            hy = m.vert[e, 1].max() - m.vert[e, 1].min()
            rhs += 0.5 * sgn * np.einsum("i,j,ij->", wg, wg, vals) * hx * hy / 4.0
    assert abs(rhs) > 1e-4
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(rhs)), (lhs, rhs)


def test_state_boundary_of_a_uniform_flow_matches_periodic():
    for bc in (1,):
        m = meshgen.channel_hexes(2, 2, 2, (0, 0, 0), (1.0, 1.0, 1.0), bc_bottom=bc, bc_top=bc)
        per = meshgen.periodic_hexes(2, 2, 2, (0, 0, 0), (1.0, 1.0, 1.0))
        a, b = Oracle(m, 3, 8, gamma=GAMMA), Oracle(per, 3, 8, gamma=GAMMA)
        U = _uniform(a.coords(), 0.9, 0.2, 0.1, 0.3, 0.8)
        a.set_bc_state(U[:, 0, 0])
        assert np.max(np.abs(a.rhs(U) - b.rhs(U))) < 1e-12
