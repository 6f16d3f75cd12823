"""The oracle's CFL time step (Euler::set_time_step, src/Euler.cpp:804-877;
Shallow::set_time_step, src/Shallow.cpp:573-637; metrics of
Element::compute_metrics_2d_tensor_prod, src/Element.cpp:485-555).

The reference holds no CFL fixture (the golden runs use a fixed dt), so the
restatement is pinned by what the formula reduces to:

* at rest on a square element of side h, the largest idt sits at a corner
  node, where both GLL spacings are delta = z_1 - z_0:
  idt = c (2/h) sqrt(2) / delta, c = sqrt(gamma p / rho);
* a uniform flow u along x on the same mesh adds u (2/h) / delta there;
* shallow water at rest: c = sqrt(phi);
* mesh scaling by s divides the CFL number by s;
* the step rule: with CFLmax the new step has CFL number CFLmax, and dtmax
  caps it.
A wrong spacing, a dropped term or a swapped metric fails one of these.
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import mesh as meshgen
from oracle_lib import Oracle

GAMMA = 1.4


def _gll(q):
    from numpy.polynomial import legendre
    c = np.zeros(q)
    c[-1] = 1.0
    return np.concatenate([[-1.0], np.sort(legendre.legroots(legendre.legder(c))), [1.0]])


def _uniform(o, rho, u, v, p):
    x = o.coords()
    one = np.ones_like(x[0])
    U = np.stack([rho * one, rho * u * one, rho * v * one, (p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)) * one])
    o.set_state(U)


@pytest.mark.parametrize("p", [2, 4, 7])
def test_rest_state_on_squares(p):
    h = 0.5
    m = meshgen.periodic_quads(4, 4, 0.0, 4 * h, 0.0, 4 * h)
    o = Oracle(m, p, 1, gamma=GAMMA)
    rho, pr = 1.3, 0.9
    _uniform(o, rho, 0.0, 0.0, pr)
    dt = 1e-3
    cfl, dtn = o.set_time_step(dt)
    z = _gll(p + 2)
    delta = z[1] - z[0]
    c = np.sqrt(GAMMA * pr / rho)
    want = dt * c * (2.0 / h) * np.sqrt(2.0) / delta
    assert abs(cfl - want) <= 1e-13 * want and dtn == dt


def test_uniform_flow_adds_the_convective_term():
    h, p = 0.25, 4
    m = meshgen.periodic_quads(3, 3, 0.0, 3 * h, 0.0, 3 * h)
    o = Oracle(m, p, 1, gamma=GAMMA)
    rho, u, pr = 1.0, 0.7, 1.0 / GAMMA
    _uniform(o, rho, u, 0.0, pr)
    cfl, _ = o.set_time_step(1.0)
    delta = _gll(p + 2)[1] - _gll(p + 2)[0]
    c = np.sqrt(GAMMA * pr / rho)
    want = (2.0 / h) * (u + c * np.sqrt(2.0)) / delta
    assert abs(cfl - want) <= 1e-13 * want


def test_shallow_rest_and_mesh_scaling():
    p = 3
    cfls = []
    for s in (1.0, 2.0):
        m = meshgen.periodic_quads(3, 2, 0.0, 3 * s, 0.0, 2 * s)
        o = Oracle(m, p, 30)
        x = o.coords()
        phi = 1.0 + 0.1 * np.sin(2 * np.pi * x[0] / (3 * s)) * np.cos(2 * np.pi * x[1] / (2 * s))
        o.set_state(np.stack([phi, 0.2 * phi, -0.1 * phi]))
        cfls.append(o.set_time_step(1e-2)[0])
    assert abs(cfls[0] / cfls[1] - 2.0) < 1e-13
    # at rest on unit squares: c = sqrt(phi)
    m = meshgen.periodic_quads(2, 2, 0.0, 2.0, 0.0, 2.0)
    o = Oracle(m, p, 30)
    x = o.coords()
    o.set_state(np.stack([np.full_like(x[0], 4.0), 0 * x[0], 0 * x[0]]))
    delta = _gll(p + 2)[1] - _gll(p + 2)[0]
    want = 1e-2 * 2.0 * 2.0 * np.sqrt(2.0) / delta
    assert abs(o.set_time_step(1e-2)[0] - want) <= 1e-13 * want


def test_step_rule():
    m = meshgen.periodic_quads(3, 3, 0.0, 1.0, 0.0, 1.0)
    o = Oracle(m, 4, 1, gamma=GAMMA)
    x = o.coords()
    rho = 1.0 + 0.1 * np.sin(2 * np.pi * x[0])
    u = 0.5 + 0.1 * np.cos(2 * np.pi * x[1])
    o.set_state(np.stack([rho, rho * u, 0.05 * rho, 1.0 / (GAMMA - 1) + 0.5 * rho * (u * u + 0.0025)]))
    c0, d0 = o.set_time_step(1e-3)
    c1, d1 = o.set_time_step(1e-3, CFLmax=0.2)
    assert d0 == 1e-3 and abs(c1 - 0.2) < 1e-14 and abs(d1 - 0.2 * 1e-3 / c0) < 1e-18
    c2, d2 = o.set_time_step(1e-3, CFLmax=0.2, dtmax=0.5 * d1)
    assert d2 == 0.5 * d1 and abs(c2 - 0.1) < 1e-14
    # hexes: Euler3d::set_time_step returns dt and keeps it
    h = Oracle(meshgen.periodic_hexes(2, 2, 2), 2, 8)
    assert h.set_time_step(3e-3, CFLmax=0.5) == (3e-3, 3e-3)
