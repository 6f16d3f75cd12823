"""BR2 viscous face gradient of the CPU oracle (3D Navier-Stokes).

The reference's viscous discretisation is BR1 only, so BR2 has no golden
vector; its parity against the reference is unpinned.  These checks pin the
oracle's BR2 by what the scheme fixes:

* continuous primitive traces make every face lifting vanish, so BR2 and BR1
  give the same residual;
* the momentum rows of the viscous operator are linear in the velocity at
  rho = 1; BR2's bilinear form is symmetric, so that operator is symmetric in
  the modal inner product (orthonormal modes, equal element Jacobians), and
  dissipative: <u, F_visc(u)> > 0 with F = -dU/dt;
* its quadratic form is exactly affine in the penalty eta with a positive
  slope (the penalty adds eta * sum_f |r_f|^2-type terms and nothing else);
* it is consistent: the conduction-only closed form converges at the spectral
  rate, as for BR1.

A plausible slip fails one of them: a wrong-side or wrong-sign face lifting
breaks symmetry or the sign of the slope; a dropped broken-gradient term breaks
consistency; lifting all faces instead of one makes BR2 depend on eta through
other faces, which the BR1 comparison at eta = 1 (where a full lifting would
reproduce BR1) rules out.
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import mesh as meshgen
from oracle_lib import Oracle

GAMMA = 1.4


def _modal(orc, F):
    """Modal coefficients [v][e][nm] of nodal fields F (via the oracle's forward transform)."""
    orc.set_state(F)
    return orc.modal()


def _visc_op(ns, eu, Uh):
    """Momentum rows of the viscous residual (NS minus Euler) for a modal state."""
    ns.set_modal(Uh)
    U = ns.state()
    F = ns.rhs(U) - eu.rhs(U)
    return _modal(ns, F)[1:4]


def _state_from_velocity(orc, vel, E0=20.0):
    ne, nm = orc.ne, orc.nm
    Uh = np.zeros((5, ne, nm))
    Uh[0, :, 0] = 1.0 / np.sqrt(0.125)       # rho = 1: mode 000 = 1 / phi_000
    Uh[4, :, 0] = E0 / np.sqrt(0.125)
    Uh[1:4] = vel
    return Uh


def _pair(p=2, n=2, mu=0.05, eta=6.0, viscous=1):
    m = meshgen.periodic_hexes(n, n, n, hi=(1.0, 1.0, 1.0))
    ns = Oracle(m, p, 9, gamma=GAMMA, vis=mu, Pr=0.72, viscous=viscous, br2_eta=eta)
    eu = Oracle(m, p, 8, gamma=GAMMA)
    return m, ns, eu


def test_br2_equals_br1_for_continuous_traces():
    """rho = 1, p = const, u a periodic continuous piecewise-linear hat in x
    (exact at p >= 2, continuous across faces): no jumps, BR2 == BR1."""
    m = meshgen.periodic_hexes(4, 2, 2, hi=(1.0, 1.0, 1.0))
    p = 3
    br1 = Oracle(m, p, 9, vis=0.05)
    br2 = Oracle(m, p, 9, vis=0.05, viscous=1, br2_eta=6.0)
    x = br1.coords()
    hat = 1.0 - np.abs(((x[0] * 4.0) % 2.0) - 1.0)       # kinks at element faces only
    zero = np.zeros_like(hat)
    p0 = 1.0
    U = np.stack([np.ones_like(hat), 0.3 * hat, zero, 0.1 * hat,
                  p0 / (GAMMA - 1) + 0.5 * (0.09 + 0.01) * hat ** 2])
    F1, F2 = br1.rhs(U), br2.rhs(U)
    scale = np.max(np.abs(F1[1:]))
    assert scale > 1e-3
    assert np.max(np.abs(F1 - F2)) <= 1e-11 * scale


@pytest.mark.parametrize("eta", [1.0, 6.0, 20.0])
def test_br2_viscous_operator_symmetric_dissipative(eta):
    m, ns, eu = _pair(eta=eta)
    rng = np.random.default_rng(11)
    a = rng.standard_normal((3, ns.ne, ns.nm)) * 0.1
    b = rng.standard_normal((3, ns.ne, ns.nm)) * 0.1
    Aa = _visc_op(ns, eu, _state_from_velocity(ns, a))
    Ab = _visc_op(ns, eu, _state_from_velocity(ns, b))
    ab, ba = np.sum(a * Ab), np.sum(b * Aa)
    assert abs(ab - ba) <= 1e-11 * (abs(ab) + abs(ba) + np.sum(a * Aa)), (ab, ba)
    assert np.sum(a * Aa) > 0 and np.sum(b * Ab) > 0
    # linear: A(a + b) = A a + A b
    Aab = _visc_op(ns, eu, _state_from_velocity(ns, a + b))
    assert np.max(np.abs(Aab - Aa - Ab)) <= 1e-10 * np.max(np.abs(Aa))


def test_br2_penalty_affine_with_positive_slope():
    rng = np.random.default_rng(5)
    a = None
    quad = []
    for eta in (1.0, 6.0, 11.0):
        m, ns, eu = _pair(eta=eta)
        if a is None:
            a = rng.standard_normal((3, ns.ne, ns.nm)) * 0.1
        quad.append(np.sum(a * _visc_op(ns, eu, _state_from_velocity(ns, a))))
    d1, d2 = quad[1] - quad[0], quad[2] - quad[1]
    assert d1 > 0
    assert abs(d1 - d2) <= 1e-9 * abs(quad[2])
    # and BR2 differs from BR1 on this discontinuous field
    m, br1, eu = _pair(viscous=0)
    q1 = np.sum(a * _visc_op(br1, eu, _state_from_velocity(br1, a)))
    assert abs(q1 - quad[1]) > 1e-6 * abs(q1)


def _heat_flux_error(nz, p, mu=0.05, Pr=0.72):
    m = meshgen.periodic_hexes(2, 2, nz, hi=(1.0, 1.0, 2 * np.pi))
    orc = Oracle(m, p, 9, vis=mu, Pr=Pr, viscous=1, br2_eta=6.0)
    z = orc.coords()[2]
    p0 = 1.0 / GAMMA
    rho = 1.0 + 0.2 * np.sin(z)
    d1, d2 = 0.2 * np.cos(z), -0.2 * np.sin(z)
    zero = np.zeros_like(z)
    U = np.stack([rho, zero, zero, zero, np.full_like(z, p0 / (GAMMA - 1))])
    F = orc.rhs(U)
    c = p0 / (GAMMA - 1)
    e2 = -c * (d2 / rho ** 2 - 2 * d1 ** 2 / rho ** 3)
    kap = mu * GAMMA / Pr
    ex = np.stack([zero, zero, zero, zero, -kap * e2])
    return np.max(np.abs(F - ex)) / np.max(np.abs(ex))


def test_br2_heat_flux_closed_form():
    """As the BR1 closed form: u = 0, p = p0, rho = 1 + 0.2 sin z; only
    conduction acts and the BR2 residual converges spectrally to -kappa e''."""
    coarse, fine = _heat_flux_error(6, 7), _heat_flux_error(12, 7)
    assert fine < 1e-6, fine
    assert coarse / fine > 16.0, (coarse, fine)
