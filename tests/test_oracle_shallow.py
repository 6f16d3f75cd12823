"""Shallow water in the oracle (Shallow::time_derivative, src/Shallow.cpp:251-525;
Shallow::WallBC, src/ShallowBC.hpp:45-62; FuncBC/FuncSource, src/BC.cpp:95-200,
src/Source.cpp:11-75) in (phi, uphi, vphi), phi the geopotential.

The reference's one shallow-water fixture (runs/shallow_mms) is on a mesh of
non-affine quads, which this oracle does not represent, so the path is pinned
by what the equations fix:
* the face flux against a closed form: with piecewise-constant states the
  residual's element integrals are sums of Lax-Friedrichs fluxes written
  independently here (a wrong wave speed, a dropped pressure term or a wrong
  normal fails it);
* conservation (periodic) and the wall's mass budget (the wall ghost keeps phi
  and the tangential momentum, so its mass flux is m.n / 2);
* uniform flow is steady, with walls parallel to it and with state and
  function boundaries holding it;
* consistency: for a manufactured solution with function boundaries and the
  matching source, the residual converges spectrally to -dU/dt;
* stage times: a uniform state forced by a source polynomial in t is
  integrated exactly by RK4 and LSRK14 (cubic) and TVD-RK3 (quadratic);
* the balance term: the first residual is subtracted from every later one.
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import mesh as meshgen
from oracle_lib import SHALLOW, Oracle
from test_oracle_invariants import _integral

TWO_PI = 2 * np.pi


def _fn(U, nx, ny):
    """Physical normal flux of (phi, m_x, m_y): [m.n, m_x u_n + phi^2/2 n_x, m_y u_n + phi^2/2 n_y]."""
    phi, mx, my = U
    mn = mx * nx + my * ny
    return np.array([mn, mx * mn / phi + 0.5 * phi * phi * nx, my * mn / phi + 0.5 * phi * phi * ny])


def _lf(a, b, nx, ny):
    lam = max(abs(a[1] * nx + a[2] * ny) / a[0] + np.sqrt(a[0]), abs(b[1] * nx + b[2] * ny) / b[0] + np.sqrt(b[0]))
    return 0.5 * (_fn(a, nx, ny) + _fn(b, nx, ny) + lam * (np.asarray(a) - np.asarray(b)))


def _smooth(x, y, t=0.0):
    phi = 1.0 + 0.1 * np.sin(x + 0.3 * t) * np.cos(y) + 0.05 * np.cos(2 * y - t)
    u = 0.2 * np.cos(y) + 0.1 * np.sin(x) * np.cos(t)
    v = 0.1 * np.sin(x + y) - 0.05 * t
    return np.stack([phi, phi * u, phi * v])


def _flux_div(x, y, t, h=1e-5):
    """div F(U) of the manufactured state, by central differences (error ~1e-10)."""
    def F(x, y):
        phi, mx, my = _smooth(x, y, t)
        p = 0.5 * phi * phi
        fx = np.stack([mx, mx * mx / phi + p, my * mx / phi])
        fy = np.stack([my, mx * my / phi, my * my / phi + p])
        return fx, fy
    return (F(x + h, y)[0] - F(x - h, y)[0]) / (2 * h) + (F(x, y + h)[1] - F(x, y - h)[1]) / (2 * h)


def _dudt(x, y, t, h=1e-5):
    return (_smooth(x, y, t + h) - _smooth(x, y, t - h)) / (2 * h)


@pytest.mark.parametrize("axis", [0, 1])
def test_face_flux_closed_form(axis):
    """Three elements in a periodic row, constant states: each element's
    residual integral is h * [Fh(self, right, +n) + Fh(self, left, -n)]."""
    n3 = (3, 1) if axis == 0 else (1, 3)
    m = meshgen.periodic_quads(*n3, 0.0, 1.5 if axis == 0 else 0.7, 0.0, 0.7 if axis == 0 else 1.5)
    o = Oracle(m, 2, SHALLOW)
    o.set_balance(False)
    states = [np.array([1.3, 0.4, -0.2]), np.array([0.8, -0.1, 0.3]), np.array([1.1, 0.25, 0.15])]
    U = np.zeros((3, 3, o.qtot))
    for e in range(3):
        U[:, e, :] = states[e][:, None]
    F = o.rhs(U)
    J, _ = o.geometry()
    b = o.basis()
    W = np.einsum("i,j->ij", b["w"], b["w"]).ravel()
    nx, ny = (1.0, 0.0) if axis == 0 else (0.0, 1.0)
    h = 0.7                                                  # length of the faces normal to the row
    for e in range(3):
        got = np.einsum("vq,q->v", F[:, e, :], W) * J[e]
        me, right, left = states[e], states[(e + 1) % 3], states[(e - 1) % 3]
        want = h * (_lf(me, right, nx, ny) + _lf(me, left, -nx, -ny))
        assert np.allclose(got, want, rtol=1e-13, atol=1e-13), (e, got, want)


def test_conservation_periodic():
    m = meshgen.periodic_quads(3, 4, 0.0, TWO_PI, 0.0, TWO_PI)
    o = Oracle(m, 5, SHALLOW)
    o.set_balance(False)
    x = o.coords()
    F = o.rhs(_smooth(x[0], x[1]))
    tot = _integral(o, F)
    scale = _integral(o, np.abs(F))
    assert np.all(np.abs(tot) <= 1e-13 * scale), (tot, scale)


def test_uniform_flow_is_steady():
    U0 = np.array([1.2, 0.3, 0.0])
    cases = [((0, None, 0, None), None),      # walls parallel to the flow
             ((1, 1, 1, 1), U0),              # state boundaries holding it
             ((2, 2, 2, 2), U0)]              # function boundaries holding it
    for bc, held in cases:
        m = meshgen.box_quads(3, 2, 0.0, 1.0, 0.0, 0.8, bc=bc)
        o = Oracle(m, 4, SHALLOW)
        o.set_balance(False)
        if held is not None:
            o.set_bc_state(held)
            o.set_functions(bc=lambda t, x, y: np.outer(held, np.ones_like(x)))
        x = o.coords()
        U = np.broadcast_to(U0[:, None, None], (3,) + x[0].shape).copy()
        F = o.rhs(U)
        assert np.max(np.abs(F)) < 1e-13, (bc, np.max(np.abs(F)))


def test_wall_mass_budget():
    p, H = 4, 0.9
    m = meshgen.box_quads(3, 3, 0.0, 1.0, 0.0, H, bc=(0, None, 0, None))
    o = Oracle(m, p, SHALLOW)
    o.set_balance(False)
    x = o.coords()
    U = np.stack([1.0 + 0.1 * np.sin(TWO_PI * x[0]) * np.cos(x[1]),
                  0.2 + 0.05 * np.cos(TWO_PI * x[0]) + 0.0 * x[1],
                  0.1 * np.sin(TWO_PI * x[0]) + 0.3 * x[1] * (H - x[1]) + 0.1 * x[1] + 0.02])
    o.set_state(U)
    o.project_state()
    U = o.state()
    F = o.rhs(U)
    lhs = _integral(o, F)[0]
    # 1/2 of the wall integral of m.n by GLL quadrature of the (degree-p) state on the wall nodes
    q = p + 2
    w = o.basis()["w"]
    rhs = 0.0
    for e in range(m.ne):
        Ue = U[2, e].reshape(q, q)                          # m_y, node (i, j) = (x index, y index)
        hx = m.vert[e, 0].max() - m.vert[e, 0].min()
        if abs(m.vert[e, 1].min()) < 1e-12:
            rhs += 0.5 * (-1.0) * np.dot(w, Ue[:, 0]) * hx / 2
        if abs(m.vert[e, 1].max() - H) < 1e-12:
            rhs += 0.5 * (+1.0) * np.dot(w, Ue[:, q - 1]) * hx / 2
    assert abs(rhs) > 1e-3
    assert abs(lhs - rhs) <= 1e-13, (lhs, rhs)


def _mms_error(p, t0=0.3, n=6):
    m = meshgen.box_quads(n, n, 0.0, TWO_PI, 0.0, TWO_PI, bc=(2, 2, 2, 2))
    o = Oracle(m, p, SHALLOW)
    o.set_balance(False)
    o.set_functions(bc=lambda t, x, y: _smooth(x, y, t),
                    src=lambda t, x, y: _dudt(x, y, t) + _flux_div(x, y, t))
    o.set_time(t0)
    x = o.coords()
    F = o.rhs(_smooth(x[0], x[1], t0))
    o.set_state(-_dudt(x[0], x[1], t0))
    o.project_state()
    ex = o.state()
    return np.sqrt(np.sum((F - ex) ** 2) / np.sum(ex ** 2))


def test_manufactured_solution_converges_spectrally():
    errs = [_mms_error(p) for p in (2, 4, 6)]
    assert errs[2] < 2e-5, errs
    assert errs[0] / errs[1] > 20 and errs[1] / errs[2] > 20, errs


@pytest.mark.parametrize("inttype,deg", [(4, 3), (14, 3), (3, 2)])
def test_stage_times_integrate_polynomial_forcing(inttype, deg):
    """Uniform state, source s(t) = -(c t^deg) on phi only: dphi/dt = c t^deg
    exactly, which RK4 / LSRK14 (order 4) and TVD-RK3 (order 3) integrate
    without error when each stage sees its own time."""
    c = 0.7
    m = meshgen.periodic_quads(2, 2)
    o = Oracle(m, 2, SHALLOW)
    o.set_balance(False)

    def src(t, x, y):
        out = np.zeros((3, len(x)))
        out[0] = c * t ** deg
        return out
    o.set_functions(src=src)
    x = o.coords()
    U0 = np.stack([np.full_like(x[0], 1.0), np.full_like(x[0], 0.1), np.full_like(x[0], -0.2)])
    o.set_state(U0)
    t0, dt, ns = 0.2, 0.05, 6
    o.set_time(t0)
    o.advance(inttype, dt, ns)
    T = t0 + ns * dt
    assert abs(o.time() - T) < 1e-14
    want = 1.0 + c * (T ** (deg + 1) - t0 ** (deg + 1)) / (deg + 1)
    got = o.state()
    assert np.max(np.abs(got[0] - want)) < 1e-13, (np.max(np.abs(got[0] - want)))
    assert np.max(np.abs(got[1:] - U0[1:])) < 1e-13


def test_balance_subtracts_the_first_residual():
    m = meshgen.periodic_quads(3, 3, 0.0, TWO_PI, 0.0, TWO_PI)
    raw, bal = Oracle(m, 3, SHALLOW), Oracle(m, 3, SHALLOW)
    raw.set_balance(False)
    x = raw.coords()
    U1, U2 = _smooth(x[0], x[1], 0.0), _smooth(x[0], x[1], 0.4)
    assert np.max(np.abs(bal.rhs(U1))) == 0.0
    want = raw.rhs(U2) - raw.rhs(U1)
    got = bal.rhs(U2)
    assert np.max(np.abs(got - want)) < 1e-14 * np.max(np.abs(want)) + 1e-15
    bal.set_balance(True)                                   # re-arming stores the next residual
    assert np.max(np.abs(bal.rhs(U2))) == 0.0
