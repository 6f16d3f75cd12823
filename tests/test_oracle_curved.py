"""Curved hexes in the oracle: nodal metrics (Hex::compute_curved_geometry,
src/Hex.cpp:3992-4080), side geometry from them, and the projection through
the element mass matrix B^T diag(wJ) B (src/Hex.cpp:289-318).

The reference's curved fixtures (runs/cyl, scyl, cvquad) are 2D and need its
curved-side files, so the 3D curved path is pinned by:
* affine nodes fed through the curved path reproduce the affine residual,
  which the golden fixtures pin (the mass matrix is then J I to rounding);
* the projection of any modal field is exact: forward(backward(u)) = u, on
  strongly curved elements (a wrong mass matrix or solve breaks it);
* consistency on a curved mesh: the residual of a smooth state converges
  spectrally to the exact flux divergence, under p-refinement; for
  Navier-Stokes (BR1) to the exact divergence of the inviscid minus the
  viscous flux, derived symbolically.
"""
from __future__ import annotations

import numpy as np

from dgm_amd import mesh as meshgen
from oracle_lib import Oracle

GAMMA = 1.4


def _state(x, y, z):
    rho = 1.0 + 0.1 * np.sin(x) * np.cos(y) * np.cos(z)
    u = 0.3 * np.sin(x) * np.cos(y)
    v = -0.3 * np.cos(x) * np.sin(y) * np.cos(z)
    w = 0.2 * np.sin(z) * np.cos(x)
    p = 1.0 + 0.1 * np.cos(2 * x) * np.cos(y + z)
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E])


def _flux_div(x, y, z, h=1e-5):
    """Exact div F(U) of the Euler flux by central differences of the analytic state (error ~1e-10)."""
    def F(x, y, z):
        U = _state(x, y, z)
        r, mx, my, mz, E = U
        u, v, w = mx / r, my / r, mz / r
        p = (GAMMA - 1) * (E - 0.5 * (mx * u + my * v + mz * w))
        fx = np.stack([mx, mx * u + p, my * u, mz * u, (E + p) * u])
        fy = np.stack([my, mx * v, my * v + p, mz * v, (E + p) * v])
        fz = np.stack([mz, mx * w, my * w, mz * w + p, (E + p) * w])
        return fx, fy, fz
    dx = (F(x + h, y, z)[0] - F(x - h, y, z)[0]) / (2 * h)
    dy = (F(x, y + h, z)[1] - F(x, y - h, z)[1]) / (2 * h)
    dz = (F(x, y, z + h)[2] - F(x, y, z - h)[2]) / (2 * h)
    return dx + dy + dz


def wavy_map(L, amp):
    """Periodic-preserving interior deformation of [0, L]^3 (boundary faces stay flat)."""
    def fn(x, y, z):
        s = lambda t: np.sin(2 * np.pi * t / L)
        return (x + amp * s(y) * s(z), y + amp * s(z) * s(x), z + amp * s(x) * s(y))
    return fn


def test_affine_nodes_reproduce_affine_path():
    m = meshgen.channel_hexes(2, 2, 2, (0, 0, 0), (2 * np.pi, 2 * np.pi, 2.0), bc_bottom=0, bc_top=1)
    bc = [1.0, 0.1, 0.0, 0.0, 2.6]
    for flux in (0, 1):
        a, b = Oracle(m, 4, 8, flux=flux), Oracle(m, 4, 8, flux=flux)
        a.set_bc_state(bc); b.set_bc_state(bc)
        b.set_nodes(meshgen.hex_nodes(m, 4))
        x = a.coords()
        U = _state(x[0], x[1], x[2])
        Fa, Fb = a.rhs(U), b.rhs(U)
        assert np.max(np.abs(Fa - Fb)) <= 1e-12 * np.max(np.abs(Fa)), np.max(np.abs(Fa - Fb))


def test_projection_exact_on_curved_elements():
    m = meshgen.channel_hexes(2, 2, 2, (0, 0, 0), (1.0, 1.0, 1.0))
    p = 4
    o = Oracle(m, p, 8)
    o.set_nodes(meshgen.mapped_nodes(m, p, meshgen.bump_map((0, 0, 0), (1, 1, 1), 0.3)))
    rng = np.random.default_rng(1)
    Uh = rng.standard_normal((5, m.ne, (p + 1) ** 3))
    o.set_modal(Uh)
    back = o.modal()
    assert np.max(np.abs(back - Uh)) < 1e-12, np.max(np.abs(back - Uh))


def _consistency_error(p, n=2, amp=0.15):
    L = 2 * np.pi
    m = meshgen.periodic_hexes(n, n, n, (0, 0, 0), (L, L, L))
    o = Oracle(m, p, 8)
    xyz = meshgen.mapped_nodes(m, p, wavy_map(L, amp))
    o.set_nodes(xyz)
    U = _state(*xyz)
    F = o.rhs(U)
    o.set_state(_flux_div(*xyz))                        # the exact divergence, projected alike
    o.project_state()
    ex = o.state()
    return np.sqrt(np.sum((F - ex) ** 2)) / np.sqrt(np.sum(ex ** 2))


def test_curved_residual_converges_spectrally():
    errs = [_consistency_error(p, n=4) for p in (3, 5, 7)]
    assert errs[2] < 2e-5, errs
    assert errs[0] / errs[1] > 10 and errs[1] / errs[2] > 10, errs
    # the curvature adds little to the resolution error of the affine mesh;
    # a wrong normal, metric or mass matrix would leave an O(1) inconsistency
    flat = _consistency_error(5, n=4, amp=0.0)
    assert errs[1] < 1.5 * flat, (errs[1], flat)


# ---- Navier-Stokes 3D (BR1) on curved hexes (src/Navier_Stokes3d.cpp:113-258) ----

PR = 0.72


def test_ns3d_affine_nodes_reproduce_affine_path():
    """Isothermal bottom wall and an imposed state on top (the viscous ghosts
    of Navier_Stokes3d's own BC database), affine nodes through the curved
    path: the affine Navier-Stokes residual is pinned by runs/chan."""
    m = meshgen.channel_hexes(2, 2, 2, (0, 0, 0), (2 * np.pi, 2 * np.pi, 2.0), bc_bottom=3, bc_top=1)
    bc = [1.0, 0.1, 0.0, 0.0, 2.6]
    for p in (3, 4):
        a, b = Oracle(m, p, 9, vis=0.05), Oracle(m, p, 9, vis=0.05)
        for o in (a, b):
            o.set_bc_state(bc)
            o.set_isothermal(1.0, 2.5)
        b.set_nodes(meshgen.hex_nodes(m, p))
        x = a.coords()
        U = _state(x[0], x[1], x[2])
        Fa, Fb = a.rhs(U), b.rhs(U)
        assert np.max(np.abs(Fa - Fb)) <= 1e-12 * np.max(np.abs(Fa)), (p, np.max(np.abs(Fa - Fb)))


def _ns_flux_div(mu):
    """Exact div (F_c - F_v) of the state _state, symbolically: the stresses
    mu (grad u + grad u^T - 2/3 div u I) and the energy flux tau.u +
    (gamma / Pr) mu grad e with e = E / rho - |u|^2 / 2."""
    import sympy as sp
    x, y, z = sp.symbols("x y z")
    X = (x, y, z)
    rho = 1 + sp.Rational(1, 10) * sp.sin(x) * sp.cos(y) * sp.cos(z)
    vel = (sp.Rational(3, 10) * sp.sin(x) * sp.cos(y), -sp.Rational(3, 10) * sp.cos(x) * sp.sin(y) * sp.cos(z),
           sp.Rational(2, 10) * sp.sin(z) * sp.cos(x))
    p = 1 + sp.Rational(1, 10) * sp.cos(2 * x) * sp.cos(y + z)
    k2 = sum(c * c for c in vel)
    E = p / (GAMMA - 1) + rho * k2 / 2
    e = E / rho - k2 / 2
    div_u = sum(sp.diff(vel[i], X[i]) for i in range(3))
    tau = [[mu * (sp.diff(vel[i], X[j]) + sp.diff(vel[j], X[i])) - (sp.Rational(2, 3) * mu * div_u if i == j else 0)
            for j in range(3)] for i in range(3)]
    Fc = ([[rho * vel[j] for j in range(3)]] +
          [[rho * vel[i] * vel[j] + (p if i == j else 0) for j in range(3)] for i in range(3)] +
          [[(E + p) * vel[j] for j in range(3)]])
    Fv = ([[0, 0, 0]] + [list(tau[i]) for i in range(3)] +
          [[sum(tau[j][k] * vel[k] for k in range(3)) + GAMMA / PR * mu * sp.diff(e, X[j]) for j in range(3)]])
    div = [sum(sp.diff(Fc[f][j] - Fv[f][j], X[j]) for j in range(3)) for f in range(5)]
    return sp.lambdify(X, div, "numpy")


def _ns_consistency_error(p, fn, mu, n=4, amp=0.15):
    L = 2 * np.pi
    m = meshgen.periodic_hexes(n, n, n, (0, 0, 0), (L, L, L))
    o = Oracle(m, p, 9, vis=mu, Pr=PR)
    xyz = meshgen.mapped_nodes(m, p, wavy_map(L, amp))
    o.set_nodes(xyz)
    F = o.rhs(_state(*xyz))
    o.set_state(np.stack([np.broadcast_to(c, xyz[0].shape) for c in fn(*xyz)]))
    o.project_state()
    ex = o.state()
    return np.sqrt(np.sum((F - ex) ** 2)) / np.sqrt(np.sum(ex ** 2))


def test_ns3d_curved_residual_converges_spectrally():
    """The Navier-Stokes residual on curved hexes converges spectrally to
    the exact div (F_c - F_v) (5.9e-3, 4.5e-4, 2.5e-5 at p = 3, 5, 7 with
    mu = 0.1, where dropping the viscous flux leaves 0.24).  A wrong metric,
    normal, lift or mass matrix in the viscous path would leave an O(1)
    inconsistency that does not converge."""
    mu = 0.1
    fn = _ns_flux_div(mu)
    errs = [_ns_consistency_error(p, fn, mu) for p in (3, 5, 7)]
    assert errs[2] < 5e-5, errs
    assert errs[0] / errs[1] > 10 and errs[1] / errs[2] > 10, errs
    flat = _ns_consistency_error(5, fn, mu, amp=0.0)
    assert errs[1] < 2.5 * flat, (errs[1], flat)
    assert _ns_consistency_error(5, _ns_flux_div(0.0), mu) > 0.1   # the viscous terms matter here


# ---- curved quads (Quad::compute_curved_geometry, src/Quad.cpp:1309-1452) ----

def _state2(x, y):
    rho = 1.0 + 0.1 * np.sin(x) * np.cos(y)
    u = 0.3 * np.sin(x) * np.cos(y)
    v = -0.3 * np.cos(x) * np.sin(y)
    p = 1.0 + 0.1 * np.cos(2 * x) * np.cos(y)
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E])


def _flux_div2(x, y, h=1e-5):
    def F(x, y):
        r, mx, my, E = _state2(x, y)
        u, v = mx / r, my / r
        p = (GAMMA - 1) * (E - 0.5 * (mx * u + my * v))
        return np.stack([mx, mx * u + p, my * u, (E + p) * u]), np.stack([my, mx * v, my * v + p, (E + p) * v])
    return (F(x + h, y)[0] - F(x - h, y)[0]) / (2 * h) + (F(x, y + h)[1] - F(x, y - h)[1]) / (2 * h)


def wavy_map2(L, amp):
    def fn(x, y):
        s = lambda t: np.sin(2 * np.pi * t / L)
        return x + amp * s(y) * s(x), y + amp * s(x) * s(y) * 0.7
    return fn


def test_quads_affine_nodes_reproduce_affine_path():
    m = meshgen.periodic_quads(3, 2, 0.0, 2 * np.pi, 0.0, 1.5)
    for eqn, kw in ((1, dict(flux=0)), (1, dict(flux=1)), (2, dict(vis=0.05)), (30, {})):
        a, b = Oracle(m, 5, eqn, **kw), Oracle(m, 5, eqn, **kw)
        if eqn == 30:
            a.set_balance(False); b.set_balance(False)
        b.set_nodes(meshgen.quad_nodes(m, 5))
        x = a.coords()
        U = _state2(x[0], x[1])
        if eqn == 30:
            U = np.stack([1.0 + 0.1 * np.sin(x[0]), 0.2 * np.cos(x[1]), 0.1 + 0 * x[0]])
        Fa, Fb = a.rhs(U), b.rhs(U)
        assert np.max(np.abs(Fa - Fb)) <= 1e-12 * np.max(np.abs(Fa)), (eqn, np.max(np.abs(Fa - Fb)))


def test_quads_projection_exact_and_free_stream():
    m = meshgen.periodic_quads(3, 3, 0.0, 1.0, 0.0, 1.0)
    p = 6
    o = Oracle(m, p, 1)
    o.set_nodes(meshgen.mapped_nodes(m, p, wavy_map2(1.0, 0.08)))
    rng = np.random.default_rng(4)
    Uh = rng.standard_normal((4, m.ne, (p + 1) ** 2))
    o.set_modal(Uh)
    assert np.max(np.abs(o.modal() - Uh)) < 1e-12
    x = o.coords()
    U = np.stack([np.full_like(x[0], v) for v in (1.2, 0.3, -0.2, 2.9)])
    assert np.max(np.abs(o.rhs(U))) < 1e-12


def _consistency_error2(p, n=4, amp=0.15):
    L = 2 * np.pi
    m = meshgen.periodic_quads(n, n, 0.0, L, 0.0, L)
    o = Oracle(m, p, 1)
    xy = meshgen.mapped_nodes(m, p, wavy_map2(L, amp))
    o.set_nodes(xy)
    F = o.rhs(_state2(*xy))
    o.set_state(_flux_div2(*xy))
    o.project_state()
    ex = o.state()
    return np.sqrt(np.sum((F - ex) ** 2)) / np.sqrt(np.sum(ex ** 2))


def test_quads_curved_residual_converges_spectrally():
    errs = [_consistency_error2(p) for p in (3, 5, 7)]
    assert errs[2] < 1e-4, errs
    assert errs[0] / errs[1] > 10 and errs[1] / errs[2] > 10, errs
    flat = _consistency_error2(5, amp=0.0)
    assert errs[1] < 1.5 * flat, (errs[1], flat)
