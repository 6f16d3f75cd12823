"""Pins of the oracle's face fluxes by the properties that define them (no GPU).

The reference's regression fixtures all use the default Lax-Friedrichs flux
(src/Euler.cpp:98), and the reference has no 3D Roe flux at all, so the Roe
fluxes are pinned here:

* consistency: F(U, U, n) = F(U).n;
* conservation: F(UL, UR, n) = -F(UR, UL, -n).  For the Roe flux, written as
  F(UL) + A-(UR - UL), this holds exactly when the Roe matrix satisfies
  F(UR) - F(UL) = A~ (UR - UL), so it pins the averages, wave strengths and
  eigenvectors;
* upwinding: with every wave speed positive the Roe flux is F(UL), with every
  wave speed negative it is F(UR) (the Roe property again, wave by wave);
* the 3D fluxes reduce to the 2D ones for w = 0, n_z = 0;
* the 3D fluxes are rotation invariant (momentum and normal rotated together).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle_lib import riemann

G = 1.4


def _state(rng, nsd, mach=0.5):
    rho = rng.uniform(0.5, 2.0)
    u = rng.normal(size=nsd) * mach
    p = rng.uniform(0.5, 2.0)
    E = p / (G - 1) + 0.5 * rho * np.dot(u, u)
    return np.concatenate([[rho], rho * u, [E]])


def _phys(U, n):
    nsd = len(n)
    rho, m, E = U[0], U[1:1 + nsd], U[-1]
    u = m / rho
    p = (G - 1) * (E - 0.5 * np.dot(m, u))
    un = np.dot(u, n)
    return np.concatenate([[rho * un], m * un + p * n, [(E + p) * un]])


def _unit(rng, nsd):
    n = rng.normal(size=nsd)
    return n / np.linalg.norm(n)


@pytest.mark.parametrize("nsd", [2, 3])
@pytest.mark.parametrize("flux", [0, 1])
def test_consistency_and_conservation(nsd, flux):
    rng = np.random.default_rng(10 * nsd + flux)
    for _ in range(50):
        UL, UR, n = _state(rng, nsd), _state(rng, nsd), _unit(rng, nsd)
        assert np.allclose(riemann(nsd, flux, UL, UL, n), _phys(UL, n), rtol=1e-13, atol=1e-13)
        a = riemann(nsd, flux, UL, UR, n)
        b = riemann(nsd, flux, UR, UL, -n)
        assert np.allclose(a, -b, rtol=1e-12, atol=1e-12), (a, b)


@pytest.mark.parametrize("nsd", [2, 3])
def test_roe_upwinds_supersonic_states(nsd):
    rng = np.random.default_rng(7 + nsd)
    for _ in range(50):
        n = _unit(rng, nsd)
        UL, UR = _state(rng, nsd, 0.1), _state(rng, nsd, 0.1)
        # add a normal velocity of Mach ~3 to both states
        for U in (UL, UR):
            rho = U[0]
            p = (G - 1) * (U[-1] - 0.5 * np.dot(U[1:-1], U[1:-1]) / rho)
            c = np.sqrt(G * p / rho)
            du = 3.0 * 2.5 * c * n                      # c <= sqrt(1.4*2/0.5)
            U[-1] += np.dot(U[1:-1], du) + 0.5 * rho * np.dot(du, du)
            U[1:-1] += rho * du
        assert np.allclose(riemann(nsd, 1, UL, UR, n), _phys(UL, n), rtol=1e-12, atol=1e-12)
        assert np.allclose(riemann(nsd, 1, UL, UR, -n), _phys(UR, -n), rtol=1e-12, atol=1e-12)
        # Lax-Friedrichs is not upwind: it must differ (the pin is not vacuous)
        assert not np.allclose(riemann(nsd, 0, UL, UR, n), _phys(UL, n), rtol=1e-6)


@pytest.mark.parametrize("flux", [0, 1])
def test_3d_reduces_to_2d(flux):
    rng = np.random.default_rng(3 + flux)
    for _ in range(50):
        UL2, UR2, n2 = _state(rng, 2), _state(rng, 2), _unit(rng, 2)
        UL3 = np.array([UL2[0], UL2[1], UL2[2], 0.0, UL2[3]])
        UR3 = np.array([UR2[0], UR2[1], UR2[2], 0.0, UR2[3]])
        F2 = riemann(2, flux, UL2, UR2, n2)
        F3 = riemann(3, flux, UL3, UR3, np.array([n2[0], n2[1], 0.0]))
        assert np.allclose(F3[[0, 1, 2, 4]], F2, rtol=1e-12, atol=1e-13)
        assert abs(F3[3]) < 1e-13


@pytest.mark.parametrize("flux", [0, 1])
def test_3d_rotation_invariance(flux):
    rng = np.random.default_rng(5 + flux)
    for _ in range(50):
        UL, UR, n = _state(rng, 3), _state(rng, 3), _unit(rng, 3)
        Rm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        if np.linalg.det(Rm) < 0:
            Rm[:, 0] *= -1
        rot = lambda U: np.concatenate([[U[0]], Rm @ U[1:4], [U[4]]])
        a = rot(riemann(3, flux, UL, UR, n))
        b = riemann(3, flux, rot(UL), rot(UR), Rm @ n)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("flux", [0, 1])
def test_face_assembly_on_anisotropic_hexes(flux):
    """The reference's golden fixtures all use 2 x 2 (x 2) unit elements,
    where every metric and every face Jacobian ratio J_side/J_vol is 1.  This
    pins the scaling on strongly anisotropic hexes: with a constant state per
    element the volume term vanishes, and each element's residual integral is
    the sum over its faces of area x F(U_e, U_neighbour, n) (the element's own
    flux integrates to zero over the closed box)."""
    from dgm_amd import mesh as meshgen
    from oracle_lib import Oracle
    nx, ny, nz = 3, 2, 2
    hi = (1.5, 0.4, 0.1)
    h = (hi[0] / nx, hi[1] / ny, hi[2] / nz)
    m = meshgen.periodic_hexes(nx, ny, nz, (0.0, 0.0, 0.0), hi)
    o = Oracle(m, 2, 8, flux=flux)
    rng = np.random.default_rng(31)
    states = [_state(rng, 3) for _ in range(m.ne)]
    U = np.zeros((5, m.ne, o.qtot))
    for e in range(m.ne):
        U[:, e, :] = states[e][:, None]
    F = o.rhs(U)
    J, _ = o.geometry()
    w = o.basis()["w"]
    W = np.einsum("i,j,k->ijk", w, w, w).ravel()
    area = {0: h[0] * h[1], 5: h[0] * h[1], 1: h[0] * h[2], 3: h[0] * h[2], 2: h[1] * h[2], 4: h[1] * h[2]}
    normal = {0: (0, 0, -1.0), 5: (0, 0, 1.0), 1: (0, -1.0, 0), 3: (0, 1.0, 0), 4: (-1.0, 0, 0), 2: (1.0, 0, 0)}
    nbr = {(int(r[0]), int(r[1])): int(r[2]) for r in m.con}
    for e in range(m.ne):
        got = np.einsum("vq,q->v", F[:, e, :], W) * J[e]
        want = sum(area[s] * riemann(3, flux, states[e], states[nbr[(e, s)]], np.array(normal[s]))
                   for s in range(6))
        assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max()), (e, got, want)
