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
