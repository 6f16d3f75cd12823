"""Pins of the CPU oracle by what the mathematics fixes (no GPU).

The golden restart files pin the Euler/NS residual and RK steppers as a whole
on the reference's own runs.  These tests pin the pieces they leave open,
each chosen so that a plausible mistake fails one of them:

* GLL / Gauss rules and the collocation derivative: exactness degrees
  (src/Polylib.cpp zwgll/zwgl/Dgll).
* Normalized Legendre modes (src/Basis.cpp:97-102): orthonormal under the
  q = p+2 GLL rule, so the modal projection is idempotent.
* Residual F = -dU/dt (src/Domain.cpp:1201-1230):
  - free-stream preservation: a constant state has F = 0 to round-off;
  - conservation: on a periodic box the domain integral of F vanishes for
    every conserved field (a one-sided or non-conservative face flux fails);
  - closed forms against smooth exact solutions, spectrally accurate at
    p = 7: an entropy wave (convective fluxes, Euler 2D/3D), a steady shear
    layer (BR1 viscous stress and viscous work, src/Navier_Stokes3d.cpp) and
    a density stratification at constant pressure (heat flux with
    kappa = mu gamma / Pr on the internal energy gradient).
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import ic, mesh as meshgen
from oracle_lib import Oracle

GAMMA = 1.4


def _integral(orc, F):
    """Domain integral of nodal fields F[v][e][q] with the GLL rule (affine elements)."""
    b = orc.basis()
    w = b["w"]
    if orc.nsd == 2:
        W = np.einsum("i,j->ij", w, w).ravel()
    else:
        W = np.einsum("i,j,k->ijk", w, w, w).ravel()
    J, _ = orc.geometry()
    return np.einsum("veq,q,e->v", F, W, J)


@pytest.mark.parametrize("p", [2, 5, 9])
def test_quadrature_and_derivative_exactness(p):
    m = meshgen.periodic_quads(2, 2)
    b = Oracle(m, p, 1).basis()
    q, qf = p + 2, p + 1
    z, w, zg, wg, D, B = b["z"], b["w"], b["zg"], b["wg"], b["D"], b["B"]
    assert z[0] == -1.0 and z[-1] == 1.0
    for k in range(2 * q - 2):                       # GLL exact to degree 2q-3
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, z ** k) - exact) < 1e-14 * max(1, q)
    k = 2 * q - 2                                    # ...and not beyond
    assert abs(np.dot(w, z ** k) - 2.0 / (k + 1)) > 1e-6
    for k in range(2 * qf):                          # Gauss exact to degree 2qf-1
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(wg, zg ** k) - exact) < 1e-14 * max(1, q)
    for k in range(q):                               # D exact on degree <= q-1
        d = D @ z ** k
        ex = k * z ** (k - 1) if k else np.zeros(q)
        assert np.max(np.abs(d - ex)) < 1e-11 * max(1, k * k)
    M = (B * w) @ B.T                                # orthonormal modes
    assert np.max(np.abs(M - np.eye(p + 1))) < 1e-13


def test_projection_idempotent():
    m = meshgen.periodic_hexes(2, 2, 2, hi=(2.0, 1.0, 1.5))
    orc = Oracle(m, 4, 8)
    x = orc.coords()
    rng = np.random.default_rng(3)
    U = 1.0 + 0.1 * rng.standard_normal((5,) + x.shape[1:])
    orc.set_state(U)
    orc.project_state()
    P1 = orc.state()
    orc.project_state()
    P2 = orc.state()
    assert np.max(np.abs(P2 - P1)) < 1e-13
    assert np.max(np.abs(P1 - U)) > 1e-3            # the projection did something


@pytest.mark.parametrize("eqn,nsd", [(1, 2), (2, 2), (8, 3), (9, 3)])
def test_free_stream_preserved(eqn, nsd):
    vis = 0.01 if eqn in (2, 9) else 0.0
    if nsd == 2:
        m = meshgen.periodic_quads(3, 2, 0.0, 1.0, 0.0, 0.7)
        U0 = [1.2, 0.3, -0.2, 2.5]
    else:
        m = meshgen.periodic_hexes(3, 2, 2, hi=(1.0, 0.7, 0.9))
        U0 = [1.2, 0.3, -0.2, 0.1, 2.5]
    orc = Oracle(m, 4, eqn, vis=vis)
    x = orc.coords()
    U = np.stack([np.full(x.shape[1:], u) for u in U0])
    F = orc.rhs(U)
    assert np.max(np.abs(F)) < 1e-12


@pytest.mark.parametrize("eqn,flux", [(1, 0), (1, 1), (2, 0), (8, 0), (8, 1), (9, 0), (9, 1)])
def test_conservation_periodic(eqn, flux):
    vis = 0.02 if eqn in (2, 9) else 0.0
    rng = np.random.default_rng(11)
    if eqn in (1, 2):
        m = meshgen.periodic_quads(3, 3, 0.0, 2 * np.pi, 0.0, 2 * np.pi)
        orc = Oracle(m, 5, eqn, vis=vis, flux=flux)
        x = orc.coords()
        pert = [ic.fourier_modes(x, (2 * np.pi,) * 2, 2, 0.05, rng) for _ in range(4)]
        U = np.stack([1.0 + pert[0], 0.2 + pert[1], -0.1 + pert[2], 2.5 + pert[3]])
    else:
        m = meshgen.periodic_hexes(3, 2, 2, hi=(2 * np.pi,) * 3)
        orc = Oracle(m, 4, eqn, vis=vis, flux=flux)
        x = orc.coords()
        pert = [ic.fourier_modes(x, (2 * np.pi,) * 3, 2, 0.05, rng) for _ in range(5)]
        U = np.stack([1.0 + pert[0], 0.2 + pert[1], -0.1 + pert[2], 0.15 + pert[3], 2.5 + pert[4]])
    F = orc.rhs(U)
    tot = _integral(orc, F)
    scale = _integral(orc, np.abs(F))
    assert np.all(scale > 1e-3)
    assert np.max(np.abs(tot) / scale) < 1e-12


def _entropy_wave(x, U0, gamma=GAMMA):
    rho = 1.0 + 0.2 * np.sin(x)
    drho = 0.2 * np.cos(x)
    p = 1.0 / gamma
    E = p / (gamma - 1) + 0.5 * rho * U0 ** 2
    return rho, drho, E


@pytest.mark.parametrize("nsd", [2, 3])
def test_entropy_wave_closed_form(nsd):
    """rho = 1 + 0.2 sin x, u = (U0, 0, 0), p const: F = rho'(x) (U0, U0^2, 0, 0, U0^3/2)."""
    U0, p = 0.5, 7
    if nsd == 2:
        m = meshgen.periodic_quads(6, 2, 0.0, 2 * np.pi, 0.0, 1.0)
        orc = Oracle(m, p, 1)
    else:
        m = meshgen.periodic_hexes(6, 2, 2, hi=(2 * np.pi, 1.0, 1.0))
        orc = Oracle(m, p, 8)
    x = orc.coords()
    rho, drho, E = _entropy_wave(x[0], U0)
    zero = np.zeros_like(rho)
    mom = [rho * U0, zero] + ([zero] if nsd == 3 else [])
    U = np.stack([rho] + mom + [E])
    F = orc.rhs(U)
    ex = [U0 * drho, U0 ** 2 * drho, zero] + ([zero] if nsd == 3 else []) + [0.5 * U0 ** 3 * drho]
    ex = np.stack(ex)
    err = np.max(np.abs(F - ex)) / np.max(np.abs(ex))
    assert err < 1e-6, err


def test_shear_layer_closed_form():
    """rho = 1, u = (A sin y, 0, 0), p const, Navier-Stokes 3D:
    F = (0, mu A sin y, 0, 0, -mu A^2 cos 2y) (viscous stress and work only)."""
    A, mu, p = 0.3, 0.05, 7
    m = meshgen.periodic_hexes(2, 6, 2, hi=(1.0, 2 * np.pi, 1.0))
    orc = Oracle(m, p, 9, vis=mu, Pr=0.72)
    x = orc.coords()
    y = x[1]
    one, zero = np.ones_like(y), np.zeros_like(y)
    u = A * np.sin(y)
    E = (1.0 / GAMMA) / (GAMMA - 1) + 0.5 * u * u
    U = np.stack([one, u, zero, zero, E])
    F = orc.rhs(U)
    ex = np.stack([zero, mu * A * np.sin(y), zero, zero, -mu * A * A * np.cos(2 * y)])
    err = np.max(np.abs(F - ex)) / np.max(np.abs(ex))
    assert err < 1e-6, err


def _heat_flux_error(nz, p, mu=0.05, Pr=0.72):
    m = meshgen.periodic_hexes(2, 2, nz, hi=(1.0, 1.0, 2 * np.pi))
    orc = Oracle(m, p, 9, vis=mu, Pr=Pr)
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


def test_heat_flux_closed_form():
    """u = 0, p = p0, rho = 1 + 0.2 sin z: only conduction acts,
    F_E = -kappa e''(z), e = p0 / ((gamma-1) rho), kappa = mu gamma / Pr.
    e is not a trigonometric polynomial, so the error is checked to converge
    at the spectral rate (h^8 at p = 7 gives a factor ~40 per halving)."""
    coarse, fine = _heat_flux_error(6, 7), _heat_flux_error(12, 7)
    assert fine < 1e-6, fine
    assert coarse / fine > 16.0, (coarse, fine)
