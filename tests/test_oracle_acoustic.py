"""The oracle's linear-system residual (2D linearized Euler about a uniform
mean flow) and the test-side discrete adjoint built on it.

* Linearization: the linear residual L U' is the derivative of the oracle's
  nonlinear Euler residual (pinned by the reference's golden vectors) at the
  uniform state, checked against central differences of that residual
  (O(eps^2) = 1e-8 at eps = 1e-4; measured 1e-10).
* Transpose: <V, L U> = <L^T V, U> in the modal Euclidean inner product, with
  L^T built from the adjoint system (-Ax^T, -Ay^T) and the element Jacobian
  scaling, on a mesh of unequal elements.
* Gradient: the discrete adjoint's dJ/dtheta against central differences of
  J along seeded directions.  J is quadratic in theta, so central differences
  are exact up to rounding.
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import mesh as meshgen
from oracle_lib import Oracle
from acoustic_ref import AcousticRef, euler_jacobians

GAMMA = 1.4
MEAN = (1.0, 0.3, -0.1, 1.0 / GAMMA)


def stretched_quads(nx, ny, L=1.0):
    """Periodic quads on [0,L]^2 with unequal column widths and row heights
    (still rectangles, so affine)."""
    m = meshgen.periodic_quads(nx, ny, 0.0, L, 0.0, L)
    f = lambda s: s + 0.06 * L * np.sin(2 * np.pi * s / L)
    m.vert[:, 0, :] = f(m.vert[:, 0, :])
    m.vert[:, 1, :] = f(m.vert[:, 1, :])
    return m


def _smooth_fields(x, rng, amp=1.0):
    out = []
    for v in range(4):
        a, b, c = rng.uniform(-1, 1, 3)
        out.append(amp * (a * np.sin(2 * np.pi * x[0] + c) + b * np.cos(2 * np.pi * x[1] - c)
                          + 0.3 * np.sin(2 * np.pi * (x[0] + x[1]))))
    return np.stack(out)


def test_linear_residual_is_the_euler_linearization():
    m = stretched_quads(4, 3)
    p = 4
    rho0, u0, v0, p0 = MEAN
    Ax, Ay = euler_jacobians(rho0, u0, v0, p0, GAMMA)
    lin = Oracle(m, p, 4, gamma=GAMMA)
    lin.set_linear(Ax, Ay, u0, v0, np.sqrt(GAMMA * p0 / rho0))
    eul = Oracle(m, p, 1, gamma=GAMMA)
    x = lin.coords()
    rng = np.random.default_rng(3)
    Up = _smooth_fields(x, rng, 0.1)
    U0 = np.stack([np.full_like(x[0], rho0), np.full_like(x[0], rho0 * u0), np.full_like(x[0], rho0 * v0),
                   np.full_like(x[0], p0 / (GAMMA - 1) + 0.5 * rho0 * (u0 * u0 + v0 * v0))])
    eps = 1e-4
    Fd = (eul.rhs(U0 + eps * Up) - eul.rhs(U0 - eps * Up)) / (2 * eps)
    Fl = lin.rhs(Up)
    rel = np.linalg.norm(Fl - Fd) / np.linalg.norm(Fd)
    assert rel < 1e-8, rel
    # and not a coincidence: a wrong Jacobian entry is far off
    Ax2 = Ax.copy()
    Ax2[1, 0] += 0.05
    lin.set_linear(Ax2, Ay, u0, v0, np.sqrt(GAMMA * p0 / rho0))
    assert np.linalg.norm(lin.rhs(Up) - Fd) / np.linalg.norm(Fd) > 1e-3


def test_transpose_identity_unequal_elements():
    m = stretched_quads(5, 4)
    ref = AcousticRef(m, 3, MEAN, [(0.31, 0.47)], f0=2.0, t0=0.5, dt=1e-3, nsteps=1, gamma=GAMMA)
    J, _ = ref.fwd.geometry()
    assert np.ptp(J) > 1e-3 * J.mean()                 # the element Jacobians differ
    rng = np.random.default_rng(8)
    U = rng.standard_normal((4, ref.ne, ref.nm))
    V = rng.standard_normal((4, ref.ne, ref.nm))
    lhs = np.sum(V * ref.apply_L(U))
    rhs = np.sum(ref.apply_LT(V) * U)
    assert abs(lhs - rhs) <= 1e-12 * (np.linalg.norm(V) * np.linalg.norm(ref.apply_L(U))), (lhs, rhs)


def _small_problem(nsteps=12):
    m = stretched_quads(4, 4)
    rec = [(0.13, 0.21), (0.77, 0.4), (0.5, 0.9)]
    return AcousticRef(m, 2, MEAN, rec, f0=3.0, t0=0.02, dt=2e-3, nsteps=nsteps, gamma=GAMMA)


def test_adjoint_gradient_matches_central_differences():
    ref = _small_problem()
    rng = np.random.default_rng(21)
    theta_true = rng.standard_normal(ref.ne)
    data = ref.forward(theta_true)
    theta = rng.standard_normal(ref.ne)
    Jv, g = ref.gradient(theta, data)
    assert Jv > 0
    assert abs(Jv - ref.objective(theta, data)) <= 1e-13 * Jv
    for k in range(3):
        d = rng.standard_normal(ref.ne)
        h = 1e-2
        fd = (ref.objective(theta + h * d, data) - ref.objective(theta - h * d, data)) / (2 * h)
        assert abs(fd - g @ d) <= 1e-9 * (abs(fd) + np.linalg.norm(g) * np.linalg.norm(d)), (fd, g @ d)
