"""Pin the oracle's Runge-Kutta steppers by mathematics (LSRK14::advance,
src/TimeInt.cpp:1870-2013; RungeKutta::advance :208-302; TVD::advance
:738-858).

* Linear operator: a small dense DG operator L (2D linearized Euler about a
  uniform mean flow, the derivative of the golden-pinned Euler residual) is
  assembled column by column from the oracle.  One oracle step of dU/dt = -L U
  is compared with exp(-dt L) U0 from an eigendecomposition; halving dt must
  shrink the one-step error by 2^(order+1) (32 for the 4th-order LSRK14 and
  RK4, 16 for TVD-RK3).
* Stability polynomial: the oracle's one-step map on that operator is a
  polynomial R(-dt L) whose Taylor coefficients must be 1, 1, 1/2, 1/6, 1/24
  (exp to 4th order) and whose degree-5 coefficient must differ from 1/120;
  this checks the 2N recurrences' A, B tables, not only quadrature conditions.
* Nonlinear order: on the nonlinear Euler residual the global error after a
  fixed time falls 16x per halving of dt (4th order), against a reference from
  RK4 at a much smaller step.

A dropped stage, a wrong sign in A or B, or a transposed register update
breaks the observed orders.
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import mesh as meshgen
from oracle_lib import Oracle
from acoustic_ref import euler_jacobians

GAMMA = 1.4
MEAN = (1.0, 0.3, -0.1, 1.0 / GAMMA)


def _linear_oracle(nx=2, ny=2, p=2):
    m = meshgen.periodic_quads(nx, ny, 0.0, 1.0, 0.0, 1.0)
    rho0, u0, v0, p0 = MEAN
    Ax, Ay = euler_jacobians(rho0, u0, v0, p0, GAMMA)
    o = Oracle(m, p, 4, gamma=GAMMA)
    o.set_linear(Ax, Ay, u0, v0, np.sqrt(GAMMA * p0 / rho0))
    return o


def _modal_rhs(o, Uh):
    o.set_modal(Uh)
    F = o.rhs(o.state())
    o.set_state(F)
    o.project_state()
    return o.modal()


def _dense_operator(o):
    shape = (o.nv, o.ne, o.nm)
    n = int(np.prod(shape))
    L = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        L[:, j] = _modal_rhs(o, e.reshape(shape)).ravel()
    return L, shape


def _expm_apply(L, dt, u0):
    lam, V = np.linalg.eig(-dt * L)
    return np.real(V @ (np.exp(lam) * np.linalg.solve(V, u0.astype(complex))))


def _step(o, shape, u0, inttype, dt, nsteps=1):
    o.set_modal(u0.reshape(shape))
    o.advance(inttype, dt, nsteps)
    return o.modal().ravel()


@pytest.fixture(scope="module")
def linear():
    o = _linear_oracle()
    L, shape = _dense_operator(o)
    rng = np.random.default_rng(11)
    u0 = rng.standard_normal(L.shape[0])
    return o, L, shape, u0


def test_dense_operator_is_linear_and_exp_is_consistent(linear):
    o, L, shape, u0 = linear
    # the assembled columns reproduce the oracle's residual of a random state
    assert np.allclose(L @ u0, _modal_rhs(o, u0.reshape(shape)).ravel(), rtol=0, atol=1e-11 * np.abs(L).max())
    # the eigendecomposition's exponential agrees with a Taylor series at small dt
    dt = 1e-3
    A = -dt * L
    t, term = u0.copy(), u0.copy()
    for k in range(1, 20):
        term = A @ term / k
        t = t + term
    assert np.allclose(_expm_apply(L, dt, u0), t, rtol=0, atol=1e-12 * np.abs(u0).max())


@pytest.mark.parametrize("inttype,order", [(14, 4), (4, 4), (3, 3)])
def test_one_step_error_order_on_linear_operator(linear, inttype, order):
    o, L, shape, u0 = linear
    lam_max = np.abs(np.linalg.eigvals(L)).max()
    errs = []
    for dt in (0.4 / lam_max, 0.2 / lam_max, 0.1 / lam_max):
        errs.append(np.linalg.norm(_step(o, shape, u0, inttype, dt) - _expm_apply(L, dt, u0)))
    ratios = [errs[0] / errs[1], errs[1] / errs[2]]
    want = 2.0 ** (order + 1)
    print(f"inttype {inttype}: one-step errors {errs}, ratios {ratios}")
    assert all(0.75 * want < r < 1.3 * want for r in ratios), (errs, ratios)


def test_lsrk14_stability_polynomial(linear):
    """Coefficients of R(z) with U1 = R(-dt L) U0, recovered from the oracle's
    step by fitting on the powers (-dt L)^k U0."""
    o, L, shape, u0 = linear
    lam_max = np.abs(np.linalg.eigvals(L)).max()
    dt = 0.3 / lam_max
    A = -dt * L
    K = [u0]
    for k in range(1, 15):
        K.append(A @ K[-1])
    Kmat = np.stack(K, axis=1)
    u1 = _step(o, shape, u0, 14, dt)
    coef, *_ = np.linalg.lstsq(Kmat, u1, rcond=None)
    want = np.array([1.0, 1.0, 0.5, 1.0 / 6.0, 1.0 / 24.0])
    print("LSRK14 R(z) coefficients", coef[:8])
    assert np.allclose(coef[:5], want, rtol=0, atol=1e-7), coef[:5]
    assert abs(coef[5] - 1.0 / 120.0) > 1e-6   # a 4th-order method, not the exponential


def test_lsrk14_global_order_on_nonlinear_euler():
    m = meshgen.periodic_quads(2, 2, 0.0, 2.0, 0.0, 2.0)
    p = 3
    o = Oracle(m, p, 1, gamma=GAMMA)
    x = o.coords()
    rho = 1.0 + 0.2 * np.sin(np.pi * x[0]) * np.cos(np.pi * x[1])
    u = 0.5 + 0.2 * np.cos(np.pi * x[1])
    v = 0.1 * np.sin(np.pi * x[0])
    pr = 1.0 + 0.1 * np.sin(np.pi * (x[0] + x[1]))
    U0 = np.stack([rho, rho * u, rho * v, pr / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)])
    o.set_state(U0)
    o.project_state()
    u0 = o.modal()
    T = 0.08

    def run(inttype, n):
        o.set_modal(u0)
        o.advance(inttype, T / n, n)
        return o.modal()

    ref = run(4, 512)                       # RK4, error ~ (1/64)^4 of the n = 8 runs
    errs = [np.linalg.norm(run(14, n) - ref) for n in (2, 4, 8)]
    ratios = [errs[0] / errs[1], errs[1] / errs[2]]
    print("LSRK14 nonlinear global errors", errs, ratios)
    assert all(11.0 < r < 22.0 for r in ratios), (errs, ratios)
