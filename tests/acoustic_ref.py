"""Test-side reference for the acoustic inverse-source problem (linearized 2D
Euler about a uniform mean flow) and its discrete adjoint, built on the CPU
oracle's linear residual.  Test infrastructure only.

Problem (all in the modal coordinates of the state, [field][elem][mode]):
  forward   LSRK14 from U = 0 with F(U) = L U - g(t) b(theta), where L is the
            DG residual of the linear system and b(theta) puts the element
            constant theta_e into the energy equation (mode 0 coefficient
            2 theta_e for orthonormal modes on [-1,1]^2);
  observe   after every step, the pressure perturbation
            p' = (gamma-1) (E' - u0 m' - v0 n' + |u0|^2 / 2 rho') at each receiver;
  objective J = 1/2 dt sum_steps sum_receivers (p' - d)^2;
  adjoint   the exact reverse of those recurrences (LSRK14 2N-storage stages
            of src/TimeInt.cpp:1989-2013), with L^T applied as J_e times the
            residual of the adjoint system (-Ax^T, -Ay^T, same speed) on V/J_e.
"""
from __future__ import annotations

import numpy as np

from oracle_lib import Oracle

LSRK_A = [0.0, -0.7188012108672410, -0.7785331173421570, -0.0053282796654044, -0.8552979934029281,
          -3.9564138245774565, -1.5780575380587385, -2.0837094552574054, -0.7483334182761610,
          -0.7032861106563359, 0.0013917096117681, -0.0932075369637460, -0.9514200470875948,
          -7.1151571693922548]
LSRK_B = [0.0367762454319673, 0.3136296607553959, 0.1531848691869027, 0.0030097086818182, 0.3326293790646110,
          0.2440251405350864, 0.3718879239592277, 0.6204126221582444, 0.1524043173028741, 0.0760894927419266,
          0.0077604214040978, 0.0024647284755382, 0.0780348340049386, 5.5059777270269628]
LSRK_C = [0.0, 0.0367762454319673, 0.1249685262725025, 0.2446177702277698, 0.2476149531070420,
          0.2969311120382472, 0.3978149645802642, 0.5270854589440328, 0.6981269994175695, 0.8190890835352128,
          0.8527059887098624, 0.8604711817462826, 0.8627060376969976, 0.8734213127600976]


def euler_jacobians(rho0, u0, v0, p0, gamma=1.4):
    """dF_x/dU, dF_y/dU of the 2D Euler flux at a uniform state (conservative U)."""
    g1 = gamma - 1.0
    q2 = u0 * u0 + v0 * v0
    E0 = p0 / g1 + 0.5 * rho0 * q2
    H = (E0 + p0) / rho0
    Ax = np.array([[0.0, 1.0, 0.0, 0.0],
                   [0.5 * g1 * q2 - u0 * u0, (3.0 - gamma) * u0, -g1 * v0, g1],
                   [-u0 * v0, v0, u0, 0.0],
                   [u0 * (0.5 * g1 * q2 - H), H - g1 * u0 * u0, -g1 * u0 * v0, gamma * u0]])
    Ay = np.array([[0.0, 0.0, 1.0, 0.0],
                   [-u0 * v0, v0, u0, 0.0],
                   [0.5 * g1 * q2 - v0 * v0, -g1 * u0, (3.0 - gamma) * v0, g1],
                   [v0 * (0.5 * g1 * q2 - H), -g1 * u0 * v0, H - g1 * v0 * v0, gamma * v0]])
    return Ax, Ay


def ricker(t, f0, t0):
    a = (np.pi * f0 * (t - t0)) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


def normleg(l, x):
    from numpy.polynomial import legendre
    c = np.zeros(l + 1)
    c[l] = 1.0
    return legendre.legval(x, c) * np.sqrt((2 * l + 1) / 2.0)


class AcousticRef:
    """Forward solve, objective and discrete-adjoint gradient on the oracle."""

    def __init__(self, mesh, p, mean, receivers, f0, t0, dt, nsteps, gamma=1.4):
        self.mesh, self.p, self.gamma = mesh, p, gamma
        rho0, u0, v0, p0 = mean
        self.mean = mean
        self.c0 = np.sqrt(gamma * p0 / rho0)
        Ax, Ay = euler_jacobians(rho0, u0, v0, p0, gamma)
        self.fwd = Oracle(mesh, p, 4, gamma=gamma)
        self.fwd.set_linear(Ax, Ay, u0, v0, self.c0)
        self.adj = Oracle(mesh, p, 4, gamma=gamma)
        self.adj.set_linear(-Ax.T, -Ay.T, u0, v0, self.c0)
        self.f0, self.t0, self.dt, self.nsteps = f0, t0, dt, nsteps
        self.ne, self.nm = mesh.ne, (p + 1) ** 2
        J, M = self.fwd.geometry()
        self.Jac = J                                  # element Jacobians
        self.obs = self._receivers(np.asarray(receivers, dtype=np.float64), M)
        g1 = gamma - 1.0
        self.cobs = g1 * np.array([0.5 * (u0 * u0 + v0 * v0), -u0, -v0, 1.0])

    def _receivers(self, xr, M):
        """(element, modal weights) of each receiver point (affine quads)."""
        out = []
        vert = self.mesh.vert
        for x in xr:
            for e in range(self.ne):
                c = vert[e].mean(axis=1)
                rx, ry, sx, sy = M[e, 0], M[e, 1], M[e, 3], M[e, 4]
                xi = rx * (x[0] - c[0]) + ry * (x[1] - c[1])
                eta = sx * (x[0] - c[0]) + sy * (x[1] - c[1])
                if abs(xi) <= 1 + 1e-12 and abs(eta) <= 1 + 1e-12:
                    L = self.p + 1
                    w = np.array([normleg(l, xi) * normleg(m, eta) for l in range(L) for m in range(L)])
                    out.append((e, w))
                    break
            else:
                raise ValueError(f"receiver {x} outside the mesh")
        return out

    def apply_L(self, U):
        self.fwd.set_modal(U)
        return self._modal_rhs(self.fwd)

    def apply_LT(self, V):
        """L^T V (Euclidean, modal): J_e * adjoint-system residual of V / J_e."""
        self.adj.set_modal(V / self.Jac[None, :, None])
        return self._modal_rhs(self.adj) * self.Jac[None, :, None]

    @staticmethod
    def _modal_rhs(orc):
        U = orc.state()
        F = orc.rhs(U)
        orc.set_state(F)
        return orc.modal()

    def source(self, theta):
        b = np.zeros((4, self.ne, self.nm))
        b[3, :, 0] = 2.0 * np.asarray(theta)
        return b

    def observe(self, U):
        return np.array([self.cobs @ U[:, e, :] @ w for e, w in self.obs])

    def forward(self, theta, keep=False):
        """Observations y[n][r] after each step (and the states if keep)."""
        U = np.zeros((4, self.ne, self.nm))
        Ut = np.zeros_like(U)
        b = self.source(theta)
        ys, states = [], [U.copy()]
        for n in range(self.nsteps):
            tn = n * self.dt
            for s in range(14):
                R = self.apply_L(U) - ricker(tn + LSRK_C[s] * self.dt, self.f0, self.t0) * b
                Ut = LSRK_A[s] * Ut - self.dt * R
                U = U + LSRK_B[s] * Ut
            ys.append(self.observe(U))
            if keep:
                states.append(U.copy())
        return (np.array(ys), states) if keep else np.array(ys)

    def objective(self, theta, data):
        y = self.forward(theta)
        return 0.5 * self.dt * np.sum((y - data) ** 2)

    def gradient(self, theta, data):
        y, states = self.forward(theta, keep=True)
        Jv = 0.5 * self.dt * np.sum((y - data) ** 2)
        Ub = np.zeros((4, self.ne, self.nm))
        Utb = np.zeros_like(Ub)
        gth = np.zeros(self.ne)
        for n in reversed(range(self.nsteps)):
            r = y[n] - data[n]
            for (e, w), rr in zip(self.obs, r):
                Ub[:, e, :] += self.dt * rr * np.outer(self.cobs, w)
            tn = n * self.dt
            for s in reversed(range(14)):
                Utb = Utb + LSRK_B[s] * Ub
                Rb = -self.dt * Utb
                Ub = Ub + self.apply_LT(Rb)
                gth += -ricker(tn + LSRK_C[s] * self.dt, self.f0, self.t0) * 2.0 * Rb[3, :, 0]
                Utb = LSRK_A[s] * Utb
        return Jv, gth
