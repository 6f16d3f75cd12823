"""Acoustic inverse-source problem and its discrete adjoint (ctypes binding of
the dgx_acoustic_* entry points of libdgx.so).

Linearized 2D Euler equations about a uniform mean flow (the reference's
LinEuler_quasi, eqntype 4, src/LinEuler_quasi.cpp:151-495, with a uniform
mean), driven by a source theta_e g(t) in the energy equation, one value per
element, g a Ricker wavelet.  The pressure perturbation is observed at
receiver points after every LSRK14 step; J = 1/2 dt sum (p' - d)^2.
``gradient`` returns J and dJ/dtheta from the discrete adjoint, with forward
states recomputed from binomial checkpoints (cf. DTK::Griewank,
src/DTK_Griewank.hpp:42-125, and LSRK14::adjoint_advance,
src/TimeInt.cpp:2064-2092).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import dgx

DGX_LINEULER = 4


class AcousticDesc(ctypes.Structure):
    _fields_ = [("rho0", ctypes.c_double), ("u0", ctypes.c_double), ("v0", ctypes.c_double),
                ("p0", ctypes.c_double), ("nrec", ctypes.c_int), ("xrec", ctypes.POINTER(ctypes.c_double)),
                ("f0", ctypes.c_double), ("t0", ctypes.c_double), ("dt", ctypes.c_double),
                ("nsteps", ctypes.c_int), ("nsnap", ctypes.c_int)]


def _bind(lib):
    if getattr(lib, "_acoustic_bound", False):
        return lib
    h = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ll = ctypes.POINTER(ctypes.c_longlong)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "dgx_acoustic_create": ([ctypes.POINTER(dgx.DomainDesc), ctypes.POINTER(AcousticDesc), h,
                                 ctypes.POINTER(h)], ctypes.c_int),
        "dgx_acoustic_destroy": ([h], ctypes.c_int),
        "dgx_acoustic_forward": ([h, dp, dp], ctypes.c_int),
        "dgx_acoustic_gradient": ([h, dp, dp, dp, dp], ctypes.c_int),
        "dgx_acoustic_stats": ([h, ll, ll, ip], ctypes.c_int),
        "dgx_revolve_count": ([ctypes.c_int, ctypes.c_int, ll, ip, ip], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    lib._acoustic_bound = True
    return lib


def _check(lib, rc):
    if rc != 0:
        raise dgx.DgxError(rc, lib.dgx_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def revolve_count(nsteps, nsnap):
    """(forward steps advanced, most slots held, reversal order correct) of the
    checkpointing schedule, simulated on the host."""
    lib = _bind(dgx.load())
    adv, mx, ok = ctypes.c_longlong(), ctypes.c_int(), ctypes.c_int()
    _check(lib, lib.dgx_revolve_count(int(nsteps), int(nsnap), ctypes.byref(adv), ctypes.byref(mx),
                                      ctypes.byref(ok)))
    return int(adv.value), int(mx.value), bool(ok.value)


class AcousticProblem:
    """One inverse-source problem on the GPU (FP64)."""

    def __init__(self, mesh, p, mean, receivers, f0, t0, dt, nsteps, nsnap=0, gamma=1.4, stream=None,
                 precision=dgx.DGX_FP64):
        self.lib = _bind(dgx.load())
        self.ne, self.nsteps = int(mesh.ne), int(nsteps)
        self._vert = np.ascontiguousarray(mesh.vert, dtype=np.float64)
        self._con = np.ascontiguousarray(mesh.con, dtype=np.int32)
        self._rec = np.ascontiguousarray(np.asarray(receivers, dtype=np.float64).reshape(-1, 2))
        self.nrec = len(self._rec)
        md = dgx.DomainDesc(2, self.ne, int(p), DGX_LINEULER, 0, int(precision), float(gamma), 0.0, 0.72,
                            _dp(self._vert), len(self._con), self._con.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                            0, 0.0)
        rho0, u0, v0, p0 = (float(v) for v in mean)
        ad = AcousticDesc(rho0, u0, v0, p0, self.nrec, _dp(self._rec), float(f0), float(t0), float(dt),
                          self.nsteps, int(nsnap))
        self.h = ctypes.c_void_p()
        _check(self.lib, self.lib.dgx_acoustic_create(ctypes.byref(md), ctypes.byref(ad),
                                                      ctypes.c_void_p(stream or 0), ctypes.byref(self.h)))
        self.dt = float(dt)
        self.control_shape = (self.ne,)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.dgx_acoustic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _theta(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        if th.shape != (self.ne,):
            raise ValueError(f"theta shape {th.shape} != {(self.ne,)}")
        return th

    def forward(self, theta):
        """Observations y[nsteps][nrec]."""
        th = self._theta(theta)
        y = np.zeros((self.nsteps, self.nrec))
        _check(self.lib, self.lib.dgx_acoustic_forward(self.h, _dp(th), _dp(y)))
        return y

    def objective(self, theta, data):
        y = self.forward(theta)
        return 0.5 * self.dt * float(np.sum((y - data) ** 2))

    def gradient(self, theta, data):
        """(J, dJ/dtheta) by the checkpointed discrete adjoint."""
        th = self._theta(theta)
        d = np.ascontiguousarray(data, dtype=np.float64)
        if d.shape != (self.nsteps, self.nrec):
            raise ValueError(f"data shape {d.shape} != {(self.nsteps, self.nrec)}")
        J = np.zeros(1)
        g = np.zeros(self.ne)
        _check(self.lib, self.lib.dgx_acoustic_gradient(self.h, _dp(th), _dp(d), _dp(J), _dp(g)))
        return float(J[0]), g

    def stats(self):
        f, a, s = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_int()
        _check(self.lib, self.lib.dgx_acoustic_stats(self.h, ctypes.byref(f), ctypes.byref(a), ctypes.byref(s)))
        return {"forward_steps": int(f.value), "adjoint_steps": int(a.value), "snapshots": int(s.value)}


class ReoOptProblem:
    """The optimisation seam of pydgm (``ReoOptProblem.computeState``,
    ``computeAdjoint``, ``getGradient``, ``getControl``; python/dgm.pyx:322-355,
    Reo::OptProblem) over an adjoint problem: the acoustic inverse-source
    problem (control: the source field theta) or the nonlinear Euler problem
    of dgm_amd.euler_adjoint (control: the initial modal state).  The state is
    a forward solve, the objective J = 1/2 dt sum (p - d)^2 against `data`, the
    gradient the checkpointed discrete adjoint's.  computeState returns J for
    the current control; computeAdjoint evaluates dJ/dcontrol, which
    getGradient then returns."""

    def __init__(self, problem, data, theta0=None):
        self.problem = problem
        self.data = np.ascontiguousarray(data, dtype=np.float64)
        shape = problem.control_shape
        self.control = np.zeros(shape) if theta0 is None else np.array(theta0, dtype=np.float64)
        self.gradient = np.zeros(shape)
        self.J = None

    def getControl(self) -> np.ndarray:
        return self.control

    def computeState(self) -> float:
        self.J = self.problem.objective(self.control, self.data)
        return self.J

    def computeAdjoint(self) -> None:
        self.J, self.gradient = self.problem.gradient(self.control, self.data)

    def getGradient(self) -> np.ndarray:
        return self.gradient

    def getObjective(self) -> float:
        return self.J

    def solve(self, iterations: int = 10, step: float = 1.0) -> float:
        """Steepest descent with a backtracking line search (the role of
        OptProblem::solve); returns the final objective."""
        J = self.computeState()
        for _ in range(iterations):
            self.computeAdjoint()
            g = self.gradient
            a = step
            while a > 1e-12:
                trial = self.control - a * g
                Jt = self.problem.objective(trial, self.data)
                if Jt < J:
                    self.control, J = trial, Jt
                    break
                a *= 0.5
            else:
                break
        self.J = J
        return J
