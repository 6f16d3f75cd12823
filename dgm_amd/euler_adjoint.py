"""Discrete adjoint of the nonlinear 2D Euler equations (ctypes binding of the
dgx_euler_adjoint_* entry points of libdgx.so).

LSRK14 steps of the 2D Euler DG residual (Euler::convection, the
Lax-Friedrichs flux, lift and projection; src/Euler.cpp:296-757,
src/Quad.cpp:1971-2038) on periodic parallelogram quads, from a modal initial
state U0 [4][ne][(p+1)^2].  The pressure is observed at receiver points after
every step; J = 1/2 dt sum (p - d)^2.  ``gradient`` returns J and dJ/dU0 from
the reversed LSRK14 recurrences (LSRK14::adjoint_advance,
src/TimeInt.cpp:2064-2092) with the hand-written transposed residual Jacobian
applied at every stage state, forward states recomputed from binomial
checkpoints (cf. DTK::Griewank, src/DTK_Griewank.hpp:42-125).  ``vjp`` is one
transposed Jacobian, (dR/dU(U))^T lam.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import dgx


class EulerAdjointDesc(ctypes.Structure):
    _fields_ = [("nrec", ctypes.c_int), ("xrec", ctypes.POINTER(ctypes.c_double)), ("dt", ctypes.c_double),
                ("nsteps", ctypes.c_int), ("nsnap", ctypes.c_int)]


def _bind(lib):
    if getattr(lib, "_euler_adjoint_bound", False):
        return lib
    h = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    ll = ctypes.POINTER(ctypes.c_longlong)
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "dgx_euler_adjoint_create": ([ctypes.POINTER(dgx.DomainDesc), ctypes.POINTER(EulerAdjointDesc), h,
                                      ctypes.POINTER(h)], ctypes.c_int),
        "dgx_euler_adjoint_destroy": ([h], ctypes.c_int),
        "dgx_euler_adjoint_forward": ([h, dp, dp, dp], ctypes.c_int),
        "dgx_euler_adjoint_gradient": ([h, dp, dp, dp, dp], ctypes.c_int),
        "dgx_euler_adjoint_vjp": ([h, dp, dp, dp], ctypes.c_int),
        "dgx_euler_adjoint_stats": ([h, ll, ll, ip], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    lib._euler_adjoint_bound = True
    return lib


def _check(lib, rc):
    if rc != 0:
        raise dgx.DgxError(rc, lib.dgx_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class EulerAdjointProblem:
    """Initial-state gradients of a pressure misfit through nonlinear Euler
    time stepping on the GPU (FP64)."""

    def __init__(self, mesh, p, receivers, dt, nsteps, nsnap=0, gamma=1.4, stream=None, flux=dgx.DGX_FLUX_LF):
        self.lib = _bind(dgx.load())
        self.ne, self.nsteps, self.p = int(mesh.ne), int(nsteps), int(p)
        self._vert = np.ascontiguousarray(mesh.vert, dtype=np.float64)
        self._con = np.ascontiguousarray(mesh.con, dtype=np.int32)
        self._rec = np.ascontiguousarray(np.asarray(receivers, dtype=np.float64).reshape(-1, 2))
        self.nrec = len(self._rec)
        md = dgx.DomainDesc(2, self.ne, self.p, dgx.DGX_EULER, int(flux), dgx.DGX_FP64, float(gamma), 0.0,
                            0.72, _dp(self._vert), len(self._con),
                            self._con.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 0, 0.0)
        ad = EulerAdjointDesc(self.nrec, _dp(self._rec), float(dt), self.nsteps, int(nsnap))
        self.h = ctypes.c_void_p()
        _check(self.lib, self.lib.dgx_euler_adjoint_create(ctypes.byref(md), ctypes.byref(ad),
                                                           ctypes.c_void_p(stream or 0), ctypes.byref(self.h)))
        self.dt = float(dt)
        self.shape = (4, self.ne, (self.p + 1) ** 2)
        self.control_shape = self.shape

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.dgx_euler_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _state(self, U, what="state"):
        u = np.ascontiguousarray(U, dtype=np.float64)
        if u.shape != self.shape:
            raise ValueError(f"{what} shape {u.shape} != {self.shape}")
        return u

    def forward(self, U0):
        """(observations y[nsteps][nrec], final modal state)."""
        u = self._state(U0)
        y = np.zeros((self.nsteps, self.nrec))
        UT = np.zeros(self.shape)
        _check(self.lib, self.lib.dgx_euler_adjoint_forward(self.h, _dp(u), _dp(y), _dp(UT)))
        return y, UT

    def objective(self, U0, data):
        y, _ = self.forward(U0)
        return 0.5 * self.dt * float(np.sum((y - data) ** 2))

    def gradient(self, U0, data):
        """(J, dJ/dU0) by the checkpointed discrete adjoint."""
        u = self._state(U0)
        d = np.ascontiguousarray(data, dtype=np.float64)
        if d.shape != (self.nsteps, self.nrec):
            raise ValueError(f"data shape {d.shape} != {(self.nsteps, self.nrec)}")
        J = np.zeros(1)
        g = np.zeros(self.shape)
        _check(self.lib, self.lib.dgx_euler_adjoint_gradient(self.h, _dp(u), _dp(d), _dp(J), _dp(g)))
        return float(J[0]), g

    def vjp(self, U, lam):
        """(dR/dU(U))^T lam for the modal residual R = -dU/dt."""
        u, l = self._state(U), self._state(lam, "lam")
        out = np.zeros(self.shape)
        _check(self.lib, self.lib.dgx_euler_adjoint_vjp(self.h, _dp(u), _dp(l), _dp(out)))
        return out

    def stats(self):
        f, a, s = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_int()
        _check(self.lib, self.lib.dgx_euler_adjoint_stats(self.h, ctypes.byref(f), ctypes.byref(a),
                                                          ctypes.byref(s)))
        return {"forward_steps": int(f.value), "adjoint_steps": int(a.value), "snapshots": int(s.value)}
