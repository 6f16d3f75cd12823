"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) -- test infrastructure.

Only tests, the smoke check and the benchmark's CPU-baseline leg import this.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None
# callback (t, n, x, y, out[3 x n]) of shallow-water boundaries and sources
FIELD_FN = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_int, _dp, _dp, _dp)
SHALLOW = 30


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    lib = ctypes.CDLL(LIB)
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip,
                               ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                               ctypes.c_int]
    for name in ("orc_destroy", "orc_project_state"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
        getattr(lib, name).restype = None
    for name in ("orc_qtot", "orc_nmodes", "orc_status"):
        getattr(lib, name).argtypes = [ctypes.c_void_p]
        getattr(lib, name).restype = ctypes.c_int
    for name in ("orc_coords", "orc_set_state", "orc_get_state", "orc_modal", "orc_set_modal",
                 "orc_norms"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, _dp]
        getattr(lib, name).restype = None
    lib.orc_geometry.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.orc_basis.argtypes = [ctypes.c_void_p] + [_dp] * 6
    lib.orc_rhs.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.orc_rhs.restype = ctypes.c_int
    lib.orc_advance.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    lib.orc_advance.restype = ctypes.c_int
    lib.orc_riemann.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp]
    lib.orc_riemann.restype = ctypes.c_int
    lib.orc_set_viscous.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
    lib.orc_set_bc_state.argtypes = [ctypes.c_void_p, _dp]
    lib.orc_set_nodes.argtypes = [ctypes.c_void_p, _dp]
    lib.orc_set_linear.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.orc_set_functions.argtypes = [ctypes.c_void_p, FIELD_FN, FIELD_FN]
    lib.orc_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
    lib.orc_get_time.argtypes = [ctypes.c_void_p]
    lib.orc_get_time.restype = ctypes.c_double
    lib.orc_set_balance.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_set_isothermal.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
    lib.orc_set_channel.argtypes = [ctypes.c_void_p, ctypes.c_int]
    lib.orc_set_time_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
    lib.orc_set_time_step.restype = ctypes.c_int
    lib.orc_varp_create.restype = ctypes.c_void_p
    lib.orc_varp_create.argtypes = [ctypes.c_int, _ip, _dp, ctypes.c_int, _ip, ctypes.c_double, ctypes.c_int]
    lib.orc_varp_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_varp_destroy.restype = None
    lib.orc_varp_qmax.argtypes = [ctypes.c_void_p]
    lib.orc_varp_qmax.restype = ctypes.c_int
    for name in ("orc_varp_coords", "orc_varp_set_state", "orc_varp_get_state", "orc_varp_modal",
                 "orc_varp_set_modal", "orc_varp_norms", "orc_varp_set_bc_state"):
        getattr(lib, name).argtypes = [ctypes.c_void_p, _dp]
        getattr(lib, name).restype = None
    lib.orc_varp_set_time_step.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
    lib.orc_varp_set_time_step.restype = ctypes.c_int
    lib.orc_varp_project_state.argtypes = [ctypes.c_void_p]
    lib.orc_varp_project_state.restype = None
    lib.orc_varp_rhs.argtypes = [ctypes.c_void_p, _dp, _dp]
    lib.orc_varp_rhs.restype = ctypes.c_int
    lib.orc_varp_advance.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    lib.orc_varp_advance.restype = ctypes.c_int
    lib.oracle_hexblock_inner_product.argtypes = [ctypes.c_longlong] + [ctypes.c_int] * 4 + [_dp, ctypes.c_int] + \
        [_dp] * 4
    lib.oracle_hexblock_inner_product.restype = None
    lib.oracle_hexblock_backward_transform.argtypes = [ctypes.c_longlong] + [ctypes.c_int] * 4 + [_dp] * 4
    lib.oracle_hexblock_backward_transform.restype = None
    _lib = lib
    return lib


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


class Oracle:
    """CPU restatement of one DGM Euler/NS domain (nodal GLL storage)."""

    def __init__(self, mesh, p, eqn, gamma=1.4, vis=0.0, Pr=0.72, flux=0, viscous=0, br2_eta=6.0):
        self.lib = _load()
        self.nsd = mesh.nsd
        self.ne = mesh.ne
        self.p = p
        self.nv = 3 if eqn == SHALLOW else self.nsd + 2
        self._vert = np.ascontiguousarray(mesh.vert, dtype=np.float64)
        self._con = np.ascontiguousarray(mesh.con, dtype=np.int32)
        self.h = self.lib.orc_create(self.nsd, self.ne, p, _ptr(self._vert), len(self._con),
                                     _ptr(self._con, _ip), eqn, gamma, vis, Pr, flux)
        self.qtot = self.lib.orc_qtot(self.h)
        self.nm = self.lib.orc_nmodes(self.h)
        if viscous:
            self.lib.orc_set_viscous(self.h, int(viscous), float(br2_eta))

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass

    def set_nodes(self, xyz):
        """Curved elements: nodal coordinates [nsd][ne][qtot] at the GLL points."""
        self._xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        assert self._xyz.shape == (self.nsd, self.ne, self.qtot), self._xyz.shape
        self.lib.orc_set_nodes(self.h, _ptr(self._xyz))

    def set_bc_state(self, U):
        """Conservative state imposed on state boundaries (3D Euler; shallow water uses 3 values)."""
        U = np.ravel(np.asarray(U, dtype=np.float64))
        self._bcU = np.zeros(5)
        self._bcU[:len(U)] = U
        self.lib.orc_set_bc_state(self.h, _ptr(self._bcU))

    def set_functions(self, bc=None, src=None):
        """Shallow water: Python callables (t, x, y) -> array [3, n] for function
        boundaries (type 2 rows) and sources; None leaves that part out."""
        def wrap(fn):
            if fn is None:
                return FIELD_FN()
            def cb(t, n, x, y, out):
                xs = np.ctypeslib.as_array(x, shape=(n,)).copy()
                ys = np.ctypeslib.as_array(y, shape=(n,)).copy()
                np.ctypeslib.as_array(out, shape=(3 * n,))[:] = np.asarray(fn(t, xs, ys), dtype=np.float64).reshape(-1)
            return FIELD_FN(cb)
        self._cbs = (wrap(bc), wrap(src))      # kept alive with the oracle
        self.lib.orc_set_functions(self.h, *self._cbs)

    def set_time(self, t):
        self.lib.orc_set_time(self.h, float(t))

    def time(self):
        return self.lib.orc_get_time(self.h)

    def set_isothermal(self, Tw, Eref):
        """Navier-Stokes 3D isothermal walls (boundary type 3): wall temperature and reference energy."""
        self.lib.orc_set_isothermal(self.h, float(Tw), float(Eref))

    def set_channel(self, on):
        """Navier-Stokes 3D: subtract the domain means of the rho, rho u and E residuals (bulk forcing)."""
        self.lib.orc_set_channel(self.h, int(bool(on)))

    def set_balance(self, on):
        """Shallow water: subtract the first residual from every residual (default on)."""
        self.lib.orc_set_balance(self.h, int(bool(on)))

    def set_linear(self, Ax, Ay, u0, v0, c0):
        """Matrices of the linear system (eqn 4); Lax-Friedrichs speed |u0 nx + v0 ny| + c0."""
        self._Ax = np.ascontiguousarray(Ax, dtype=np.float64).reshape(16)
        self._Ay = np.ascontiguousarray(Ay, dtype=np.float64).reshape(16)
        self.lib.orc_set_linear(self.h, _ptr(self._Ax), _ptr(self._Ay), float(u0), float(v0), float(c0))

    def coords(self):
        x = np.zeros((self.nsd, self.ne, self.qtot))
        self.lib.orc_coords(self.h, _ptr(x))
        return x

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        self._check(U)
        self.lib.orc_set_state(self.h, _ptr(U))

    def state(self):
        U = np.zeros((self.nv, self.ne, self.qtot))
        self.lib.orc_get_state(self.h, _ptr(U))
        return U

    def project_state(self):
        self.lib.orc_project_state(self.h)

    def modal(self):
        Uh = np.zeros((self.nv, self.ne, self.nm))
        self.lib.orc_modal(self.h, _ptr(Uh))
        return Uh

    def set_modal(self, Uh):
        Uh = np.ascontiguousarray(Uh, dtype=np.float64)
        if Uh.shape != (self.nv, self.ne, self.nm):
            raise ValueError(f"modal shape {Uh.shape} != {(self.nv, self.ne, self.nm)}")
        self.lib.orc_set_modal(self.h, _ptr(Uh))

    def _check(self, U):
        if U.shape != (self.nv, self.ne, self.qtot):
            raise ValueError(f"state shape {U.shape} != {(self.nv, self.ne, self.qtot)}")

    def rhs(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        self._check(U)
        F = np.zeros_like(U)
        st = self.lib.orc_rhs(self.h, _ptr(U), _ptr(F))
        if st:
            raise FloatingPointError("oracle: non-realizable state")
        return F

    def advance(self, inttype, dt, nsteps):
        st = self.lib.orc_advance(self.h, inttype, dt, nsteps)
        if st:
            raise FloatingPointError(f"oracle advance failed ({st})")

    def set_time_step(self, dt, CFLmax=0.0, dtmax=0.0):
        """(CFL, new dt) of the current state (Euler::set_time_step)."""
        out = np.zeros(2)
        if self.lib.orc_set_time_step(self.h, float(dt), float(CFLmax), float(dtmax), _ptr(out)):
            raise FloatingPointError("oracle: non-realizable state in the CFL evaluation")
        return float(out[0]), float(out[1])

    def norms(self):
        out = np.zeros((self.nv, 3))
        self.lib.orc_norms(self.h, _ptr(out))
        return out

    def geometry(self):
        J = np.zeros(self.ne)
        M = np.zeros((self.ne, 9))
        self.lib.orc_geometry(self.h, _ptr(J), _ptr(M))
        return J, M

    def basis(self):
        q, qf, L = self.p + 2, self.p + 1, self.p + 1
        z, w, zg, wg = np.zeros(q), np.zeros(q), np.zeros(qf), np.zeros(qf)
        D, B = np.zeros((q, q)), np.zeros((L, q))
        self.lib.orc_basis(self.h, _ptr(z), _ptr(w), _ptr(zg), _ptr(wg), _ptr(D), _ptr(B))
        return dict(z=z, w=w, zg=zg, wg=wg, D=D, B=B)


class VarOracle:
    """The oracle's variable-order quads (2D Euler, affine): one order per
    element (root.ord).  Arrays are padded per element: nodal [4][ne][qmax^2]
    with element e's (p_e+2)^2 nodes first, modal [4][ne][(pmax+1)^2] with its
    (p_e+1)^2 modes first; the rest is zero."""

    def __init__(self, mesh, orders, gamma=1.4, flux=0):
        self.lib = _load()
        if mesh.nsd != 2:
            raise ValueError("variable order: quads only")
        self.nsd, self.ne, self.nv = 2, mesh.ne, 4
        self.orders = np.ascontiguousarray(orders, dtype=np.int32)
        self.p = int(self.orders.max())
        self._vert = np.ascontiguousarray(mesh.vert, dtype=np.float64)
        self._con = np.ascontiguousarray(mesh.con, dtype=np.int32)
        self.h = self.lib.orc_varp_create(self.ne, _ptr(self.orders, _ip), _ptr(self._vert), len(self._con),
                                          _ptr(self._con, _ip), gamma, flux)
        self.qtot = (self.p + 2) ** 2
        self.nm = (self.p + 1) ** 2

    def __del__(self):
        try:
            self.lib.orc_varp_destroy(self.h)
        except Exception:
            pass

    def coords(self):
        x = np.zeros((2, self.ne, self.qtot))
        self.lib.orc_varp_coords(self.h, _ptr(x))
        return x

    def set_bc_state(self, U):
        """Conservative state of state boundaries (type 1 rows)."""
        self._bcU = np.zeros(5)
        self._bcU[:4] = np.ravel(np.asarray(U, dtype=np.float64))[:4]
        self.lib.orc_varp_set_bc_state(self.h, _ptr(self._bcU))

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (4, self.ne, self.qtot), U.shape
        self.lib.orc_varp_set_state(self.h, _ptr(U))

    def state(self):
        U = np.zeros((4, self.ne, self.qtot))
        self.lib.orc_varp_get_state(self.h, _ptr(U))
        return U

    def project_state(self):
        self.lib.orc_varp_project_state(self.h)

    def modal(self):
        Uh = np.zeros((4, self.ne, self.nm))
        self.lib.orc_varp_modal(self.h, _ptr(Uh))
        return Uh

    def set_modal(self, Uh):
        Uh = np.ascontiguousarray(Uh, dtype=np.float64)
        assert Uh.shape == (4, self.ne, self.nm), Uh.shape
        self.lib.orc_varp_set_modal(self.h, _ptr(Uh))

    def rhs(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64)
        assert U.shape == (4, self.ne, self.qtot), U.shape
        F = np.zeros_like(U)
        if self.lib.orc_varp_rhs(self.h, _ptr(U), _ptr(F)):
            raise FloatingPointError("oracle: non-realizable state")
        return F

    def advance(self, inttype, dt, nsteps):
        st = self.lib.orc_varp_advance(self.h, inttype, dt, nsteps)
        if st:
            raise FloatingPointError(f"oracle advance failed ({st})")

    def norms(self):
        out = np.zeros((4, 3))
        self.lib.orc_varp_norms(self.h, _ptr(out))
        return out

    def set_time_step(self, dt, CFLmax=0.0, dtmax=0.0):
        """(CFL, new dt) of the current state (Euler::set_time_step)."""
        out = np.zeros(2)
        if self.lib.orc_varp_set_time_step(self.h, float(dt), float(CFLmax), float(dtmax), _ptr(out)):
            raise FloatingPointError("oracle: non-realizable state in the CFL evaluation")
        return float(out[0]), float(out[1])


def riemann(nsd, flux, UL, UR, n, gamma=1.4):
    """Numerical flux of the oracle at one face point (Cartesian components)."""
    lib = _load()
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    if UL.shape != (nsd + 2,) or UR.shape != (nsd + 2,) or n.shape != (nsd,):
        raise ValueError("riemann: state or normal size")
    F = np.zeros(nsd + 2)
    if lib.orc_riemann(nsd, flux, gamma, _ptr(UL), _ptr(UR), _ptr(n), _ptr(F)):
        raise FloatingPointError("non-realizable state")
    return F


def dgm_diff(oracle_a: Oracle, Uh_ref, Uh_test):
    """``Domain::diff`` norms (src/Domain.cpp:1000-1071): returns (ref, diff) norm arrays."""
    oracle_a.set_modal(Uh_ref)
    ref = oracle_a.norms()
    oracle_a.set_modal(Uh_ref - Uh_test)
    dif = oracle_a.norms()
    return ref, dif


def hexblock_inner_product(u, Ba, Bb, Bc, wJ, per_element=False):
    """The oracle's element-block inner product (GpGpu::hex_kernels /
    Hex::inner_product), on a copy: u [ne][qa*qb*qc] slots; returns the slots
    with each element's nm^3 inner products first."""
    lib = _load()
    nm, qa = Ba.shape
    qb, qc = Bb.shape[1], Bc.shape[1]
    out = np.array(u, dtype=np.float64, order="C", copy=True)
    wJ = np.ascontiguousarray(wJ, dtype=np.float64)
    B = [np.ascontiguousarray(x, dtype=np.float64) for x in (Ba, Bb, Bc)]
    lib.oracle_hexblock_inner_product(out.shape[0], nm, qa, qb, qc, _ptr(wJ), qa * qb * qc if per_element else 0,
                                      _ptr(B[0]), _ptr(B[1]), _ptr(B[2]), _ptr(out))
    return out


def hexblock_backward_transform(u, Ba, Bb, Bc):
    """The oracle's element-block backward transform (Hex::backward_transform),
    on a copy of the slots u [ne][qa*qb*qc]."""
    lib = _load()
    nm, qa = Ba.shape
    qb, qc = Bb.shape[1], Bc.shape[1]
    out = np.array(u, dtype=np.float64, order="C", copy=True)
    B = [np.ascontiguousarray(x, dtype=np.float64) for x in (Ba, Bb, Bc)]
    lib.oracle_hexblock_backward_transform(out.shape[0], nm, qa, qb, qc, _ptr(B[0]), _ptr(B[1]), _ptr(B[2]),
                                           _ptr(out))
    return out
