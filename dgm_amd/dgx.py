"""ctypes binding of libdgx.so (include/dgx.h): one DG domain on one GPU.

The library is the product path; this module fails loudly when it is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DGX_LIBRARY") or os.path.join(HERE, "libdgx.so")

DGX_EULER, DGX_NAVIER_STOKES, DGX_EULER3D, DGX_NAVIER_STOKES3D, DGX_SHALLOW = 1, 2, 8, 9, 30
DGX_FLUX_LF, DGX_FLUX_ROE = 0, 1
DGX_VISCOUS_BR1, DGX_VISCOUS_BR2 = 0, 1
DGX_BC_WALL, DGX_BC_STATE, DGX_BC_FUNCTION, DGX_BC_ISOTHERMAL = 0, 1, 2, 3
DGX_FP64, DGX_MIXED = 0, 1
DGX_TVD, DGX_RK4, DGX_LSRK14 = 3, 4, 14
DGX_NODAL, DGX_MODAL = 0, 1

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class Function(ctypes.Structure):
    """dgx_function: per-component postfix programs of a function of (x, y, z, t)."""
    _fields_ = [("ncomp", ctypes.c_int), ("start", _ip), ("code", _ip), ("nconst", ctypes.c_int),
                ("consts", _dp)]


class DomainDesc(ctypes.Structure):
    _fields_ = [("nsd", ctypes.c_int), ("ne", ctypes.c_int), ("p", ctypes.c_int),
                ("eqntype", ctypes.c_int), ("flux", ctypes.c_int), ("precision", ctypes.c_int),
                ("gamma", ctypes.c_double), ("vis", ctypes.c_double), ("Pr", ctypes.c_double),
                ("vert", _dp), ("ncon", ctypes.c_int), ("con", _ip),
                ("viscous", ctypes.c_int), ("br2_eta", ctypes.c_double), ("bc_state", ctypes.c_double * 5),
                ("xnodes", _dp), ("no_balance", ctypes.c_int),
                ("bc_function", ctypes.POINTER(Function)), ("source", ctypes.POINTER(Function)),
                ("wall_T", ctypes.c_double), ("wall_Eref", ctypes.c_double), ("bulk_forcing", ctypes.c_int),
                ("porder", _ip)]


class DgxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"dgx error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load libdgx.so; raise if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    h = ctypes.c_void_p
    sig = {
        "dgx_domain_create": ([ctypes.POINTER(DomainDesc), ctypes.c_void_p, ctypes.POINTER(h)], ctypes.c_int),
        "dgx_domain_destroy": ([h], ctypes.c_int),
        "dgx_sizes": ([h, _ip, _ip, _ip, _ip], ctypes.c_int),
        "dgx_coordinates": ([h, _dp], ctypes.c_int),
        "dgx_vfield_sizes": ([h, _ip, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                              ctypes.POINTER(ctypes.c_longlong), _ip], ctypes.c_int),
        "dgx_state_upload": ([h, _dp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "dgx_state_download": ([h, _dp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "dgx_state_device_ptr": ([h, ctypes.POINTER(_dp)], ctypes.c_int),
        "dgx_rhs": ([h], ctypes.c_int),
        "dgx_solver_iterations": ([h, _ip], ctypes.c_int),
        "dgx_set_time": ([h, ctypes.c_double], ctypes.c_int),
        "dgx_get_time": ([h, _dp], ctypes.c_int),
        "dgx_residual_download": ([h, _dp, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "dgx_advance": ([h, ctypes.c_int, ctypes.c_double, ctypes.c_int], ctypes.c_int),
        "dgx_stage": ([h, ctypes.c_int, ctypes.c_int, ctypes.c_double], ctypes.c_int),
        "dgx_synchronize": ([h], ctypes.c_int),
        "dgx_set_time_step": ([h, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp], ctypes.c_int),
        "dgx_norms": ([h, _dp], ctypes.c_int),
        "dgx_stage_count": ([h], ctypes.c_longlong),
        "dgx_profile": ([h, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _ip, ctypes.c_int],
                        ctypes.c_int),
        "dgx_hexblock_create": ([ctypes.c_longlong, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.POINTER(h)], ctypes.c_int),
        "dgx_hexblock_destroy": ([h], ctypes.c_int),
        "dgx_hexblock_set_wJ": ([h, _dp, ctypes.c_int], ctypes.c_int),
        "dgx_hexblock_set_basis": ([h, _dp, _dp, _dp], ctypes.c_int),
        "dgx_hexblock_inner_product": ([h, _dp, ctypes.c_int], ctypes.c_int),
        "dgx_hexblock_backward_transform": ([h, _dp, ctypes.c_int], ctypes.c_int),
        "dgx_hexblock_synchronize": ([h], ctypes.c_int),
        "dgx_last_error": ([], ctypes.c_char_p),
        "dgx_version": ([], ctypes.c_char_p),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise DgxError(rc, load().dgx_last_error().decode())


def function_arrays(fn):
    """(start, code, consts) int32/int32/float64 arrays of a function: a
    formats.read_functions block (name -> expression tree) or the arrays
    themselves."""
    if isinstance(fn, dict):
        from .formats import compile_function
        return compile_function(fn)
    start, code, consts = fn
    return (np.ascontiguousarray(start, dtype=np.int32), np.ascontiguousarray(code, dtype=np.int32),
            np.ascontiguousarray(consts, dtype=np.float64))


def make_function(fn, keep):
    """A dgx_function pointer for fn (None: NULL); keep holds the arrays alive."""
    if fn is None:
        return None
    start, code, consts = function_arrays(fn)
    if consts.size == 0:
        consts = np.zeros(1)
    keep += [start, code, consts]
    f = Function(len(start) - 1, start.ctypes.data_as(_ip), code.ctypes.data_as(_ip), int(consts.size),
                 consts.ctypes.data_as(_dp))
    keep.append(f)
    return ctypes.pointer(f)


class GPUDomain:
    """A DG domain on the GPU: modal FP64 state, residual and RK stepping."""

    def __init__(self, mesh, p, eqntype, gamma=1.4, vis=0.0, Pr=0.72, flux=DGX_FLUX_LF,
                 precision=DGX_FP64, stream=None, viscous=DGX_VISCOUS_BR1, br2_eta=0.0, bc_state=None,
                 xnodes=None, balance=True, bc_function=None, source=None, wall=None, bulk_forcing=False,
                 porder=None):
        lib = load()
        self.lib = lib
        self.nsd = int(mesh.nsd)
        self._vert = np.ascontiguousarray(mesh.vert, dtype=np.float64)
        self._con = np.ascontiguousarray(mesh.con, dtype=np.int32)
        desc = DomainDesc(self.nsd, int(mesh.ne), int(p), int(eqntype), int(flux), int(precision),
                          float(gamma), float(vis), float(Pr), _ptr(self._vert), len(self._con),
                          self._con.ctypes.data_as(_ip), int(viscous), float(br2_eta),
                          (ctypes.c_double * 5)(*([0.0] * 5 if bc_state is None else [float(v) for v in bc_state])),
                          None if xnodes is None else _ptr(self._nodes(xnodes, mesh, p)),
                          0 if balance else 1)
        self._fkeep = []
        desc.bc_function = make_function(bc_function, self._fkeep)
        desc.source = make_function(source, self._fkeep)
        if wall is not None:       # isothermal walls: (Tw, Eref)
            desc.wall_T, desc.wall_Eref = float(wall[0]), float(wall[1])
        desc.bulk_forcing = int(bool(bulk_forcing))
        self.porder = None
        if porder is not None:     # variable order: one order per element, p the largest
            self.porder = np.ascontiguousarray(porder, dtype=np.int32)
            desc.porder = self.porder.ctypes.data_as(_ip)
        self.h = ctypes.c_void_p()
        _check(lib.dgx_domain_create(ctypes.byref(desc), ctypes.c_void_p(stream or 0),
                                     ctypes.byref(self.h)))
        q, m, f, ne = (ctypes.c_int() for _ in range(4))
        _check(lib.dgx_sizes(self.h, ctypes.byref(q), ctypes.byref(m), ctypes.byref(f), ctypes.byref(ne)))
        self.qtot, self.nm, self.nf, self.ne = q.value, m.value, f.value, ne.value
        self.p = int(p)

    def _nodes(self, xnodes, mesh, p):
        q, nsd = int(p) + 2, int(mesh.nsd)
        self._xn = np.ascontiguousarray(xnodes, dtype=np.float64)
        if self._xn.shape != (nsd, int(mesh.ne), q ** nsd):
            raise ValueError(f"xnodes shape {self._xn.shape} != {(nsd, int(mesh.ne), q ** nsd)}")
        return self._xn

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.dgx_domain_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def vfield_sizes(self):
        """dict of the state's vField totals: size, qtot, nmodes, sdof, nsd."""
        sz, nsd = ctypes.c_int(), ctypes.c_int()
        q, m, s = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        _check(self.lib.dgx_vfield_sizes(self.h, ctypes.byref(sz), ctypes.byref(q), ctypes.byref(m), ctypes.byref(s),
                                         ctypes.byref(nsd)))
        return {"size": sz.value, "qtot": q.value, "nmodes": m.value, "sdof": s.value, "nsd": nsd.value}

    def coords(self):
        x = np.zeros((self.nsd, self.ne, self.qtot))
        _check(self.lib.dgx_coordinates(self.h, _ptr(x)))
        return x

    def upload(self, U, layout=DGX_NODAL):
        U = np.ascontiguousarray(U, dtype=np.float64)
        n = self.qtot if layout == DGX_NODAL else self.nm
        if U.shape != (self.nf, self.ne, n):
            raise ValueError(f"state shape {U.shape} != {(self.nf, self.ne, n)}")
        _check(self.lib.dgx_state_upload(self.h, _ptr(U), layout, 0))

    def download(self, layout=DGX_NODAL):
        n = self.qtot if layout == DGX_NODAL else self.nm
        U = np.zeros((self.nf, self.ne, n))
        _check(self.lib.dgx_state_download(self.h, _ptr(U), layout, 0))
        return U

    def rhs(self, layout=DGX_NODAL):
        _check(self.lib.dgx_rhs(self.h))
        n = self.qtot if layout == DGX_NODAL else self.nm
        F = np.zeros((self.nf, self.ne, n))
        _check(self.lib.dgx_residual_download(self.h, _ptr(F), layout, 0))
        self.synchronize()
        return F

    def solver_iterations(self):
        """Most conjugate-gradient iterations of a curved hex mass solve so far."""
        it = ctypes.c_int()
        _check(self.lib.dgx_solver_iterations(self.h, ctypes.byref(it)))
        return it.value

    def set_time(self, t):
        _check(self.lib.dgx_set_time(self.h, float(t)))

    @property
    def time(self):
        t = ctypes.c_double()
        _check(self.lib.dgx_get_time(self.h, ctypes.byref(t)))
        return t.value

    def advance(self, inttype, dt, nsteps):
        _check(self.lib.dgx_advance(self.h, int(inttype), float(dt), int(nsteps)))

    def set_time_step(self, dt, CFLmax=0.0, dtmax=0.0):
        """(CFL number, new dt) of the current state (Domain::set_time_step)."""
        c, d = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.dgx_set_time_step(self.h, float(dt), float(CFLmax), float(dtmax), ctypes.byref(c),
                                          ctypes.byref(d)))
        return c.value, d.value

    def stage(self, inttype, stage, dt):
        """Stage `stage` of one step (dgx_stage)."""
        _check(self.lib.dgx_stage(self.h, int(inttype), int(stage), float(dt)))

    def synchronize(self):
        _check(self.lib.dgx_synchronize(self.h))

    def norms(self):
        out = np.zeros((self.nf, 3))
        _check(self.lib.dgx_norms(self.h, _ptr(out)))
        return out

    def device_state_ptr(self):
        p = _dp()
        _check(self.lib.dgx_state_device_ptr(self.h, ctypes.byref(p)))
        return ctypes.cast(p, ctypes.c_void_p).value

    @property
    def stage_count(self):
        return int(self.lib.dgx_stage_count(self.h))
