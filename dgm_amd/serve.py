"""Serving loop for ensembles of independent forward solves (ctypes binding of
the dgx_server_* entry points of libdgx.so).

Requests arrive at any time (``submit``).  The library admits them in arrival
order as slots and the device-memory budget free up.  It runs each on its own
stream, so the work of different requests overlaps on the GPU, and replays
each request's step slices from a captured hipGraph.  ``progress`` runs one
non-blocking scheduling pass; ``run_until_done`` drives it.  There is no
counterpart in the reference: the loop sits on its step seam (TimeInt::advance,
src/TimeInt.hpp:177).
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np

from . import dgx

QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_size_t), ("max_active", ctypes.c_int),
                ("steps_per_slice", ctypes.c_int), ("use_graphs", ctypes.c_int), ("batch", ctypes.c_int),
                ("persistent", ctypes.c_int)]


def _bind(lib):
    if getattr(lib, "_serve_bound", False):
        return lib
    h = ctypes.c_void_p
    _dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "dgx_domain_bytes": ([ctypes.POINTER(dgx.DomainDesc), ctypes.POINTER(ctypes.c_size_t)], ctypes.c_int),
        "dgx_server_create": ([ctypes.POINTER(ServerConfig), ctypes.POINTER(h)], ctypes.c_int),
        "dgx_server_destroy": ([h], ctypes.c_int),
        "dgx_server_submit": ([h, ctypes.POINTER(dgx.DomainDesc), _dp, ctypes.c_int, ctypes.c_int,
                               ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_longlong)], ctypes.c_int),
        "dgx_server_progress": ([h, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "dgx_server_status": ([h, ctypes.c_longlong, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "dgx_server_result": ([h, ctypes.c_longlong, _dp, ctypes.c_int, _dp], ctypes.c_int),
        "dgx_server_peak": ([h, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    lib._serve_bound = True
    return lib


def _check(lib, rc):
    if rc != 0:
        raise dgx.DgxError(rc, lib.dgx_last_error().decode())


def _desc(mesh, p, eqntype, gamma, vis, Pr, flux, precision, viscous=dgx.DGX_VISCOUS_BR1, br2_eta=0.0,
          balance=True, bc_state=None):
    vert = np.ascontiguousarray(mesh.vert, dtype=np.float64)
    con = np.ascontiguousarray(mesh.con, dtype=np.int32)
    d = dgx.DomainDesc(int(mesh.nsd), int(mesh.ne), int(p), int(eqntype), int(flux), int(precision),
                       float(gamma), float(vis), float(Pr),
                       vert.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), len(con),
                       con.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), int(viscous), float(br2_eta))
    d.no_balance = 0 if balance else 1
    if bc_state is not None:
        vals = [float(v) for v in bc_state] + [0.0] * (5 - len(bc_state))
        d.bc_state = (ctypes.c_double * 5)(*vals[:5])
    return d, (vert, con)


def domain_bytes(mesh, p, eqntype, precision=dgx.DGX_FP64):
    """Device bytes a domain of this mesh and order holds (admission size)."""
    lib = _bind(dgx.load())
    d, keep = _desc(mesh, p, eqntype, 1.4, 0.0, 0.72, 0, precision)
    out = ctypes.c_size_t()
    _check(lib, lib.dgx_domain_bytes(ctypes.byref(d), ctypes.byref(out)))
    return int(out.value)


@dataclass
class Request:
    id: int
    nf: int
    ne: int
    per_elem: int
    layout: int


class Server:
    """Continuous-batching server of forward solves on one GPU."""

    def __init__(self, hbm_budget: int, max_active: int = 64, steps_per_slice: int = 10,
                 use_graphs: bool = True, batch: bool = True, persistent: bool = False):
        self.lib = _bind(dgx.load())
        cfg = ServerConfig(int(hbm_budget), int(max_active), int(steps_per_slice), int(bool(use_graphs)),
                           int(bool(batch)), int(bool(persistent)))
        self.h = ctypes.c_void_p()
        _check(self.lib, self.lib.dgx_server_create(ctypes.byref(cfg), ctypes.byref(self.h)))
        self.requests: Dict[int, Request] = {}

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.dgx_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, mesh, p, eqntype, U0, inttype, dt, nsteps, gamma=1.4, vis=0.0, Pr=0.72,
               flux=dgx.DGX_FLUX_LF, precision=dgx.DGX_FP64, layout=dgx.DGX_NODAL,
               viscous=dgx.DGX_VISCOUS_BR1, br2_eta=0.0, balance=True, bc_state=None) -> int:
        d, keep = _desc(mesh, p, eqntype, gamma, vis, Pr, flux, precision, viscous, br2_eta, balance, bc_state)
        U0 = np.ascontiguousarray(U0, dtype=np.float64)
        nf = 3 if eqntype == dgx.DGX_SHALLOW else mesh.nsd + 2
        L, Q = p + 1, p + 2
        per = (Q if layout == dgx.DGX_NODAL else L) ** mesh.nsd
        if U0.shape != (nf, mesh.ne, per):
            raise ValueError(f"initial state shape {U0.shape} != {(nf, mesh.ne, per)}")
        rid = ctypes.c_longlong()
        _check(self.lib, self.lib.dgx_server_submit(self.h, ctypes.byref(d),
                                                    U0.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                    int(layout), int(inttype), float(dt), int(nsteps),
                                                    ctypes.byref(rid)))
        self.requests[rid.value] = Request(rid.value, nf, mesh.ne, per, layout)
        return rid.value

    def progress(self) -> int:
        pending = ctypes.c_int()
        _check(self.lib, self.lib.dgx_server_progress(self.h, ctypes.byref(pending)))
        return pending.value

    def status(self, rid: int) -> int:
        st = ctypes.c_int()
        _check(self.lib, self.lib.dgx_server_status(self.h, rid, ctypes.byref(st)))
        return st.value

    def result(self, rid: int):
        """(final state, (t_submit, t_admit, t_done)) of a finished request."""
        r = self.requests[rid]
        U = np.zeros((r.nf, r.ne, r.per_elem))
        t = np.zeros(3)
        _check(self.lib, self.lib.dgx_server_result(self.h, rid, U.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                    r.layout, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return U, tuple(t)

    def peak(self):
        b, a = ctypes.c_size_t(), ctypes.c_int()
        _check(self.lib, self.lib.dgx_server_peak(self.h, ctypes.byref(b), ctypes.byref(a)))
        return int(b.value), int(a.value)

    def run_until_done(self, poll_s: float = 0.0002, timeout_s: Optional[float] = None) -> None:
        t0 = time.perf_counter()
        while self.progress() > 0:
            if timeout_s is not None and time.perf_counter() - t0 > timeout_s:
                raise TimeoutError("serving loop did not finish")
            time.sleep(poll_s)


def poisson_arrivals(n: int, rate: float, rng) -> List[float]:
    """Arrival times (seconds) of a Poisson process of the given rate."""
    return list(np.cumsum(rng.exponential(1.0 / rate, size=n)))


def serve_timeline(server: Server, jobs, arrivals, poll_s: float = 0.0002, timeout_s: float = 600.0,
                   report_every: int = 0):
    """Submit job i at arrivals[i] seconds after the start (jobs[i] is a dict of
    Server.submit keyword arguments) while driving the scheduler; returns the
    request ids in submission order.  report_every > 0 prints a line each time
    that many more requests have finished."""
    ids: List[int] = []
    t0 = time.perf_counter()
    i = 0
    reported = 0
    while True:
        now = time.perf_counter() - t0
        while i < len(jobs) and arrivals[i] <= now:
            ids.append(server.submit(**jobs[i]))
            i += 1
        pending = server.progress()
        done = i - pending
        if report_every > 0 and done >= reported + report_every:
            reported = done - done % report_every
            print(f"[serve] {done} of {len(jobs)} requests finished at {now:.1f} s", flush=True)
        if i == len(jobs) and pending == 0:
            return ids
        if now > timeout_s:
            raise TimeoutError("serving loop did not finish")
        time.sleep(poll_s)
