"""libdgx.so loads and exports the C ABI of include/dgx.h (no GPU needed).

Only argument validation runs here: every call below is rejected before the
library touches a device.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from dgm_amd import dgx, mesh as meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dgx.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return sorted(set(re.findall(r"\b(dgx_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(dgx.LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "dgm_amd", "csrc"), "-j8"])
    return dgx.load()


def test_every_declared_symbol_is_exported(lib):
    names = _declared()
    assert "dgx_rhs" in names and "dgx_advance" in names and len(names) >= 15
    raw = ctypes.CDLL(dgx.LIB_PATH)
    missing = [n for n in names if not hasattr(raw, n)]
    assert not missing, missing


def test_version_string(lib):
    import dgm_amd
    v = lib.dgx_version().decode()
    assert re.fullmatch(r"\d+\.\d+\.\d+", v)
    assert v == dgm_amd.__version__


def _desc(m, p, eqn, **kw):
    vert = np.ascontiguousarray(m.vert, dtype=np.float64)
    con = np.ascontiguousarray(m.con, dtype=np.int32)
    d = dgx.DomainDesc(nsd=m.nsd, ne=m.ne, p=p, eqntype=eqn, flux=kw.get("flux", 0),
                       precision=kw.get("precision", 0), gamma=1.4, vis=kw.get("vis", 0.0),
                       Pr=0.72, vert=vert.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                       ncon=len(con), con=con.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                       viscous=kw.get("viscous", 0), br2_eta=kw.get("br2_eta", 0.0))
    return d, (vert, con)


@pytest.mark.parametrize("case,msg", [
    (dict(p=0), b"p must be"),
    (dict(p=11), b"p must be"),
    (dict(eqn=1), b"eqntype"),
    (dict(flux=7), b"flux must be"),
    (dict(p=10, eqn=9), b"p <= 9"),
    (dict(p=9, eqn=8, precision=1), b"DGX_MIXED"),
    (dict(p=3, eqn=8, precision=7), b"precision"),
    (dict(p=3, eqn=9, viscous=2), b"viscous must be"),
    (dict(p=3, eqn=8, viscous=1), b"DGX_VISCOUS_BR2"),
    (dict(p=3, eqn=9, viscous=1, br2_eta=-1.0), b"br2_eta"),
    (dict(p=3, eqn=9, viscous=1, br2_eta=float("nan")), b"br2_eta"),
    (dict(p=3, eqn=9, boundary=0), b"no slip-wall viscous flux"),
    (dict(p=3, eqn=8, boundary=3), b"DGX_BC_ISOTHERMAL is a Navier_Stokes3d boundary"),
    (dict(p=3, eqn=9, boundary=3, top=3, precision=1), b"Navier_Stokes3d boundaries are available"),
    (dict(p=3, eqn=9, boundary=3, top=3, viscous=1), b"Navier_Stokes3d boundaries are available"),
    (dict(p=10, eqn=9, boundary=3, top=3), b"Navier_Stokes3d boundaries are available"),
    (dict(p=3, eqn=9, bulk=1, precision=1), b"bulk_forcing is available"),
    (dict(p=3, eqn=8, boundary=5), b"unknown boundary type"),
    (dict(p=8, eqn=9, curved=True), b"curved elements"),
    (dict(p=3, eqn=9, viscous=1, curved=True), b"curved elements"),
    (dict(p=3, eqn=9, precision=1, curved=True), b"DGX_MIXED"),
    (dict(p=8, eqn=9, boundary=3, top=3, curved=True), b"curved elements"),
    (dict(p=9, eqn=8, curved=True), b"curved elements"),
    (dict(p=9, eqn=8, precision=1, curved=True), b"DGX_MIXED"),
])
def test_create_rejects_bad_arguments(lib, case, msg):
    m = meshgen.periodic_hexes(2, 2, 2)
    if "boundary" in case:
        m = meshgen.channel_hexes(2, 2, 2, bc_bottom=case["boundary"], bc_top=case.get("top", 0))
    d, keep = _desc(m, case.get("p", 3), case.get("eqn", 8), flux=case.get("flux", 0),
                    precision=case.get("precision", 0), viscous=case.get("viscous", 0),
                    br2_eta=case.get("br2_eta", 0.0))
    d.bulk_forcing = case.get("bulk", 0)
    if case.get("curved"):
        xn = np.ascontiguousarray(meshgen.hex_nodes(m, case.get("p", 3)))
        keep = keep + (xn,)
        d.xnodes = xn.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    h = ctypes.c_void_p()
    rc = lib.dgx_domain_create(ctypes.byref(d), None, ctypes.byref(h))
    assert rc == -1
    assert msg in lib.dgx_last_error()
    assert not h.value


@pytest.mark.parametrize("case,msg", [
    (dict(flux=1), b"Lax-Friedrichs flux only"),
    (dict(precision=7), b"precision must be"),
    (dict(nsd3=True), b"eqntype does not match nsd"),
    (dict(boundary=4), b"unknown boundary type"),
    (dict(boundary=0, eqn=2), b"boundary sides are available"),
])
def test_shallow_create_rejects_bad_arguments(lib, case, msg):
    if case.get("nsd3"):
        m = meshgen.periodic_hexes(2, 2, 2)
    elif "boundary" in case:
        m = meshgen.box_quads(2, 2, bc=(case["boundary"], None, 0, None))
    else:
        m = meshgen.periodic_quads(2, 2)
    d, keep = _desc(m, 3, case.get("eqn", 30), flux=case.get("flux", 0), precision=case.get("precision", 0))
    h = ctypes.c_void_p()
    rc = lib.dgx_domain_create(ctypes.byref(d), None, ctypes.byref(h))
    assert rc == -1
    assert msg in lib.dgx_last_error(), lib.dgx_last_error()
    assert not h.value


def test_curved_navier_stokes_quads_rejected_with_boundaries(lib):
    m = meshgen.box_quads(2, 2, bc=(0, None, 0, None))
    d, keep = _desc(m, 3, 2, vis=0.01)
    xn = np.ascontiguousarray(meshgen.quad_nodes(m, 3))
    d.xnodes = xn.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    h = ctypes.c_void_p()
    assert lib.dgx_domain_create(ctypes.byref(d), None, ctypes.byref(h)) == -1
    assert b"boundary sides are available" in lib.dgx_last_error()


def test_mixed_rejected_for_2d_navier_stokes_and_curved_quads(lib):
    m = meshgen.periodic_quads(2, 2)
    d, keep = _desc(m, 3, 2, precision=1, vis=0.01)
    h = ctypes.c_void_p()
    assert lib.dgx_domain_create(ctypes.byref(d), None, ctypes.byref(h)) == -1
    assert b"DGX_MIXED" in lib.dgx_last_error()
    d, keep = _desc(m, 3, 1, precision=1)
    xn = np.ascontiguousarray(meshgen.quad_nodes(m, 3))
    d.xnodes = xn.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.dgx_domain_create(ctypes.byref(d), None, ctypes.byref(h)) == -1
    assert b"DGX_FP64" in lib.dgx_last_error()


def test_null_arguments_rejected(lib):
    h = ctypes.c_void_p()
    assert lib.dgx_domain_create(None, None, ctypes.byref(h)) == -1
    assert lib.dgx_rhs(None) == -1
    assert lib.dgx_advance(None, 14, 1e-3, 1) == -1
    assert lib.dgx_domain_destroy(None) == 0


def test_server_argument_checks(lib):
    from dgm_amd import serve
    serve._bind(lib)
    h = ctypes.c_void_p()
    bad = serve.ServerConfig(0, 4, 5, 1, 1, 0)
    assert lib.dgx_server_create(ctypes.byref(bad), ctypes.byref(h)) == -1
    assert b"positive" in lib.dgx_last_error()
    good = serve.ServerConfig(1 << 30, 4, 5, 1, 1, 0)
    assert lib.dgx_server_create(ctypes.byref(good), ctypes.byref(h)) == 0
    m = meshgen.periodic_quads(2, 2)
    d, keep = _desc(m, 3, 1)
    U = np.zeros((4, 4, 25))
    rid = ctypes.c_longlong()
    dp = U.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.dgx_server_submit(h, ctypes.byref(d), dp, 0, 14, 1e-3, 0, ctypes.byref(rid)) == -1
    assert lib.dgx_server_submit(h, ctypes.byref(d), dp, 0, 5, 1e-3, 10, ctypes.byref(rid)) == -1
    assert lib.dgx_server_submit(h, ctypes.byref(d), dp, 0, 14, 1e-3, 10, ctypes.byref(rid)) == 0
    st = ctypes.c_int()
    assert lib.dgx_server_status(h, rid.value, ctypes.byref(st)) == 0 and st.value == 0   # queued
    assert lib.dgx_server_status(h, 999, ctypes.byref(st)) == -1
    assert lib.dgx_server_destroy(h) == 0


def test_domain_bytes_estimate(lib):
    from dgm_amd import serve
    m = meshgen.periodic_hexes(4, 4, 4)
    b = serve.domain_bytes(m, 7, 9)
    state = 5 * 64 * 512 * 8
    assert 5 * state < b < 20 * state          # registers, traces, viscous terms, nodal buffer
    bm = serve.domain_bytes(m, 7, 9, precision=1)
    assert bm < b                               # FP32 viscous terms


@pytest.mark.parametrize("case,msg", [
    (dict(nsd=3), b"2D mesh"),
    (dict(eqn=1), b"DGX_LINEULER"),
    (dict(precision=7), b"DGX_FP64 or DGX_MIXED"),
    (dict(rho0=0.0), b"mean flow"),
    (dict(dt=0.0), b"dt > 0"),
    (dict(nsteps=0), b"nsteps"),
    (dict(nsnap=-1), b"nsnap"),
])
def test_acoustic_create_rejects_bad_arguments(lib, case, msg):
    from dgm_amd import acoustic
    acoustic._bind(lib)
    m = meshgen.periodic_hexes(2, 2, 2) if case.get("nsd") == 3 else meshgen.periodic_quads(3, 3)
    d, keep = _desc(m, 2, case.get("eqn", acoustic.DGX_LINEULER), precision=case.get("precision", 0))
    xr = np.array([[0.5, 0.5]])
    ad = acoustic.AcousticDesc(case.get("rho0", 1.0), 0.0, 0.0, 1.0, 1,
                               xr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 10.0, 0.1,
                               case.get("dt", 1e-3), case.get("nsteps", 5), case.get("nsnap", 0))
    h = ctypes.c_void_p()
    rc = lib.dgx_acoustic_create(ctypes.byref(d), ctypes.byref(ad), None, ctypes.byref(h))
    assert rc == -1 and msg in lib.dgx_last_error()
    assert not h.value


@pytest.mark.parametrize("case,msg", [
    (dict(nsd=3), b"2D mesh"),
    (dict(eqn=30), b"DGX_EULER"),
    (dict(vis=1e-3), b"vis = 0"),
    (dict(precision=1), b"DGX_FP64"),
    (dict(flux=7), b"Lax-Friedrichs and Roe"),
    (dict(walls=True), b"periodic"),
    (dict(dt=0.0), b"dt > 0"),
    (dict(nsteps=0), b"nsteps"),
    (dict(nrec=-1), b"receivers"),
    (dict(nsnap=-1), b"nsnap"),
])
def test_euler_adjoint_create_rejects_bad_arguments(lib, case, msg):
    from dgm_amd import euler_adjoint
    euler_adjoint._bind(lib)
    if case.get("nsd") == 3:
        m = meshgen.periodic_hexes(2, 2, 2)
    elif case.get("walls"):
        m = meshgen.box_quads(3, 3, bc=(0, 1, 0, 1))
    else:
        m = meshgen.periodic_quads(3, 3)
    d, keep = _desc(m, 2, case.get("eqn", 1), precision=case.get("precision", 0), flux=case.get("flux", 0),
                    vis=case.get("vis", 0.0))
    xr = np.array([[0.5, 0.5]])
    ad = euler_adjoint.EulerAdjointDesc(case.get("nrec", 1), xr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        case.get("dt", 1e-3), case.get("nsteps", 5), case.get("nsnap", 0))
    h = ctypes.c_void_p()
    rc = lib.dgx_euler_adjoint_create(ctypes.byref(d), ctypes.byref(ad), None, ctypes.byref(h))
    assert rc == -1 and msg in lib.dgx_last_error()
    assert not h.value


def _fn(start, code, consts, keep):
    start = np.asarray(start, dtype=np.int32)
    code = np.asarray(code, dtype=np.int32)
    consts = np.asarray(consts if len(consts) else [0.0], dtype=np.float64)
    keep += [start, code, consts]
    f = dgx.Function(len(start) - 1, start.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                     code.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(consts),
                     consts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    keep.append(f)
    return ctypes.pointer(f)


X, Y, T, ADD, SIN = 1, 2, 4, 5, 11


@pytest.mark.parametrize("which,start,code,consts,msg", [
    ("source", [0, 1, 2], [X, Y], [], b"3 components"),
    ("source", [0, 1, 2, 2], [X, Y], [], b"empty"),
    ("source", [0, 1, 2, 3], [X, Y, 99], [], b"bad instruction"),
    ("source", [0, 1, 2, 3], [X, Y, 0 | (3 << 8)], [1.0], b"constant index"),
    ("bc", [0, 1, 2, 3], [X, Y, ADD], [], b"stack underflow"),
    ("bc", [0, 1, 2, 4], [X, Y, X, Y], [], b"one value"),
    ("bc", [0, 1, 2, 20], [X, Y] + [T] * 17 + [SIN], [], b"DGX_FN_STACK"),
    ("bc", [0, 1, 2, 3], [X, Y, SIN | (1 << 8)], [], b"bad instruction"),
])
def test_functions_validated(lib, which, start, code, consts, msg):
    m = meshgen.box_quads(2, 2, bc=(2, None, 2, None))
    d, kd = _desc(m, 2, 30)
    keep = []
    good = _fn([0, 1, 2, 3], [X, Y, T], [], keep)
    bad = _fn(start, code, consts, keep)
    d.bc_function = bad if which == "bc" else good
    d.source = bad if which == "source" else None
    h = ctypes.c_void_p()
    assert lib.dgx_domain_create(ctypes.byref(d), None, ctypes.byref(h)) == -1
    assert msg in lib.dgx_last_error(), lib.dgx_last_error()
    assert not h.value


def test_functions_need_shallow_and_function_rows_need_a_function(lib):
    keep = []
    m = meshgen.periodic_quads(2, 2)
    d, k1 = _desc(m, 2, 1)
    d.source = _fn([0, 1, 2, 3], [X, Y, T], [], keep)
    h = ctypes.c_void_p()
    assert lib.dgx_domain_create(ctypes.byref(d), None, ctypes.byref(h)) == -1
    assert b"available for Shallow" in lib.dgx_last_error()
    m = meshgen.box_quads(2, 2, bc=(2, None, 0, None))
    d, k2 = _desc(m, 2, 30)
    assert lib.dgx_domain_create(ctypes.byref(d), None, ctypes.byref(h)) == -1
    assert b"without bc_function" in lib.dgx_last_error()


def test_time_entry_points_reject_null(lib):
    t = ctypes.c_double()
    assert lib.dgx_set_time(None, 0.0) == -1
    assert lib.dgx_get_time(None, ctypes.byref(t)) == -1


@pytest.mark.parametrize("args,msg", [((0, 3, 4, 4, 4), b"ne must be positive"),
                                      ((8, 0, 4, 4, 4), b"must be in 1..16"),
                                      ((8, 3, 17, 4, 4), b"must be in 1..16"),
                                      ((8, 5, 4, 6, 6), b"must not exceed qa"),
                                      ((8, 4, 6, 6, 3), b"must not exceed qa")])
def test_hexblock_create_rejects_bad_sizes(lib, args, msg):
    h = ctypes.c_void_p()
    assert lib.dgx_hexblock_create(*args, None, ctypes.byref(h)) == -1
    assert msg in lib.dgx_last_error()
    assert not h.value


def test_hexblock_entry_points_reject_null(lib):
    x = np.zeros(4)
    p = x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.dgx_hexblock_create(8, 3, 4, 4, 4, None, None) == -1
    assert lib.dgx_hexblock_set_wJ(None, p, 0) == -1
    assert lib.dgx_hexblock_set_basis(None, p, p, p) == -1
    assert lib.dgx_hexblock_inner_product(None, p, 0) == -1
    assert lib.dgx_hexblock_backward_transform(None, p, 1) == -1
    assert lib.dgx_hexblock_synchronize(None) == -1
    assert lib.dgx_hexblock_destroy(None) == 0
