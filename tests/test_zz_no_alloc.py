"""The allocation-free CFL entry point (dgx_set_time_step_noalloc) against
dgx_set_time_step through the C ABI: the same kernels on a file-scope device
scratch array instead of a library-allocated buffer, so both must agree
bitwise, on two calls in a row, the second on a smaller domain (the array
keeps the first call's contents), and sizes beyond the array are refused."""
import ctypes

import numpy as np
import pytest

from dgm_amd import dgx, ic, mesh as meshgen

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    fn = lib.dgx_set_time_step_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    fn.restype = ctypes.c_int
    return fn


def _vortex_domain(n, p, eqn=1):
    m = meshgen.periodic_quads(n, n, 0.0, 10.0, 0.0, 10.0)
    d = dgx.GPUDomain(m, p, eqn, balance=False)
    x = d.coords()
    if eqn == 30:
        U = np.stack([1.0 + 0.05 * np.sin(x[0]) * np.cos(x[1]), 0.1 * np.cos(x[1]), 0.05 * np.sin(x[0])])
    else:
        U = ic.isentropic_vortex(x[0], x[1], 3.3, 6.1)
    d.upload(U)
    return d


def test_noalloc_matches_the_allocating_entry_point():
    lib = dgx.load()
    fn = _noalloc(lib)
    for n, p, eqn in ((32, 4, 1), (6, 3, 30)):        # second call: a smaller domain
        print(f"[noalloc] {n}x{n} quads p={p} eqn={eqn}", flush=True)
        d = _vortex_domain(n, p, eqn)
        for cflmax, dtmax in ((0.0, 0.0), (0.2, 0.0), (0.2, 1e-4)):
            want = d.set_time_step(1e-3, cflmax, dtmax)
            c, dn = ctypes.c_double(), ctypes.c_double()
            assert fn(d.h, 1e-3, cflmax, dtmax, ctypes.byref(c), ctypes.byref(dn)) == 0, \
                lib.dgx_last_error().decode()
            print(f"[noalloc] allocating {want}, allocation-free {(c.value, dn.value)}", flush=True)
            assert (c.value, dn.value) == want
        d.close()


def test_noalloc_refuses_domains_beyond_its_scratch():
    lib = dgx.load()
    fn = _noalloc(lib)
    m = meshgen.periodic_quads(725, 725, 0.0, 1.0, 0.0, 1.0)   # 525,625 elements > 524,287
    d = dgx.GPUDomain(m, 1, 1)
    c, dn = ctypes.c_double(), ctypes.c_double()
    assert fn(d.h, 1e-3, 0.0, 0.0, ctypes.byref(c), ctypes.byref(dn)) == -1   # DGX_ERR_ARGUMENT
    assert b"exceed its scratch" in lib.dgx_last_error()
    d.close()
