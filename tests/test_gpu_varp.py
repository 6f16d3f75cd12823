"""Variable-order quads on the GPU (dgx_varp2d.hip) against the oracle's
restatement (VarOracle: Field::Field's root.ord, Side::match's side sizing,
Quad::fill_sides and Quad::local_lift with the side's own Gauss points,
src/Field.cpp:2290-2320, src/Side.cpp:106-117, src/Quad.cpp:845-869,
1971-2038), the oracle itself pinned by runs/vortex.ref.

Bar: relative L2 per field <= 1e-12 for one residual, <= 1e-11 after steps.
Meshes: periodic and walled boxes, with element orders drawn at random so
that every pair of neighbouring orders (equal, lower, higher) occurs, on
rectangles and on sheared (parallelogram) quads; sides across the periodic
seam are flipped."""
import dataclasses

import numpy as np
import pytest

from dgm_amd import dgx, mesh as meshgen
from oracle_lib import VarOracle
from test_gpu_parity import TOL_RHS, TOL_STEP, rel_l2

pytestmark = pytest.mark.gpu


def _state(x, seed):
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(0.5, 1.5, 2)
    rho = 1.0 + 0.1 * np.sin(a * x[0]) * np.cos(b * x[1])
    u = 0.4 + 0.1 * np.sin(x[0] + x[1])
    v = -0.2 * np.cos(b * x[0]) * np.sin(a * x[1])
    p = 1.0 / 1.4 + 0.05 * np.cos(2 * x[0]) * np.cos(x[1])
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])


def _shear(mesh, s):
    vert = mesh.vert.copy()
    vert[:, 0, :] += s * vert[:, 1, :]
    return dataclasses.replace(mesh, vert=vert)


def _pair(mesh, orders, flux, seed, bc_state=None):
    orc = VarOracle(mesh, orders, flux=flux)
    if bc_state is not None:
        orc.set_bc_state(bc_state)
    x = orc.coords()
    orc.set_state(_state(x, seed))
    orc.project_state()
    Up = orc.state()
    gpu = dgx.GPUDomain(mesh, int(orders.max()), dgx.DGX_EULER, flux=flux, porder=orders, bc_state=bc_state)
    qmax = int(orders.max()) + 2
    for e in range(mesh.ne):   # the padded coordinate slots agree on each element's own nodes
        q2 = (int(orders[e]) + 2) ** 2
        assert np.allclose(gpu.coords()[:, e, :q2], x[:, e, :q2], rtol=0, atol=1e-13)
    assert Up.shape[2] == qmax * qmax
    gpu.upload(Up)
    return orc, gpu, Up


MESHES = {
    "periodic": lambda: meshgen.periodic_quads(5, 4, 0.0, 2 * np.pi, 0.0, 2 * np.pi),
    "sheared": lambda: _shear(meshgen.periodic_quads(4, 4, 0.0, 2 * np.pi, 0.0, 2 * np.pi), 0.3),
    "walls": lambda: meshgen.box_quads(4, 3, 0.0, 2 * np.pi, 0.0, 2 * np.pi, bc=(0, 1, 0, 1)),
}


@pytest.mark.parametrize("kind,flux,lo,hi", [("periodic", 0, 1, 6), ("periodic", 1, 2, 10),
                                             ("sheared", 0, 3, 8), ("walls", 0, 2, 7), ("walls", 1, 4, 9)])
def test_residual_matches_oracle(kind, flux, lo, hi):
    mesh = MESHES[kind]()
    rng = np.random.default_rng(lo * 10 + hi)
    orders = rng.integers(lo, hi + 1, mesh.ne).astype(np.int32)
    orders[0] = hi
    print(f"[varp] {kind} flux={flux} orders {orders.tolist()}", flush=True)
    bc = [1.0, 0.4, 0.0, 1.0 / 0.56 + 0.08] if kind == "walls" else None   # slip walls and a free stream
    orc, gpu, Up = _pair(mesh, orders, flux, seed=hi, bc_state=bc)
    F_ref = orc.rhs(Up)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"[varp] rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    gpu.close()


@pytest.mark.parametrize("inttype", [dgx.DGX_RK4, dgx.DGX_LSRK14, dgx.DGX_TVD])
def test_steps_match_oracle(inttype):
    mesh = meshgen.periodic_quads(4, 4, 0.0, 2 * np.pi, 0.0, 2 * np.pi)
    orders = np.array([2, 5, 3, 7] * 4, dtype=np.int32)
    print(f"[varp] steps inttype={inttype}", flush=True)
    orc, gpu, Up = _pair(mesh, orders, 0, seed=3)
    orc.advance(inttype, 2e-3, 3)
    gpu.advance(inttype, 2e-3, 3)
    err = rel_l2(gpu.download(), orc.state())
    print(f"[varp] 3 steps rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_STEP), err
    gpu.close()


def test_uniform_orders_match_the_uniform_path():
    """Every element at one order: the variable-order kernels and the uniform
    kernels compute the same residual (to rounding: the side tables differ)."""
    mesh = meshgen.periodic_quads(4, 3, 0.0, 2 * np.pi, 0.0, 2 * np.pi)
    orders = np.full(mesh.ne, 6, dtype=np.int32)
    gpu_v = dgx.GPUDomain(mesh, 6, dgx.DGX_EULER, porder=orders)
    gpu_u = dgx.GPUDomain(mesh, 6, dgx.DGX_EULER)
    x = gpu_u.coords()
    U = _state(x, 1)
    gpu_u.upload(U)
    gpu_v.upload(U)
    assert np.array_equal(gpu_u.download(dgx.DGX_MODAL), gpu_v.download(dgx.DGX_MODAL))
    err = rel_l2(gpu_v.rhs(), gpu_u.rhs())
    print(f"[varp] variable-order vs uniform kernels: {err}", flush=True)
    assert np.all(err <= 1e-13), err
    gpu_u.close()
    gpu_v.close()


def test_cfl_matches_oracle():
    """Euler::set_time_step over elements of different orders (each with its
    own GLL spacings), the allocating and the allocation-free entry points."""
    import ctypes
    mesh = meshgen.periodic_quads(5, 4, 0.0, 10.0, 0.0, 8.0)
    orders = np.random.default_rng(9).integers(2, 9, mesh.ne).astype(np.int32)
    orders[3] = 9
    orc, gpu, Up = _pair(mesh, orders, 0, seed=5)
    orc.set_state(Up)
    lib = gpu.lib
    fn = lib.dgx_set_time_step_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    for cflmax, dtmax in ((0.0, 0.0), (0.3, 0.0), (0.3, 1e-4)):
        want = orc.set_time_step(1e-3, cflmax, dtmax)
        got = gpu.set_time_step(1e-3, cflmax, dtmax)
        c, dn = ctypes.c_double(), ctypes.c_double()
        assert fn(gpu.h, 1e-3, cflmax, dtmax, ctypes.byref(c), ctypes.byref(dn)) == 0
        print(f"[varp] CFL oracle {want} GPU {got} allocation-free {(c.value, dn.value)}", flush=True)
        assert (c.value, dn.value) == got
        assert abs(got[0] - want[0]) <= 1e-13 * want[0] and abs(got[1] - want[1]) <= 1e-13 * want[1]
    gpu.close()
