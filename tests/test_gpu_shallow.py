"""Shallow water on the GPU (k_shallow2 behind the C ABI, eqntype 30) against
the oracle's restatement of Shallow::time_derivative (src/Shallow.cpp:251-525).

Bar: relative L2 per field <= 1e-12 for one residual, <= 1e-11 after steps;
the balance term (first residual stored and subtracted) as the reference
compiles it; walls and state boundaries; the non-realizable state error.
"""
import numpy as np
import pytest

from dgm_amd import dgx, mesh as meshgen
from oracle_lib import SHALLOW, Oracle
from test_gpu_parity import projected_state, rel_l2

pytestmark = pytest.mark.gpu

TWO_PI = 2 * np.pi


def _state(x, t=0.0, amp=0.08):
    phi = 1.0 + amp * np.sin(x[0] + 0.4 * t) * np.cos(x[1]) + 0.5 * amp * np.cos(2 * x[1] - t)
    u = 0.2 * np.cos(x[1]) + amp * np.sin(x[0] + t)
    v = amp * np.sin(x[0] + x[1]) - 0.1
    return np.stack([phi, phi * u, phi * v])


def _pair(m, p, balance=False, bc_state=None):
    orc = Oracle(m, p, SHALLOW)
    orc.set_balance(balance)
    if bc_state is not None:
        orc.set_bc_state(bc_state)
    gpu = dgx.GPUDomain(m, p, dgx.DGX_SHALLOW, balance=balance,
                        bc_state=None if bc_state is None else list(bc_state) + [0.0, 0.0])
    assert gpu.nf == 3
    return orc, gpu


@pytest.mark.parametrize("p", [1, 3, 6, 10])
def test_rhs_periodic(p):
    m = meshgen.periodic_quads(4, 3, 0.0, TWO_PI, 0.0, TWO_PI)
    orc, gpu = _pair(m, p)
    x = orc.coords()
    assert np.allclose(gpu.coords(), x, rtol=0, atol=1e-13)
    U = projected_state(orc, _state(x))
    F_ref = orc.rhs(U)
    gpu.upload(U)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"shallow periodic p={p} rel L2 {err}", flush=True)
    assert np.all(err <= 1e-12), err
    gpu.close()


@pytest.mark.parametrize("p", [2, 5])
def test_rhs_walls_and_state_boundaries(p):
    m = meshgen.box_quads(4, 4, 0.0, 2.0, 0.0, 1.5, bc=(0, 1, 0, 1))
    orc, gpu = _pair(m, p, bc_state=[1.05, 0.2, -0.05])
    x = orc.coords()
    U = projected_state(orc, _state(x))
    F_ref = orc.rhs(U)
    gpu.upload(U)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"shallow box p={p} rel L2 {err}", flush=True)
    assert np.all(err <= 1e-12), err
    gpu.close()


def test_balance_term():
    m = meshgen.periodic_quads(3, 3, 0.0, TWO_PI, 0.0, TWO_PI)
    orc, gpu = _pair(m, 4, balance=True)
    x = orc.coords()
    U1 = projected_state(orc, _state(x, 0.0))
    U2 = projected_state(orc, _state(x, 0.7))
    gpu.upload(U1)
    assert np.max(np.abs(gpu.rhs())) == 0.0              # the first residual is stored and subtracted
    assert np.max(np.abs(orc.rhs(U1))) == 0.0
    gpu.upload(U2)
    F_ref = orc.rhs(U2)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"shallow balance rel L2 {err}", flush=True)
    assert np.all(err <= 1e-12), err
    gpu.close()


@pytest.mark.parametrize("inttype,balance", [(dgx.DGX_LSRK14, False), (dgx.DGX_RK4, False), (dgx.DGX_LSRK14, True)])
def test_steps(inttype, balance):
    m = meshgen.box_quads(4, 3, 0.0, TWO_PI, 0.0, TWO_PI, bc=(0, None, 0, None))
    orc, gpu = _pair(m, 5, balance=balance)
    x = orc.coords()
    U = projected_state(orc, _state(x))
    orc.set_state(U)
    gpu.upload(U)
    dt, n = 2e-3, 8
    orc.advance(inttype, dt, n)
    gpu.advance(inttype, dt, n)
    ref = orc.state()
    err = rel_l2(gpu.download(), ref)
    print(f"shallow steps inttype={inttype} balance={balance} rel L2 {err}", flush=True)
    assert np.all(err <= 1e-11), err
    if balance:                                          # the initial state is then a fixed point
        assert np.max(np.abs(ref - U)) < 1e-12
    else:
        assert np.max(np.abs(ref - U)) > 1e-4
    gpu.close()


def test_nonrealizable_state_is_reported():
    m = meshgen.periodic_quads(2, 2)
    gpu = dgx.GPUDomain(m, 2, dgx.DGX_SHALLOW, balance=False)
    x = gpu.coords()
    U = _state(x)
    U[0] -= 2.0                                          # negative geopotential
    gpu.upload(U)
    with pytest.raises(dgx.DgxError) as ei:
        gpu.rhs()
    assert ei.value.code == -3
    gpu.close()
