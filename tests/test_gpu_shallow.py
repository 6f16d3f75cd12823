"""Shallow water on the GPU (k_shallow2 behind the C ABI, eqntype 30) against
the oracle's restatement of Shallow::time_derivative (src/Shallow.cpp:251-525).

Bar: relative L2 per field <= 1e-12 for one residual, <= 1e-11 after steps;
the balance term (first residual stored and subtracted) as the reference
compiles it; walls, state and function boundaries; function sources at each
stage's time; the non-realizable state error; the reference's shallow_mms run
within its own tolerances.
"""
import numpy as np
import pytest

from dgm_amd import dgx, formats, ic, mesh as meshgen
from fixture_runs import GOLDEN, case_setup, load_case, run_oracle_case
from oracle_lib import SHALLOW, Oracle, dgm_diff
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


def _tree(text):
    p = formats._ExprParser(formats._tokens(text))
    return p.expr()


def _block(*exprs):
    return {f"c{i}": _tree(e) for i, e in enumerate(exprs)}


def test_shallow_mms_golden_run():
    """runs/shallow_mms end to end on the GPU: 400 non-parallelogram quads
    (the curved path), p = 1, function boundaries and source, RK4, 3 steps with
    the state projected after each (Ntout = 1), against shallow_mms.ref with the
    reference's aeps = 1e-9, reps = 1e-10 and against the oracle's run."""
    prm, m, _ = load_case("shallow_mms")
    mesh, _, _, xn = case_setup("shallow_mms", prm, m)
    assert xn is not None
    bct = formats.read_functions(f"{GOLDEN}/shallow_mms.bct")["dansMMSFunction"]
    src = next(iter(formats.read_functions(f"{GOLDEN}/shallow_mms.src").values()))
    p, dt = int(prm["p"]), float(prm["dt"])
    gpu = dgx.GPUDomain(mesh, p, dgx.DGX_SHALLOW, xnodes=xn, bc_function=bct, source=src)
    x = gpu.coords()
    gpu.upload(ic.shallow_mms(x[0], x[1]))
    gpu.upload(gpu.download())
    for _ in range(int(prm["Nt"])):
        gpu.advance(int(prm["inttype"]), dt, 1)
        gpu.upload(gpu.download())
    assert abs(gpu.time - 3 * dt) < 1e-18
    Uh = gpu.download(dgx.DGX_MODAL)
    gpu.close()
    o, Uo = run_oracle_case("shallow_mms")
    err = rel_l2(Uh, Uo)
    print(f"shallow_mms GPU vs oracle rel L2 {err}", flush=True)
    assert np.all(err <= 1e-12), err
    ref = formats.read_restart(f"{GOLDEN}/shallow_mms.ref", mesh.ne)
    rn, dn = dgm_diff(o, ref.data, Uh)
    print(f"shallow_mms dgm_diff absolute {dn.max():.3e}", flush=True)
    assert np.all(dn <= 1e-9), dn
    assert np.all(dn <= 1e-10 * rn + 1e-300), dn / rn


@pytest.mark.parametrize("p", [2, 5])
def test_function_boundaries_on_an_affine_box(p):
    """Function boundaries and a source on a parallelogram mesh without nodal
    coordinates: the library takes the curved path on the bilinear map."""
    m = meshgen.box_quads(4, 3, 0.0, 2.0, 0.0, 1.5, bc=(2, 0, 2, 1))
    bfn = _block("1.0 + 0.05*sin(x + t)*cos(y)", "0.2 + 0.01*x*t", "-0.03*cos(y - t)")
    sfn = _block("0.1*x*y*t", "sin(x)*exp(-t)", "0.2 - y^2 + t")
    orc = Oracle(m, p, SHALLOW)
    orc.set_balance(False)
    orc.set_bc_state([1.02, 0.1, -0.02])
    orc.set_functions(bc=formats.field_function(bfn), src=formats.field_function(sfn))
    orc.set_time(0.3)
    x = orc.coords()
    U = projected_state(orc, _state(x))
    F_ref = orc.rhs(U)
    gpu = dgx.GPUDomain(m, p, dgx.DGX_SHALLOW, balance=False, bc_state=[1.02, 0.1, -0.02, 0.0, 0.0],
                        bc_function=bfn, source=sfn)
    assert np.allclose(gpu.coords(), x, rtol=0, atol=1e-13)
    gpu.set_time(0.3)
    gpu.upload(U)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"shallow function boundaries p={p} rel L2 {err}", flush=True)
    assert np.all(err <= 1e-12), err
    gpu.close()


@pytest.mark.parametrize("inttype,exact_degree", [(dgx.DGX_RK4, 3), (dgx.DGX_LSRK14, 3), (dgx.DGX_TVD, 2)])
def test_stage_times(inttype, exact_degree):
    """A uniform state forced by a source polynomial in t: dU/dt = s(t) is
    integrated exactly by an integrator whose stage times are right."""
    m = meshgen.periodic_quads(3, 3, 0.0, 10.0, 0.0, 10.0)    # CFL about 0.3 at dt = 0.05
    coef = [0.3, -0.8, 1.2, 0.9][:exact_degree + 1]
    poly = " + ".join(f"{c}*t^{k}" for k, c in enumerate(coef))
    src = _block(poly, "0.0*x", f"0.5*({poly})")              # F -= s and F = -dU/dt: dU/dt = s
    gpu = dgx.GPUDomain(m, 3, dgx.DGX_SHALLOW, balance=False, source=src)
    x = gpu.coords()
    U0 = np.stack([np.full_like(x[0], 1.0), np.full_like(x[0], 0.1), np.full_like(x[0], -0.2)])
    gpu.upload(U0)
    gpu.set_time(0.2)
    dt, n = 0.05, 4
    gpu.advance(inttype, dt, n)
    assert abs(gpu.time - (0.2 + n * dt)) < 1e-14
    integral = sum(c * ((0.2 + n * dt) ** (k + 1) - 0.2 ** (k + 1)) / (k + 1) for k, c in enumerate(coef))
    got = gpu.download()
    assert np.allclose(got[0], 1.0 + integral, rtol=0, atol=1e-13), np.abs(got[0] - 1.0 - integral).max()
    assert np.allclose(got[1], 0.1, rtol=0, atol=1e-13)
    assert np.allclose(got[2], -0.2 + 0.5 * integral, rtol=0, atol=1e-13)
    gpu.close()
    orc = Oracle(m, 3, SHALLOW)                          # the oracle's stage times agree
    orc.set_balance(False)
    orc.set_functions(src=formats.field_function(src))
    orc.set_state(U0)
    orc.set_time(0.2)
    orc.advance(inttype, dt, n)
    assert abs(orc.time() - (0.2 + n * dt)) < 1e-14
    assert np.all(rel_l2(got, orc.state()) <= 1e-13)


def test_balance_term_on_curved_quads():
    """The balance term also applies on the curved path (stored at the first
    residual, subtracted from every later one)."""
    p = 4
    m = meshgen.box_quads(4, 3, 0.0, 2.0, 0.0, 1.0, bc=(0, 1, 0, 1))
    xn = meshgen.mapped_nodes(m, p, lambda x, y: (x, y + 0.15 * np.sin(np.pi * x) * (1.0 - y)))
    bc = [1.05, 0.21, 0.0]
    orc = Oracle(m, p, SHALLOW)
    orc.set_bc_state(bc)
    orc.set_nodes(xn)
    x = orc.coords()
    U1 = projected_state(orc, _state(x, 0.0))
    U2 = projected_state(orc, _state(x, 0.7))
    gpu = dgx.GPUDomain(m, p, dgx.DGX_SHALLOW, bc_state=bc + [0.0, 0.0], xnodes=xn)
    gpu.upload(U1)
    assert np.max(np.abs(gpu.rhs())) == 0.0
    assert np.max(np.abs(orc.rhs(U1))) == 0.0
    gpu.upload(U2)
    err = rel_l2(gpu.rhs(), orc.rhs(U2))
    print(f"curved shallow balance rel L2 {err}", flush=True)
    assert np.all(err <= 1e-12), err
    orc.set_state(U2)
    orc.advance(dgx.DGX_LSRK14, 2e-3, 3)
    gpu.advance(dgx.DGX_LSRK14, 2e-3, 3)
    err = rel_l2(gpu.download(), orc.state())
    assert np.all(err <= 1e-11), err
    gpu.close()
