"""Curved quads on the GPU (k_rhs2c: metrics at the nodes, mass-matrix
projection with each element's Cholesky factor, as the reference solves it)
against the oracle's restatement of Quad::compute_curved_geometry
(src/Quad.cpp:1309-1452).

Bar: relative L2 per field <= 1e-12 for one residual at every order (for
Navier-Stokes, or three times the residual's own rounding floor where that is
larger: at p = 10 a one-ulp perturbation of the input state moves the energy
residual by about 8e-13), <= 1e-11 after steps; the reference's cvquad run (2D Navier-Stokes on arc-sided quads)
within its own dgm_diff tolerances.
"""
import dataclasses

import numpy as np
import pytest

from dgm_amd import dgx, formats, mesh as meshgen
from fixture_runs import GOLDEN, initial_state, load_case
from oracle_lib import SHALLOW, Oracle, dgm_diff
from test_gpu_parity import TOL_RHS, TOL_STEP, projected_state, rel_l2

pytestmark = pytest.mark.gpu

GAMMA = 1.4
TWO_PI = 2 * np.pi


def _wavy(L, amp):
    def fn(x, y):
        s = lambda t: np.sin(2 * np.pi * t / L)
        return x + amp * s(y) * s(x), y + 0.7 * amp * s(x) * s(y)
    return fn


def _bump(x0, x1, H, amp):
    def fn(x, y):
        b = (1 - np.cos(2 * np.pi * (x - x0) / (x1 - x0))) / 2
        return x, y + amp * b * (H - y) / H
    return fn


def _euler_state(x):
    rho = 1.0 + 0.1 * np.sin(x[0]) * np.cos(x[1])
    u = 0.3 + 0.1 * np.sin(x[0] + x[1])
    v = -0.2 * np.cos(x[0]) * np.sin(x[1])
    p = 1.0 / GAMMA + 0.05 * np.cos(2 * x[0]) * np.cos(x[1])
    E = p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E])


def _shallow_state(x):
    phi = 1.0 + 0.08 * np.sin(x[0]) * np.cos(2 * x[1])
    return np.stack([phi, phi * (0.2 + 0.05 * np.cos(x[1])), phi * 0.05 * np.sin(x[0])])


@pytest.mark.parametrize("p,flux", [(2, 0), (5, 1), (8, 0), (10, 1)])
def test_rhs_euler2d_curved_periodic(p, flux):
    print(f"[curved2d] euler p={p} flux={flux}", flush=True)
    m = meshgen.periodic_quads(4, 3, 0.0, TWO_PI, 0.0, TWO_PI)
    xn = meshgen.mapped_nodes(m, p, _wavy(TWO_PI, 0.25))
    orc = Oracle(m, p, 1, flux=flux)
    orc.set_nodes(xn)
    Up = projected_state(orc, _euler_state(xn))
    F_ref = orc.rhs(Up)
    gpu = dgx.GPUDomain(m, p, dgx.DGX_EULER, flux=flux, xnodes=xn)
    assert np.allclose(gpu.coords(), orc.coords(), rtol=0, atol=1e-13)
    gpu.upload(Up)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"[curved2d] rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    gpu.close()


@pytest.mark.parametrize("p", [3, 6])
def test_rhs_shallow_curved_channel(p):
    """A channel whose bottom wall carries a bump: curved walls (slip) below
    and above, imposed states at the ends."""
    print(f"[curved2d] shallow p={p}", flush=True)
    H = 1.0
    m = meshgen.box_quads(4, 3, 0.0, 2.0, 0.0, H, bc=(0, 1, 0, 1))
    xn = meshgen.mapped_nodes(m, p, _bump(0.0, 2.0, H, 0.2))
    bc = [1.05, 0.21, 0.0]
    orc = Oracle(m, p, SHALLOW)
    orc.set_balance(False)
    orc.set_bc_state(bc)
    orc.set_nodes(xn)
    Up = projected_state(orc, _shallow_state(xn))
    F_ref = orc.rhs(Up)
    gpu = dgx.GPUDomain(m, p, dgx.DGX_SHALLOW, balance=False, bc_state=bc + [0.0, 0.0], xnodes=xn)
    gpu.upload(Up)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"[curved2d] rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    gpu.close()


def test_nodal_upload_projects_through_the_mass_matrix():
    p = 6
    m = meshgen.periodic_quads(3, 3, 0.0, 1.0, 0.0, 1.0)
    xn = meshgen.mapped_nodes(m, p, _wavy(1.0, 0.06))
    orc = Oracle(m, p, 1)
    orc.set_nodes(xn)
    rng = np.random.default_rng(8)
    Uh = rng.standard_normal((4, m.ne, (p + 1) ** 2))
    Uh[0, :, 0] += 5.0
    orc.set_modal(Uh)
    gpu = dgx.GPUDomain(m, p, dgx.DGX_EULER, xnodes=xn)
    gpu.upload(orc.state())
    back = gpu.download(dgx.DGX_MODAL)
    err = np.max(np.abs(back - Uh))
    print(f"[curved2d] projection max error {err:.3e}", flush=True)
    assert err < 1e-12, err
    gpu.close()


def test_advance_curved2d_lsrk14():
    p = 4
    m = meshgen.periodic_quads(3, 3, 0.0, TWO_PI, 0.0, TWO_PI)
    xn = meshgen.mapped_nodes(m, p, _wavy(TWO_PI, 0.2))
    orc = Oracle(m, p, 1)
    orc.set_nodes(xn)
    Up = projected_state(orc, _euler_state(xn))
    gpu = dgx.GPUDomain(m, p, dgx.DGX_EULER, xnodes=xn)
    gpu.upload(Up)
    orc.set_state(Up)
    orc.advance(14, 2e-3, 3)
    gpu.advance(14, 2e-3, 3)
    err = rel_l2(gpu.download(), orc.state())
    print(f"[curved2d] LSRK14 3 steps {err}", flush=True)
    assert np.all(err <= TOL_STEP), err
    gpu.close()


@pytest.mark.parametrize("p", [2, 5, 8, 10])
def test_rhs_ns2d_curved_periodic(p):
    """BR1 Navier-Stokes on curved quads: lifted gradients, viscous flux with
    each side's own normal, one projection of the sum."""
    print(f"[curved2d] navier-stokes p={p}", flush=True)
    vis = 0.05
    m = meshgen.periodic_quads(4, 3, 0.0, TWO_PI, 0.0, TWO_PI)
    xn = meshgen.mapped_nodes(m, p, _wavy(TWO_PI, 0.25))
    orc = Oracle(m, p, 2, vis=vis)
    orc.set_nodes(xn)
    Up = projected_state(orc, _euler_state(xn))
    F_ref = orc.rhs(Up)
    eul = Oracle(m, p, 1)
    eul.set_nodes(xn)
    visc = rel_l2(eul.rhs(Up), F_ref)
    assert np.all(visc[1:] > 1e-3), visc                 # the viscous terms are resolved by the bar
    # The residual's own rounding floor: its change when the input state is
    # perturbed by one ulp (random relative 2.2e-16), on the oracle.  At p = 10
    # the second derivatives of the viscous terms put it near 8e-13 for the
    # energy, so two FP64 evaluations in different operation orders can differ
    # by more than 1e-12 there; the bar is 1e-12 or three times that floor.
    rng = np.random.default_rng(p)
    floor = np.max([rel_l2(orc.rhs(Up * (1 + 2.2e-16 * rng.standard_normal(Up.shape))), F_ref)
                    for _ in range(2)], axis=0)
    bar = np.maximum(TOL_RHS, 3 * floor)
    gpu = dgx.GPUDomain(m, p, dgx.DGX_NAVIER_STOKES, vis=vis, xnodes=xn)
    gpu.upload(Up)
    err = rel_l2(gpu.rhs(), F_ref)
    print(f"[curved2d] rel L2 per field: {err}; rounding floor {floor}; bar {bar}", flush=True)
    assert np.all(err <= bar), err
    gpu.close()


def test_cvquad_golden_run():
    """runs/cvquad end to end on the GPU: 2D Navier-Stokes, TVD-RK3, p = 10,
    four quads with circular-arc sides, 1000 steps with the state projected at
    each restart write (every 500 steps), against cvquad.ref with the
    reference's own aeps = reps = 5e-13 (runs/cvquad.tst:22-23)."""
    prm, m, icd = load_case("cvquad")
    p = int(prm["p"])
    xn, _ = meshgen.curved_quad_nodes(m, p, formats.read_crv(f"{GOLDEN}/cvquad.crv"))
    gpu = dgx.GPUDomain(m, p, dgx.DGX_NAVIER_STOKES, vis=float(prm["vis"]), xnodes=xn)
    gpu.upload(initial_state(icd, gpu.coords(), 2))
    gpu.upload(gpu.download())                           # Domain::solve projects the initial state
    Nt, Ntout, dt, it = int(prm["Nt"]), int(prm["Ntout"]), float(prm["dt"]), int(prm["inttype"])
    done = 0
    while done < Nt:
        n = min(Ntout, Nt - done)
        gpu.advance(it, dt, n)
        done += n
        gpu.upload(gpu.download())                       # each restart write projects the state
        print(f"[curved2d] cvquad {done} steps", flush=True)
    Uh = gpu.download(dgx.DGX_MODAL)
    gpu.close()
    o = Oracle(m, p, 2, vis=float(prm["vis"]))
    o.set_nodes(xn)
    ref = formats.read_restart(f"{GOLDEN}/cvquad.ref", m.ne)
    rn, dn = dgm_diff(o, ref.data, Uh)
    print(f"[curved2d] cvquad dgm_diff absolute {dn.max():.3e}", flush=True)
    assert np.all(dn <= 5e-13), dn
    mask = rn > np.sqrt(np.finfo(float).eps)
    assert np.all(np.where(mask, dn / np.where(mask, rn, 1.0), 0.0) <= 5e-13)


def test_non_parallelogram_quads_take_the_curved_path():
    """Quads that are not parallelograms, given by their vertices only, are
    curved as in the reference (a straight-sided curve on the bilinear map):
    the residual and the norms match the oracle fed the bilinear nodes."""
    p = 4
    m = meshgen.periodic_quads(4, 4, 0.0, TWO_PI, 0.0, TWO_PI)
    v = np.array(m.vert, dtype=np.float64)
    # move interior vertices consistently (shared vertices move together)
    def shift(x, y):
        return x + 0.15 * np.sin(x) * np.sin(y), y
    v[:, 0, :], v[:, 1, :] = shift(v[:, 0, :], v[:, 1, :])
    m2 = dataclasses.replace(m, vert=v)
    assert not np.all(meshgen.affine_quads(m2))
    xn = meshgen.quad_nodes(m2, p)
    orc = Oracle(m2, p, 1)
    orc.set_nodes(xn)
    Up = projected_state(orc, _euler_state(xn))
    gpu = dgx.GPUDomain(m2, p, dgx.DGX_EULER)
    assert np.allclose(gpu.coords(), xn, rtol=0, atol=1e-13)
    gpu.upload(Up)
    err = rel_l2(gpu.rhs(), orc.rhs(Up))
    print(f"[curved2d] non-parallelogram rel L2 {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    orc.set_state(Up)
    assert np.allclose(gpu.norms(), orc.norms(), rtol=1e-12, atol=0)
    gpu.close()


def test_served_non_parallelogram_request_steps_on_its_own():
    """A request on non-parallelogram quads is curved, so the batching server
    steps it on its own; its result equals the direct domain's."""
    from dgm_amd import serve
    p, dt, n = 3, 1e-3, 10
    m = meshgen.periodic_quads(4, 4, 0.0, TWO_PI, 0.0, TWO_PI)
    v = np.array(m.vert, dtype=np.float64)
    v[:, 0, :] = v[:, 0, :] + 0.15 * np.sin(v[:, 0, :]) * np.sin(v[:, 1, :])
    m2 = dataclasses.replace(m, vert=v)
    xn = meshgen.quad_nodes(m2, p)
    U = _euler_state(xn)
    gpu = dgx.GPUDomain(m2, p, dgx.DGX_EULER)
    gpu.upload(U)
    gpu.advance(dgx.DGX_LSRK14, dt, n)
    want = gpu.download()
    gpu.close()
    srv = serve.Server(hbm_budget=1 << 34, max_active=2, steps_per_slice=5, batch=True)
    rid = srv.submit(m2, p, dgx.DGX_EULER, U, dgx.DGX_LSRK14, dt, n)
    srv.run_until_done(timeout_s=60.0)
    out = srv.result(rid)[0]
    srv.close()
    err = rel_l2(out, want)
    print(f"[curved2d] served non-parallelogram rel L2 {err}", flush=True)
    assert np.all(err <= 1e-13), err
