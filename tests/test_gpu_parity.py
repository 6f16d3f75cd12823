"""Parity of the HIP residual and steppers (libdgx.so) with the CPU oracle.

Bar: relative L2 per conserved field <= 1e-12 for one residual evaluation on
the same projected state (FP64 path), and <= 1e-11 after several RK steps.
"""
import numpy as np
import pytest

from dgm_amd import dgx, ic, mesh as meshgen
from fixture_runs import initial_state, load_case
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

TOL_RHS = 1e-12
TOL_STEP = 1e-11


def rel_l2(a, b):
    """Per-field relative L2 of nodal arrays [field][elem][node]."""
    num = np.sqrt(((a - b) ** 2).reshape(a.shape[0], -1).sum(1))
    den = np.sqrt((b ** 2).reshape(b.shape[0], -1).sum(1))
    return num / np.maximum(den, 1e-300)


def projected_state(orc, U):
    orc.set_state(U)
    orc.project_state()
    return orc.state()


def perturbed(U, x, amp=0.05):
    """Add a smooth O(amp) perturbation to every field (all residual fields nonzero)."""
    out = U.copy()
    for v in range(U.shape[0]):
        ph = 0.3 * (v + 1)
        w = np.sin(2.1 * x[0] + ph) * np.cos(1.7 * x[1] - ph)
        if len(x) == 3:
            w = w * np.cos(1.3 * x[2] + 0.5 * ph)
        out[v] = out[v] + amp * (1.0 + np.abs(U[v]).max()) * w
    return out


def check_rhs(m, p, eqn, U_fn, vis=0.0, flux=0, tol=TOL_RHS, viscous=0, br2_eta=6.0):
    orc = Oracle(m, p, eqn, vis=vis, flux=flux, viscous=viscous, br2_eta=br2_eta)
    x = orc.coords()
    Up = projected_state(orc, U_fn(x))
    F_ref = orc.rhs(Up)
    gpu = dgx.GPUDomain(m, p, eqn, vis=vis, flux=flux, viscous=viscous, br2_eta=br2_eta)
    gx = gpu.coords()
    assert np.allclose(gx, x, rtol=0, atol=1e-13)
    gpu.upload(Up)
    F = gpu.rhs()
    err = rel_l2(F, F_ref)
    print(f"eqn={eqn} p={p} ne={m.ne} rel L2 per field: {err}")
    assert np.all(err <= tol), err
    gpu.close()


def test_rhs_vhex_euler3d_p10():
    prm, m, icd = load_case("vhex")
    check_rhs(m, 10, 8, lambda x: perturbed(initial_state(icd, x, 3), x))


@pytest.mark.parametrize("p", [2, 4, 7])
def test_rhs_taylor_green_euler3d(p):
    m = meshgen.periodic_hexes(3, 3, 3, (0, 0, 0), (2 * np.pi,) * 3)
    check_rhs(m, p, 8, lambda x: perturbed(ic.taylor_green(x[0], x[1], x[2]), x))


@pytest.mark.parametrize("p", [3, 7])
def test_rhs_taylor_green_ns3d(p):
    m = meshgen.periodic_hexes(3, 3, 2, (0, 0, 0), (2 * np.pi,) * 3)
    check_rhs(m, p, 9, lambda x: perturbed(ic.taylor_green(x[0], x[1], x[2]), x), vis=1.0 / 100.0)


@pytest.mark.parametrize("p", [3, 7, 10])
def test_rhs_roe_euler3d(p):
    """3D Roe flux (fused kernels for p <= 9, unfused for p = 10)."""
    m = meshgen.periodic_hexes(3, 2, 2, (0, 0, 0), (2 * np.pi,) * 3)
    check_rhs(m, p, 8, lambda x: perturbed(ic.taylor_green(x[0], x[1], x[2]), x), flux=1)


def test_rhs_roe_ns3d():
    m = meshgen.periodic_hexes(3, 3, 2, (0, 0, 0), (2 * np.pi,) * 3)
    check_rhs(m, 4, 9, lambda x: perturbed(ic.taylor_green(x[0], x[1], x[2]), x), vis=0.01, flux=1)


@pytest.mark.parametrize("p,flux,eta", [(2, 0, 6.0), (3, 1, 6.0), (5, 0, 2.5), (7, 1, 6.0), (8, 0, 6.0),
                                        (9, 1, 20.0)])
def test_rhs_br2_ns3d(p, flux, eta):
    """BR2 viscous face gradient (no reference counterpart: the oracle's BR2 is
    pinned by tests/test_oracle_br2.py); anisotropic elements."""
    print(f"[br2] p={p} flux={flux} eta={eta}", flush=True)
    m = meshgen.periodic_hexes(3, 3, 2, (0, 0, 0), (2 * np.pi, 2 * np.pi, 4.0))
    check_rhs(m, p, 9, lambda x: perturbed(ic.taylor_green(x[0], x[1], x[2]), x), vis=0.02, flux=flux,
              viscous=1, br2_eta=eta)


def test_rhs_br2_ns3d_vhex_mesh():
    """BR2 on the vhex fixture's mesh (general side links and orientations)."""
    prm, m, icd = load_case("vhex")
    check_rhs(m, 4, 9, lambda x: perturbed(initial_state(icd, x, 3), x), vis=0.01, viscous=1)


def test_rhs_vquad_ns2d_p10():
    prm, m, icd = load_case("vquad")
    check_rhs(m, 10, 2, lambda x: perturbed(initial_state(icd, x, 2), x), vis=0.02)


@pytest.mark.parametrize("flux", [0, 1])
def test_rhs_isentropic_vortex_euler2d(flux):
    m = meshgen.periodic_quads(6, 5, 0, 10, 0, 10)
    check_rhs(m, 4, 1, lambda x: ic.isentropic_vortex(x[0], x[1], 4.0, 6.0, period=(10, 10)),
              flux=flux)


def test_advance_vquad_tvd():
    prm, m, icd = load_case("vquad")
    orc = Oracle(m, 10, 2, vis=0.02)
    x = orc.coords()
    Up = projected_state(orc, initial_state(icd, x, 2))
    gpu = dgx.GPUDomain(m, 10, 2, vis=0.02)
    gpu.upload(Up)
    orc.set_state(Up)
    orc.advance(3, 1e-3, 20)
    gpu.advance(dgx.DGX_TVD, 1e-3, 20)
    gpu.synchronize()
    err = rel_l2(gpu.download(), orc.state())
    print("vquad TVD 20 steps", err)
    assert np.all(err <= TOL_STEP), err


@pytest.mark.parametrize("inttype", [4, 14])
def test_advance_tgv_ns3d(inttype):
    m = meshgen.periodic_hexes(2, 2, 2, (0, 0, 0), (2 * np.pi,) * 3)
    orc = Oracle(m, 4, 9, vis=1.0 / 1600)
    x = orc.coords()
    Up = projected_state(orc, perturbed(ic.taylor_green(x[0], x[1], x[2]), x))
    gpu = dgx.GPUDomain(m, 4, 9, vis=1.0 / 1600)
    gpu.upload(Up)
    orc.set_state(Up)
    orc.advance(inttype, 2e-3, 3)
    gpu.advance(inttype, 2e-3, 3)
    gpu.synchronize()
    err = rel_l2(gpu.download(), orc.state())
    print(f"TGV NS3d inttype={inttype} 3 steps", err)
    assert np.all(err <= TOL_STEP), err


def _channel_state(x, H):
    """Smooth state at Mach ~0.3 (so that the momentum residual is not a small
    difference of O(1/Ma^2) pressure fluxes) with nonzero wall-normal momentum
    at the walls."""
    rho = 1.0 + 0.1 * np.sin(x[0]) * np.cos(x[1])
    u = 0.3 * np.sin(x[0]) * np.cos(x[1]) * np.cos(np.pi * x[2] / H)
    v = -0.3 * np.cos(x[0]) * np.sin(x[1])
    w = 0.1 * np.cos(np.pi * x[2] / H) * np.sin(x[0]) + 0.05 * x[2] / H
    p = 1.0 + 0.1 * np.cos(2 * x[0]) * np.cos(x[2])
    E = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E])


@pytest.mark.parametrize("p,flux,bcs", [(3, 0, (0, 1)), (7, 1, (0, 0)), (5, 1, (1, 0)), (10, 0, (0, 1))])
def test_rhs_euler3d_walls_and_state_boundaries(p, flux, bcs):
    """Slip walls and imposed-state boundaries on z = 0 and z = H (fused kernels
    for p <= 9, unfused at p = 10)."""
    print(f"[bc] p={p} flux={flux} bc={bcs}", flush=True)
    H = 2.0
    m = meshgen.channel_hexes(3, 2, 3, (0, 0, 0), (2 * np.pi, 2 * np.pi, H), bc_bottom=bcs[0], bc_top=bcs[1])
    bc_state = [1.05, 0.1, -0.05, 0.02, 1.0 / 0.4 + 0.01]
    orc = Oracle(m, p, 8, flux=flux)
    orc.set_bc_state(bc_state)
    x = orc.coords()
    Up = projected_state(orc, _channel_state(x, H))
    F_ref = orc.rhs(Up)
    gpu = dgx.GPUDomain(m, p, 8, flux=flux, bc_state=bc_state)
    gpu.upload(Up)
    F = gpu.rhs()
    err = rel_l2(F, F_ref)
    print(f"[bc] rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    gpu.close()


def test_advance_euler3d_channel_lsrk14():
    H = 2.0
    m = meshgen.channel_hexes(2, 2, 3, (0, 0, 0), (2 * np.pi, 2 * np.pi, H), bc_bottom=0, bc_top=1)
    bc_state = [1.0, 0.0, 0.0, 0.0, 1.0 / 0.4]
    orc = Oracle(m, 4, 8)
    orc.set_bc_state(bc_state)
    x = orc.coords()
    Up = projected_state(orc, _channel_state(x, H))
    gpu = dgx.GPUDomain(m, 4, 8, bc_state=bc_state)
    gpu.upload(Up)
    orc.set_state(Up)
    orc.advance(14, 2e-3, 3)
    gpu.advance(14, 2e-3, 3)
    gpu.synchronize()
    err = rel_l2(gpu.download(), orc.state())
    print("channel LSRK14 3 steps", err)
    assert np.all(err <= TOL_STEP), err


@pytest.mark.parametrize("p,flux", [(3, 0), (5, 1), (6, 0), (7, 1), (8, 0)])
def test_rhs_euler3d_curved_bump_channel(p, flux):
    """Curved hexes: a channel with a bump on its bottom wall (slip wall) and
    an imposed state on top; metrics at the nodes, mass-matrix projection."""
    print(f"[curved] bump p={p} flux={flux}", flush=True)
    lo, hi = (0.0, 0.0, 0.0), (2.0, 2.0, 1.0)
    m = meshgen.channel_hexes(3, 2, 2, lo, hi, bc_bottom=0, bc_top=1)
    xn = meshgen.mapped_nodes(m, p, meshgen.bump_map(lo, hi, 0.15))
    bc_state = [1.0, 0.3, 0.0, 0.0, 1.0 / 0.4 + 0.045]
    orc = Oracle(m, p, 8, flux=flux)
    orc.set_bc_state(bc_state)
    orc.set_nodes(xn)
    Up = projected_state(orc, _channel_state(xn, 1.0))
    F_ref = orc.rhs(Up)
    gpu = dgx.GPUDomain(m, p, 8, flux=flux, bc_state=bc_state, xnodes=xn)
    assert np.allclose(gpu.coords(), orc.coords(), rtol=0, atol=1e-13)
    gpu.upload(Up)
    F = gpu.rhs()
    err = rel_l2(F, F_ref)
    print(f"[curved] rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    gpu.close()


def test_advance_euler3d_curved_lsrk14():
    lo, hi = (0.0, 0.0, 0.0), (2.0, 2.0, 1.0)
    m = meshgen.channel_hexes(2, 2, 2, lo, hi, bc_bottom=0, bc_top=1)
    p = 4
    xn = meshgen.mapped_nodes(m, p, meshgen.bump_map(lo, hi, 0.1))
    bc_state = [1.0, 0.3, 0.0, 0.0, 1.0 / 0.4 + 0.045]
    orc = Oracle(m, p, 8)
    orc.set_bc_state(bc_state)
    orc.set_nodes(xn)
    Up = projected_state(orc, _channel_state(xn, 1.0))
    gpu = dgx.GPUDomain(m, p, 8, bc_state=bc_state, xnodes=xn)
    gpu.upload(Up)
    orc.set_state(Up)
    orc.advance(14, 1e-3, 3)
    gpu.advance(14, 1e-3, 3)
    gpu.synchronize()
    err = rel_l2(gpu.download(), orc.state())
    print("curved LSRK14 3 steps", err)
    assert np.all(err <= TOL_STEP), err
