"""Navier-Stokes 3D (BR1) on curved hexes on the GPU against the oracle.

The oracle's curved Navier-Stokes path (nodal metrics in the gradients and
divergences, each side's own normal in the lifted jumps and the side fluxes,
the mass-matrix projection of the lifted gradients and of the whole residual;
src/Navier_Stokes3d.cpp:113-258, src/Euler3d.cpp:1144-1199,
src/Hex.cpp:3992-4080) is pinned in tests/test_oracle_curved.py: affine nodes
through it reproduce the golden-pinned affine residual, and on a curved mesh
the residual converges spectrally to the exact divergence of the inviscid and
viscous fluxes.

Bar: relative L2 per field <= 1e-12 for one residual, <= 1e-11 after steps.
"""
import numpy as np
import pytest

from dgm_amd import dgx, mesh as meshgen
from oracle_lib import Oracle
from test_oracle_curved import wavy_map

pytestmark = pytest.mark.gpu

TOL_RHS = 1e-12
TOL_STEP = 1e-11
WALL = (1.0, 3.0)      # Tw, Eref
BC_STATE = [1.05, 0.25, -0.05, 0.0, 1.0 / 0.4 + 0.04]


def rel_l2(a, b):
    num = np.sqrt(((a - b) ** 2).reshape(a.shape[0], -1).sum(1))
    den = np.sqrt((b ** 2).reshape(b.shape[0], -1).sum(1))
    return num / np.maximum(den, 1e-300)


def _state(x, H):
    rho = 1.0 + 0.1 * np.sin(x[0]) * np.cos(x[1]) * np.cos(x[2])
    u = 0.3 * np.sin(x[0]) * np.cos(x[1]) * np.sin(np.pi * x[2] / H) + 0.2
    v = -0.3 * np.cos(x[0]) * np.sin(x[1]) * np.sin(np.pi * x[2] / H)
    w = 0.05 * np.sin(x[0]) * np.sin(np.pi * x[2] / H) + 0.02 * x[2] / H
    p = 1.0 + 0.1 * np.cos(2 * x[0]) * np.cos(x[2])
    E = p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)
    return np.stack([rho, rho * u, rho * v, rho * w, E])


def _bump(p, bcs, amp=0.15, H=2.0, vis=0.02):
    lo, hi = (0.0, 0.0, 0.0), (2 * np.pi, 2 * np.pi, H)
    m = meshgen.channel_hexes(3, 2, 2, lo, hi, bc_bottom=bcs[0], bc_top=bcs[1])
    xn = meshgen.mapped_nodes(m, p, meshgen.bump_map(lo, hi, amp))
    orc = Oracle(m, p, 9, vis=vis)
    orc.set_isothermal(*WALL)
    orc.set_bc_state(BC_STATE)
    orc.set_nodes(xn)
    orc.set_state(_state(xn, H))
    orc.project_state()
    Up = orc.state()
    gpu = dgx.GPUDomain(m, p, 9, vis=vis, bc_state=BC_STATE, wall=WALL, xnodes=xn)
    assert np.allclose(gpu.coords(), orc.coords(), rtol=0, atol=1e-13)
    gpu.upload(Up)
    return orc, gpu, Up


@pytest.mark.parametrize("p,bcs", [(1, (3, 1)), (2, (3, 3)), (3, (3, 1)), (5, (1, 3)), (6, (3, 3)), (7, (3, 1))])
def test_rhs_ns3d_curved_bump_channel(p, bcs):
    """A channel whose bottom wall carries a bump: isothermal walls or an
    imposed state on the bottom and top, periodic sides."""
    print(f"[curved-ns] bump p={p} bc={bcs}", flush=True)
    orc, gpu, Up = _bump(p, bcs)
    F_ref = orc.rhs(Up)
    F = gpu.rhs()
    err = rel_l2(F, F_ref)
    print(f"[curved-ns] rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    gpu.close()


@pytest.mark.parametrize("p", [2, 4])
def test_rhs_ns3d_curved_periodic(p):
    """Every element curved (a periodic box deformed inside), no boundaries."""
    print(f"[curved-ns] wavy periodic p={p}", flush=True)
    L = 2 * np.pi
    m = meshgen.periodic_hexes(3, 2, 2, (0, 0, 0), (L, L, L))
    xn = meshgen.mapped_nodes(m, p, wavy_map(L, 0.2))
    orc = Oracle(m, p, 9, vis=0.05)
    orc.set_nodes(xn)
    orc.set_state(_state(xn, L))
    orc.project_state()
    Up = orc.state()
    gpu = dgx.GPUDomain(m, p, 9, vis=0.05, xnodes=xn)
    gpu.upload(Up)
    err = rel_l2(gpu.rhs(), orc.rhs(Up))
    print(f"[curved-ns] rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_RHS), err
    gpu.close()


def test_advance_ns3d_curved_lsrk14():
    """Three LSRK14 steps (the fused curved stage with the viscous kernels before it)."""
    print("[curved-ns] LSRK14 3 steps p=4", flush=True)
    orc, gpu, Up = _bump(4, (3, 1), amp=0.1)
    orc.set_state(Up)
    orc.advance(14, 1e-3, 3)
    gpu.advance(14, 1e-3, 3)
    gpu.synchronize()
    err = rel_l2(gpu.download(), orc.state())
    print(f"[curved-ns] LSRK14 rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_STEP), err
    gpu.close()


def test_rk4_ns3d_curved_matches_residual_path():
    """RK4 runs the residual (k_grad3c, k_visc3c, k_rhs3c) and the BLAS
    updates; it must agree with the oracle too."""
    orc, gpu, Up = _bump(3, (3, 3), amp=0.1)
    orc.set_state(Up)
    orc.advance(4, 1e-3, 2)
    gpu.advance(4, 1e-3, 2)
    gpu.synchronize()
    err = rel_l2(gpu.download(), orc.state())
    print(f"[curved-ns] RK4 rel L2 per field: {err}", flush=True)
    assert np.all(err <= TOL_STEP), err
    gpu.close()
