"""Discrete adjoint of the nonlinear 2D Euler equations on the GPU
(dgx_euler_adjoint_*: the backward kernels of dgx_adjoint2d.hip driven by
LSRK14::adjoint_advance's recurrence, src/TimeInt.cpp:2064-2092).

The oracle's forward residual and time stepping (Euler::convection with the
Lax-Friedrichs or Roe flux, src/Euler.cpp:296-757, pinned by the golden runs
and the Roe flux's closed-form checks) are the checker; the transposed Jacobian has no counterpart in the reference to
compare with, so it is pinned by the mathematics:

* every entry of (dR/dU)^T, one vjp per unit vector, against central
  differences of the oracle's residual on a small mesh (1e-7 of the largest
  entry; the differences carry O(eps^2) truncation);
* <vjp(lam), v> against <lam, dR/dU v> by central differences of the oracle
  residual along smooth seeded directions on larger meshes (1e-9 of
  |lam| |dR/dU v|;
  the directions are smooth so that the differences do not straddle the kinks
  of the Lax-Friedrichs speed, |u.n| and max(lambda_L, lambda_R), or of the
  Roe flux's upwinded wave speeds min(u.n - c, 0), min(u.n, 0), min(u.n + c, 0),
  where the transposed Jacobian takes one side's derivative; the test state
  keeps every face's normal velocity away from zero);
* the forward observations against the oracle's LSRK14 solution, read at GLL
  nodes chosen as receivers (1e-12), and the final state against dgx_advance
  (bitwise: both run the fused stage kernel);
* dJ/dU0 against central differences of the oracle's objective along smooth
  seeded directions (1e-6 relative: second-order differences at eps = 1e-5 of a
  smooth objective), and, with the data the forward solve itself produced,
  J = 0 and dJ/dU0 = 0 exactly;
* the checkpoint count changes the recomputation schedule, not the result:
  gradients with 2, 3 and the default number of slots are bitwise equal.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from dgm_amd import dgx, euler_adjoint, mesh as meshgen
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def _state(x, seed):
    """A smooth periodic state whose velocity (u >= 0.3, v >= 0.05) keeps every
    face's normal velocity away from zero, where |u.n| in the Lax-Friedrichs
    speed has its kink."""
    rng = np.random.default_rng(seed)
    a, b = rng.integers(1, 3, 2)
    rho = 1.0 + 0.1 * np.sin(a * x[0]) * np.cos(b * x[1])
    u = 0.4 + 0.1 * np.sin(x[0] + x[1])
    v = 0.25 - 0.2 * np.cos(b * x[0]) * np.sin(a * x[1])
    p = 1.0 / GAMMA + 0.05 * np.cos(2 * x[0]) * np.cos(x[1])
    return np.stack([rho, rho * u, rho * v, p / (GAMMA - 1) + 0.5 * rho * (u * u + v * v)])


def _direction(orc, shear, seed):
    """A smooth periodic perturbation direction (modal): its jumps across faces
    are projection errors, so differences along it do not cross the kink of
    the flux's max(lambda_L, lambda_R) where the two speeds nearly agree."""
    rng = np.random.default_rng(seed)
    x = _logical(orc.coords(), shear)
    scale = (1.0, 0.4, 0.25, 2.5)
    f = np.stack([sum(scale[c] * rng.uniform(0.2, 1.0) *
                      np.sin(rng.integers(0, 3) * x[0] + rng.integers(1, 3) * x[1] + rng.uniform(0, 2 * np.pi))
                      for _ in range(3)) for c in range(4)])
    orc.set_state(f)
    orc.project_state()
    return orc.modal()


def _logical(x, shear):
    """Coordinates of the unsheared box, in which the periodic mesh repeats."""
    return np.stack([x[0] - shear * x[1], x[1]])


def _shear(mesh, s):
    vert = mesh.vert.copy()
    vert[:, 0, :] += s * vert[:, 1, :]
    return dataclasses.replace(mesh, vert=vert)


def _oracle(mesh, p, seed, shear=0.0, flux=dgx.DGX_FLUX_LF):
    orc = Oracle(mesh, p, dgx.DGX_EULER, gamma=GAMMA, flux=flux)
    orc.set_state(_state(_logical(orc.coords(), shear), seed))
    orc.project_state()
    return orc, orc.modal()


def _orc_rhs(orc, Uh):
    """Modal residual R = -dU/dt of the oracle at modal state Uh."""
    orc.set_modal(Uh)
    F = orc.rhs(orc.state())
    orc.set_state(F)
    return orc.modal()


def _pressure(U):
    return (GAMMA - 1) * (U[3] - 0.5 * (U[1] ** 2 + U[2] ** 2) / U[0])


@pytest.mark.parametrize("flux", [dgx.DGX_FLUX_LF, dgx.DGX_FLUX_ROE])
def test_transposed_jacobian_every_entry(flux):
    """(dR/dU)^T column by column on a sheared 3 x 3 mesh at p = 2."""
    mesh = _shear(meshgen.periodic_quads(3, 3, 0.0, 2 * np.pi, 0.0, 2 * np.pi), 0.25)
    p = 2
    orc, Uh = _oracle(mesh, p, seed=1, shear=0.25, flux=flux)
    prob = euler_adjoint.EulerAdjointProblem(mesh, p, [], 1e-3, 1, flux=flux)
    n = Uh.size
    print(f"[euler-adjoint] full Jacobian, {n} unknowns", flush=True)
    JT = np.zeros((n, n))
    for i in range(n):
        lam = np.zeros(n)
        lam[i] = 1.0
        JT[i] = prob.vjp(Uh, lam.reshape(Uh.shape)).ravel()   # (dR/dU)^T e_i: dR_i/dU_j over j
    eps = 1e-6
    Jfd = np.zeros((n, n))
    for j in range(n):
        dv = np.zeros(n)
        dv[j] = eps
        dv = dv.reshape(Uh.shape)
        Jfd[:, j] = ((_orc_rhs(orc, Uh + dv) - _orc_rhs(orc, Uh - dv)) / (2 * eps)).ravel()
    err = np.max(np.abs(JT - Jfd)) / np.max(np.abs(Jfd))
    print(f"[euler-adjoint] max |J^T - J_fd| / max |J| = {err:.2e}", flush=True)
    assert err <= 1e-7, err
    prob.close()


@pytest.mark.parametrize("nx,p,shear,flux", [(4, 3, 0.0, dgx.DGX_FLUX_LF), (5, 5, 0.3, dgx.DGX_FLUX_LF),
                                             (3, 8, 0.0, dgx.DGX_FLUX_LF), (4, 3, 0.0, dgx.DGX_FLUX_ROE),
                                             (5, 5, 0.3, dgx.DGX_FLUX_ROE)])
def test_vjp_against_residual_differences(nx, p, shear, flux):
    mesh = _shear(meshgen.periodic_quads(nx, nx, 0.0, 2 * np.pi, 0.0, 2 * np.pi), shear)
    orc, Uh = _oracle(mesh, p, seed=nx + p, shear=shear, flux=flux)
    prob = euler_adjoint.EulerAdjointProblem(mesh, p, [], 1e-3, 1, flux=flux)
    rng = np.random.default_rng(p)
    eps = 1e-6
    for t in range(3):
        lam = rng.standard_normal(Uh.shape)
        v = _direction(orc, shear, 100 * p + t)
        lhs = float(np.sum(prob.vjp(Uh, lam) * v))
        dR = (_orc_rhs(orc, Uh + eps * v) - _orc_rhs(orc, Uh - eps * v)) / (2 * eps)
        rhs = float(np.sum(lam * dR))
        # relative to |lam| |dR v|, the scale of the inner product's rounding and
        # difference errors (a random lam can be nearly orthogonal to dR v)
        rel = abs(lhs - rhs) / (np.linalg.norm(lam) * np.linalg.norm(dR))
        print(f"[euler-adjoint] nx={nx} p={p} shear={shear} flux={flux} draw {t}: <vjp, v> {lhs:.12e} "
              f"<lam, dR v> {rhs:.12e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-9, rel
    prob.close()


def _setup(nx=4, p=3, nsteps=20, dt=0.01, nrec=5, nsnap=0, seed=3, flux=dgx.DGX_FLUX_LF):
    mesh = meshgen.periodic_quads(nx, nx, 0.0, 2 * np.pi, 0.0, 2 * np.pi)
    orc, Uh = _oracle(mesh, p, seed, flux=flux)
    x = orc.coords()
    q = p + 2
    rng = np.random.default_rng(seed)
    interior = [i * q + j for i in range(1, q - 1) for j in range(1, q - 1)]
    rec = [(int(rng.integers(mesh.ne)), int(rng.choice(interior))) for _ in range(nrec)]
    pts = np.array([[x[0, e, k], x[1, e, k]] for e, k in rec])
    prob = euler_adjoint.EulerAdjointProblem(mesh, p, pts, dt, nsteps, nsnap=nsnap, gamma=GAMMA, flux=flux)
    return mesh, orc, Uh, rec, prob


def _oracle_obs(orc, Uh, rec, dt, nsteps):
    orc.set_modal(Uh)
    y = np.zeros((nsteps, len(rec)))
    for k in range(nsteps):
        orc.advance(dgx.DGX_LSRK14, dt, 1)
        U = orc.state()
        y[k] = [_pressure(U[:, e, i]) for e, i in rec]
    return y


def test_forward_matches_oracle_and_advance():
    mesh, orc, Uh, rec, prob = _setup()
    y, UT = prob.forward(Uh)
    y_ref = _oracle_obs(orc, Uh, rec, prob.dt, prob.nsteps)
    rel = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    print(f"[euler-adjoint] observations vs oracle rel {rel:.2e}", flush=True)
    assert rel <= 1e-12, rel
    dom = dgx.GPUDomain(mesh, prob.p, dgx.DGX_EULER, gamma=GAMMA)
    dom.upload(Uh, dgx.DGX_MODAL)
    dom.advance(dgx.DGX_LSRK14, prob.dt, prob.nsteps)
    ref = dom.download(dgx.DGX_MODAL)
    print(f"[euler-adjoint] final state vs dgx_advance: max abs diff {np.abs(UT - ref).max()}", flush=True)
    assert np.array_equal(UT, ref)   # the same stage kernel (k_step2) in both
    dom.close()
    prob.close()


@pytest.mark.parametrize("flux", [dgx.DGX_FLUX_LF, dgx.DGX_FLUX_ROE])
def test_gradient_against_oracle_differences(flux):
    mesh, orc, Uh, rec, prob = _setup(flux=flux)
    rng = np.random.default_rng(11)
    U_true = Uh * (1.0 + 0.01 * rng.standard_normal(Uh.shape))
    data = _oracle_obs(orc, U_true, rec, prob.dt, prob.nsteps)
    J, g = prob.gradient(Uh, data)
    y_ref = _oracle_obs(orc, Uh, rec, prob.dt, prob.nsteps)
    J_ref = 0.5 * prob.dt * float(np.sum((y_ref - data) ** 2))
    print(f"[euler-adjoint] J {J:.15e} oracle {J_ref:.15e}", flush=True)
    assert abs(J - J_ref) <= 1e-10 * J_ref
    eps = 1e-5
    for t in range(3):
        v = _direction(orc, 0.0, 7 + t)
        Jp = 0.5 * prob.dt * float(np.sum((_oracle_obs(orc, Uh + eps * v, rec, prob.dt, prob.nsteps) - data) ** 2))
        Jm = 0.5 * prob.dt * float(np.sum((_oracle_obs(orc, Uh - eps * v, rec, prob.dt, prob.nsteps) - data) ** 2))
        fd = (Jp - Jm) / (2 * eps)
        ad = float(np.sum(g * v))
        rel = abs(ad - fd) / abs(fd)
        print(f"[euler-adjoint] direction {t}: adjoint {ad:.12e} differences {fd:.12e} rel {rel:.2e}", flush=True)
        assert rel <= 1e-6, rel
    st = prob.stats()
    assert st["adjoint_steps"] == prob.nsteps
    prob.close()


def test_zero_misfit_and_checkpoint_invariance():
    mesh, orc, Uh, rec, prob = _setup(nsteps=12)
    y, _ = prob.forward(Uh)
    J0, g0 = prob.gradient(Uh, y)
    print(f"[euler-adjoint] zero misfit: J {J0}, max |g| {np.abs(g0).max()}", flush=True)
    assert J0 == 0.0 and not np.any(g0)
    data = y + 1e-3 * np.random.default_rng(5).standard_normal(y.shape)
    grads = []
    for nsnap in (2, 3, 0):
        pr = euler_adjoint.EulerAdjointProblem(mesh, prob.p, prob._rec, prob.dt, prob.nsteps, nsnap=nsnap, gamma=GAMMA)
        J, g = pr.gradient(Uh, data)
        st = pr.stats()
        print(f"[euler-adjoint] nsnap={nsnap}: J {J:.15e}, {st}", flush=True)
        assert st["adjoint_steps"] == prob.nsteps and st["snapshots"] <= (nsnap or 5)
        grads.append((J, g))
        pr.close()
    for J, g in grads[1:]:
        assert J == grads[0][0] and np.array_equal(g, grads[0][1])
    prob.close()


def test_reo_opt_problem_descends():
    """ReoOptProblem (python/dgm.pyx:322-355) over the Euler problem: the
    control is the initial state, and steepest descent on the adjoint gradient
    lowers the misfit."""
    from dgm_amd.acoustic import ReoOptProblem
    mesh, orc, Uh, rec, prob = _setup(nsteps=10)
    U_true = Uh + 0.02 * _direction(orc, 0.0, 21)
    data, _ = prob.forward(U_true)
    opt = ReoOptProblem(prob, data, theta0=Uh)
    J0 = opt.computeState()
    opt.computeAdjoint()
    assert opt.getGradient().shape == Uh.shape and np.any(opt.getGradient())
    J = opt.solve(iterations=3, step=10.0)
    print(f"[euler-adjoint] descent: J {J0:.6e} -> {J:.6e}", flush=True)
    assert J < J0
    prob.close()
