"""Acoustic inverse-source problem on the GPU (dgx_acoustic_*).

* Forward observations and the discrete-adjoint gradient against the test-side
  reference built on the CPU oracle (tests/acoustic_ref.py, itself pinned by
  tests/test_oracle_acoustic.py): 1e-12 / 1e-10 relative.
* The gradient against central differences of the GPU objective along seeded
  directions.  J is quadratic in theta, so the differences are exact up to
  rounding: 1e-8.
* The checkpoint count changes the schedule of recomputation, not the result:
  gradients with 2, 5 and the default number of slots are bitwise equal.
"""
from __future__ import annotations

import numpy as np
import pytest

from dgm_amd import acoustic, dgx, mesh as meshgen
from acoustic_ref import AcousticRef
from test_oracle_acoustic import stretched_quads

pytestmark = pytest.mark.gpu

GAMMA = 1.4
MEAN = (1.0, 0.3, -0.1, 1.0 / GAMMA)


def _pair(m, p, rec, f0, t0, dt, nsteps, nsnap=0):
    gpu = acoustic.AcousticProblem(m, p, MEAN, rec, f0, t0, dt, nsteps, nsnap=nsnap, gamma=GAMMA)
    ref = AcousticRef(m, p, MEAN, rec, f0, t0, dt, nsteps, gamma=GAMMA)
    return gpu, ref


@pytest.mark.parametrize("n,p", [(4, 2), (5, 5)])
def test_forward_and_gradient_match_oracle(n, p):
    print(f"[acoustic] oracle parity n={n} p={p}", flush=True)
    m = stretched_quads(n, n)
    rec = [(0.13, 0.21), (0.77, 0.4), (0.5, 0.9)]
    gpu, ref = _pair(m, p, rec, f0=3.0, t0=0.02, dt=2e-3, nsteps=10, nsnap=3)
    rng = np.random.default_rng(n + p)
    theta_true = rng.standard_normal(m.ne)
    y_gpu, y_ref = gpu.forward(theta_true), ref.forward(theta_true)
    rel = np.linalg.norm(y_gpu - y_ref) / np.linalg.norm(y_ref)
    print(f"[acoustic] observations rel {rel:.2e}", flush=True)
    assert rel <= 1e-12, rel
    theta = rng.standard_normal(m.ne)
    J_gpu, g_gpu = gpu.gradient(theta, y_ref)
    J_ref, g_ref = ref.gradient(theta, y_ref)
    gr = np.linalg.norm(g_gpu - g_ref) / np.linalg.norm(g_ref)
    print(f"[acoustic] J rel {abs(J_gpu - J_ref) / J_ref:.2e}, gradient rel {gr:.2e}", flush=True)
    assert abs(J_gpu - J_ref) <= 1e-10 * J_ref
    assert gr <= 1e-10, gr
    st = gpu.stats()
    assert st["adjoint_steps"] == 10
    gpu.close()


def _medium(nsnap=0):
    m = meshgen.periodic_quads(16, 16, 0.0, 1.0, 0.0, 1.0)
    rng = np.random.default_rng(99)
    rec = rng.uniform(0.0, 1.0, size=(16, 2))
    p = 5
    dt = 0.3 * (1.0 / 16) / ((p + 1) ** 2 * 1.35)
    return m, p, rec, dt


def test_gradient_central_differences_medium():
    m, p, rec, dt = _medium()
    nsteps = 200
    print(f"[acoustic] FD check 16x16 p={p} {nsteps} steps", flush=True)
    gpu = acoustic.AcousticProblem(m, p, MEAN, rec, 20.0, 0.05, dt, nsteps, gamma=GAMMA)
    rng = np.random.default_rng(4096)
    data = gpu.forward(rng.standard_normal(m.ne))
    theta = rng.standard_normal(m.ne)
    J, g = gpu.gradient(theta, data)
    assert J > 0 and np.all(np.isfinite(g))
    assert abs(J - gpu.objective(theta, data)) <= 1e-12 * J
    for k in range(8):
        d = rng.standard_normal(m.ne)
        h = 1e-1
        fd = (gpu.objective(theta + h * d, data) - gpu.objective(theta - h * d, data)) / (2 * h)
        err = abs(fd - g @ d) / (abs(fd) + 1e-300)
        print(f"[acoustic] direction {k}: fd {fd:.6e} adjoint {g @ d:.6e} rel {err:.2e}", flush=True)
        assert err <= 1e-8, err
    gpu.close()


def test_checkpoint_count_does_not_change_the_gradient():
    m, p, rec, dt = _medium()
    nsteps = 60
    rng = np.random.default_rng(7)
    theta = rng.standard_normal(m.ne)
    out = []
    for nsnap in (2, 5, 0):
        gpu = acoustic.AcousticProblem(m, p, MEAN, rec, 20.0, 0.01, dt, nsteps, nsnap=nsnap, gamma=GAMMA)
        data = np.zeros((nsteps, len(rec)))
        out.append(gpu.gradient(theta, data))
        st = gpu.stats()
        adv, used, ok = acoustic.revolve_count(nsteps, nsnap or int(np.ceil(np.sqrt(nsteps))) + 1)
        assert ok and st["forward_steps"] == adv and st["adjoint_steps"] == nsteps
        gpu.close()
    for J, g in out[1:]:
        assert J == out[0][0] and np.array_equal(g, out[0][1])


def test_mixed_against_fp64():
    """DGX_MIXED (FP32 element contractions in the forward and adjoint
    kernels; FP64 state, registers, observations and gradient sums) against
    the FP64 problem on the same inputs, and the mixed gradient against central
    differences of the mixed objective.  Stated bars (measured in brackets):
    observations 1e-5 [1.9e-6], J 1e-6 [3.9e-8], gradient 1e-5 [8.2e-7]; the
    mixed gradient against its own objective's differences, |fd - g.d| /
    (|g| |d|), 1e-6 (relative to |fd| alone it is large in directions nearly
    orthogonal to g, where fd is small)."""
    m, p, rec, dt = _medium()
    nsteps = 200
    print(f"[acoustic] mixed vs FP64 16x16 p={p} {nsteps} steps", flush=True)
    g64 = acoustic.AcousticProblem(m, p, MEAN, rec, 20.0, 0.05, dt, nsteps, gamma=GAMMA)
    g32 = acoustic.AcousticProblem(m, p, MEAN, rec, 20.0, 0.05, dt, nsteps, gamma=GAMMA,
                                   precision=dgx.DGX_MIXED)
    rng = np.random.default_rng(4097)
    theta_true = rng.standard_normal(m.ne)
    y64, y32 = g64.forward(theta_true), g32.forward(theta_true)
    ey = np.linalg.norm(y32 - y64) / np.linalg.norm(y64)
    theta = rng.standard_normal(m.ne)
    J64, gr64 = g64.gradient(theta, y64)
    J32, gr32 = g32.gradient(theta, y64)
    eJ, eg = abs(J32 - J64) / J64, np.linalg.norm(gr32 - gr64) / np.linalg.norm(gr64)
    print(f"[acoustic] mixed: observations rel {ey:.2e}, J rel {eJ:.2e}, gradient rel {eg:.2e}", flush=True)
    errs = []
    for k in range(4):
        d = rng.standard_normal(m.ne)
        h = 1e-1
        fd = (g32.objective(theta + h * d, y64) - g32.objective(theta - h * d, y64)) / (2 * h)
        errs.append(abs(fd - gr32 @ d) / (np.linalg.norm(gr32) * np.linalg.norm(d)))
    print(f"[acoustic] mixed gradient vs mixed differences: {['%.2e' % e for e in errs]}", flush=True)
    assert ey <= 1e-5 and eJ <= 1e-6 and eg <= 1e-5, (ey, eJ, eg)
    assert max(errs) <= 1e-6, errs
    g64.close()
    g32.close()
