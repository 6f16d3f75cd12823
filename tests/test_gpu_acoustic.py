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

from dgm_amd import acoustic, mesh as meshgen
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
