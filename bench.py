"""Benchmark: DG DOF-stage updates per second on one MI355X.

Default workload: 3D compressible Navier-Stokes Taylor-Green vortex on the
periodic box [0, 2pi]^3, 32^3 affine hexes at p = 7 (8^3 modes per element,
16.8 M modal DOF per field, 5 fields), Re = 1600 (mu = 1/1600), Ma = 0.1,
Pr = 0.71, Lax-Friedrichs flux with BR1 viscous terms (the reference's
Navier_Stokes3d operator), LSRK14 stepping (14 stages per step), FP64.

A step is one LSRK14 step of the whole domain with the state resident in HBM.
value = modal DOF x stages / wall time of the timed steps.

Also printed in the same JSON line:
  roofline      dominant kernel's algorithmic HBM bytes per launch (32 B per
                DOF-stage: read U, read Ut, write Ut, write U of a fused 2N-storage
                stage) / its mean launch time from HIP events on the library's
                stream, against MI355X's 8 TB/s
  cpu_baseline  the CPU oracle (1 core) on a bounded sample of the same
                workload (same p and physics on a 3^3 periodic box)
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from dgm_amd import dgx, ic, mesh as meshgen  # noqa: E402

METRIC = ("DG DOF-stage updates/s, mixed vs FP64 at ≤1e-6 rel err; p50/p99 solve latency")
HBM_PEAK_GBS = 8000.0
BYTES_PER_DOF_STAGE = 32.0
KINDS = ["traces", "residual", "grad", "visc", "forward", "backward", "update"]


def tgv_domain(n, p, mu):
    m = meshgen.periodic_hexes(n, n, n, (0.0, 0.0, 0.0), (2 * np.pi,) * 3)
    return m


def tgv_dt(n, p, gamma=1.4, Ma=0.1, cfl=0.5):
    h = 2 * np.pi / n
    c = np.sqrt(gamma * (1.0 / (gamma * Ma * Ma) + 0.25)) + 1.0
    return cfl * h / ((p + 1) ** 2 * c)


def cpu_baseline(p, mu, seconds=15.0):
    """Time the oracle's LSRK14 on a 3^3-element box of the same order/physics."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    n = 3
    m = tgv_domain(n, p, mu)
    orc = Oracle(m, p, 9, vis=mu, Pr=0.71)
    x = orc.coords()
    orc.set_state(ic.taylor_green(x[0], x[1], x[2]))
    orc.project_state()
    dt = tgv_dt(32, p)
    dofs = 5 * m.ne * (p + 1) ** 3
    orc.advance(14, dt, 1)                 # first step excluded, as src/Domain.cpp:937-981
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        orc.advance(14, dt, 1)
        steps += 1
    el = time.perf_counter() - t0
    return {"value": dofs * 14 * steps / el, "unit": "DOF-stage/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle, NS3d TGV {n}^3 hexes p={p}, {steps} LSRK14 steps "
                      f"({14 * steps} residuals) in {el:.1f} s on 1 thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=32, help="elements per direction")
    ap.add_argument("--p", type=int, default=7)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with one device-to-device copy of the modal state (a plain copy "
                         "kernel of known size for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this engine has no multi-GPU path in this build "
              "(sharding is described in DESIGN.md); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    n, p, mu = args.n, args.p, 1.0 / 1600.0
    m = tgv_domain(n, p, mu)
    gpu = dgx.GPUDomain(m, p, dgx.DGX_NAVIER_STOKES3D, gamma=1.4, vis=mu, Pr=0.71)
    x = gpu.coords()
    U = ic.taylor_green(x[0], x[1], x[2])
    del x
    gpu.upload(U)
    del U
    gpu.synchronize()
    dt = tgv_dt(n, p)
    dofs = 5 * m.ne * (p + 1) ** 3
    print(f"[bench] TGV NS3d {n}^3 p={p}: {dofs / 1e6:.1f} M DOF, dt={dt:.3e}", file=sys.stderr, flush=True)

    gpu.advance(dgx.DGX_LSRK14, dt, args.warmup)
    gpu.synchronize()
    t0 = time.perf_counter()
    gpu.advance(dgx.DGX_LSRK14, dt, args.steps)
    gpu.synchronize()
    el = time.perf_counter() - t0
    stages = 14 * args.steps
    value = dofs * stages / el
    print(f"[bench] timed {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)

    # per-kernel times from HIP events on the library's stream
    ms = np.zeros(7)
    cnt = np.zeros(7, dtype=np.int32)
    prof_steps = 1
    rc = gpu.lib.dgx_profile(gpu.h, dgx.DGX_LSRK14, dt, prof_steps,
                             ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                             cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), 7)
    if rc != 0:
        raise RuntimeError(gpu.lib.dgx_last_error().decode())
    gpu.synchronize()
    kern = {KINDS[k]: {"ms_total": float(ms[k]), "launches": int(cnt[k]),
                       "ms_mean": (ms[k] / cnt[k]) if cnt[k] else 0.0} for k in range(7)}
    dom = max(kern, key=lambda k: kern[k]["ms_total"])
    mean_s = kern[dom]["ms_mean"] / 1e3
    algo_bytes = BYTES_PER_DOF_STAGE * dofs
    achieved = algo_bytes / mean_s / 1e9
    norms = gpu.norms()
    finite = bool(np.all(np.isfinite(norms)))
    if args.copy_calibration:
        nbytes = 8 * dofs
        rc = gpu.lib.dgx_residual_download(gpu.h, ctypes.cast(gpu.device_state_ptr(),
                                           ctypes.POINTER(ctypes.c_double)), dgx.DGX_MODAL, 1)
        if rc != 0:
            raise RuntimeError(gpu.lib.dgx_last_error().decode())
        print(f"[bench] calibration copy: {nbytes} bytes read and written", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D Navier-Stokes Taylor-Green vortex, periodic [0,2pi]^3, {n}^3 hexes, "
                               f"p={p}, Re=1600, Ma=0.1, Pr=0.71, LF + BR1, LSRK14, FP64",
                   "dofs": dofs, "stages_per_step": 14, "dt": dt, "precision": "fp64",
                   "state_finite": finite},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom,
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "stage_rate_GBs": value * BYTES_PER_DOF_STAGE / 1e9},
        "kernels": kern,
    }
    gpu.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(p, mu, args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
