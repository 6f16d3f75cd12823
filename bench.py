"""Benchmark: DG DOF-stage updates per second on one MI355X.

Workload: 3D compressible Navier-Stokes Taylor-Green vortex on the periodic
box [0, 2pi]^3, 32^3 affine hexes at p = 7 (8^3 modes per element, 16.8 M
modal DOF per field, 5 fields), Re = 1600 (mu = 1/1600), Ma = 0.1, Pr = 0.71,
Roe flux with BR2 viscous terms by default (--flux lf --viscous br1 selects
the reference's own Navier_Stokes3d operator), LSRK14 stepping (14 stages per
step).

Both paths run on the same initial state: the FP64 path (`value`) and the
mixed path (FP32 MFMA contractions, FP64 state and registers; `mixed`), whose
final state is compared with the FP64 one per conserved field (relative L2 of
the modal coefficients, the Domain::diff L2).

A step is one LSRK14 step of the whole domain with the state resident in HBM.
value = modal DOF x stages / wall time of the timed steps.

Also printed in the same JSON line:
  roofline      the dominant kernel (largest share of a step's time, from HIP
                events on the library's stream) against the compute roofline:
                its algorithmic flops per launch / its mean launch time, against
                MI355X's 78.6 TFLOP/s FP64 peak (157.3 TFLOP/s FP32 matrix for the
                mixed path).  Algorithmic flops per DOF-stage at p = 7 (minimal
                sum-factorized operator): 787 in all, 256 of them in the inviscid
                residual kernel, the other 531 in the BR1 viscous kernel; BR2's
                face pass adds 85 to the viscous kernel.
                `bound` is "valu" for the FP64 kernels (every flop on the FP64
                vector pipe) and "mfma" for the mixed ones.  `traffic` is the
                kernel's HBM bytes per launch measured with rocprofv3 FETCH_SIZE
                and WRITE_SIZE (scaled by a plain copy of known size) in a
                separate profiling run of this workload: a fixed calibration,
                not read from this run's counters.
  roofline_hbm  the HBM view of a whole stage: 32 B per DOF-stage (read U,
                read Ut, write Ut, write U of a fused 2N-storage stage) / the
                mean stage time, against 8 TB/s
  adjoint       the acoustic inverse-source workload (64x64 quads, p=5, 4096
                source coefficients, 16 receivers, 2000 LSRK14 steps, 45
                checkpoints): time of one forward solve and of one gradient
                (checkpointed forward sweep + discrete adjoint), and the
                gradient's agreement with central differences in 8 directions;
                and the nonlinear Euler initial-state gradient (64x64 quads,
                p=5, 16 pressure receivers, 200 LSRK14 steps, 15 checkpoints):
                forward and gradient times, central differences in 2 directions
  serving       p50/p99 latency and throughput of a serving-loop sample: Poisson
                arrivals of 2D Euler and shallow-water forward solves of random
                size and order, continuously batched (dgm_amd.serve), in the
                FP32-mixed precision the ensemble configuration names;
                serving_fp64 the same sample in FP64
  element_block the element-block seam (GpGpu::hex_kernels' inner_product and
                backward_transform) on 32768 hexes at 8 modes / 9 points per
                direction, device-resident, with its HBM GB/s
  cpu_baseline  the CPU oracle on the host cores (one worker per core, at
                most 16) on a bounded sample of the same workload: same p and
                physics on a 3^3 periodic box per worker, ~15 s
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
FP64_PEAK_TFLOPS = 78.6
FP32_MFMA_PEAK_TFLOPS = 157.3
ARITHMETIC = {"fp64": "FP64 VALU (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)",
              "mixed": "FP32 MFMA contractions with FP64 pointwise work; peak is the FP32 matrix rate"}
BYTES_PER_DOF_STAGE = 32.0
# algorithmic flops per DOF-stage by kernel kind, per order (3D NS, BR1); BR2
# adds its face pass to the viscous kernel: nodal face lifts (2 x 29,376 FMA),
# the per-face-node correction (486 nodes x ~154 flops), the lift of the flux
# difference (29,376 FMA) and its face sum (12,288 FMA) per element of 2,560 DOF
FLOPS_PER_DOF_STAGE = {7: {"residual": 256.0, "visc": 531.0}}
BR2_VISC_FLOPS = {7: 85.0}
# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on
# the 32^3 p=7 workload, scaled by the counters of a plain copy of known size
# (profiles/tgv_ns3d_32cubed_p7_roe_br2.md, profiles/tgv_ns3d_32cubed_p7_lf_br1.md)
# keys: path, n, p, kernel, flux, viscous scheme
MEASURED_TRAFFIC = {("fp64", 32, 7, "visc", "roe", "br2"): 2.308e9, ("fp64", 32, 7, "residual", "roe", "br2"): 5.789e9,
                    ("mixed", 32, 7, "visc", "roe", "br2"): 1.836e9, ("mixed", 32, 7, "residual", "roe", "br2"): 5.197e9,
                    ("fp64", 32, 7, "visc", "lf", "br1"): 2.308e9, ("fp64", 32, 7, "residual", "lf", "br1"): 5.789e9}
KINDS = ["traces", "residual", "grad", "visc", "forward", "backward", "update"]
FLUXES = {"lf": dgx.DGX_FLUX_LF, "roe": dgx.DGX_FLUX_ROE}
VISCOUS = {"br1": dgx.DGX_VISCOUS_BR1, "br2": dgx.DGX_VISCOUS_BR2}
OPNAME = {"lf": "LF", "roe": "Roe"}


def tgv_domain(n, p, mu):
    m = meshgen.periodic_hexes(n, n, n, (0.0, 0.0, 0.0), (2 * np.pi,) * 3)
    return m


def tgv_dt(n, p, gamma=1.4, Ma=0.1, cfl=0.5):
    h = 2 * np.pi / n
    c = np.sqrt(gamma * (1.0 / (gamma * Ma * Ma) + 0.25)) + 1.0
    return cfl * h / ((p + 1) ** 2 * c)


def _oracle_worker(args):
    """One CPU worker: the oracle's LSRK14 on its own share of the TGV mesh, a
    periodic n x n x nz slab of the configuration's hexes (same element size and
    order), timed over `steps` steps after one untimed step."""
    p, mu, n, nz, steps, flux, viscous = args
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    h = 2 * np.pi / n
    m = meshgen.periodic_hexes(n, n, nz, (0.0, 0.0, 0.0), (2 * np.pi, 2 * np.pi, nz * h))
    orc = Oracle(m, p, 9, vis=mu, Pr=0.71, flux=flux, viscous=viscous)
    x = orc.coords()
    orc.set_state(ic.taylor_green(x[0], x[1], x[2]))
    orc.project_state()
    dt = tgv_dt(n, p)
    dofs = 5 * m.ne * (p + 1) ** 3
    orc.advance(14, dt, 1)                 # first step excluded, as src/Domain.cpp:937-981
    t0 = time.perf_counter()
    orc.advance(14, dt, steps)
    return dofs * 14 * steps, time.perf_counter() - t0, steps, m.ne


def cpu_baseline(p, mu, n=32, workers=None, steps=1, flux=0, viscous=0):
    """The reference's path (the oracle, a C++ restatement of
    Navier_Stokes3d's LF + BR1 residual and LSRK14) on the host cores: the
    n^3 mesh cut into `workers` periodic slabs of n x n x n/workers hexes, one
    per core, as the reference runs one MPI rank per core with single-threaded
    BLAS (src/dgm.cpp:165), here without the halo exchange.  Together the
    slabs hold the configuration's elements.  Spawned before this process
    touches the GPU; the first step is excluded (src/Domain.cpp:937-981)."""
    import multiprocessing as mp
    nproc = os.cpu_count()
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else nproc
    workers = workers or max(1, min(16, ncpu))
    while n % workers:
        workers -= 1
    nz = n // workers
    with mp.get_context("spawn").Pool(workers) as pool:
        res = pool.map(_oracle_worker, [(p, mu, n, nz, steps, flux, viscous)] * workers)
    total = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    per_core = [r[0] / r[1] for r in res]
    ne = sum(r[3] for r in res)
    return {"value": total / wall, "unit": "DOF-stage/s", "cores": workers, "kind": "port",
            "nproc": nproc, "affinity_cpus": ncpu, "one_core": float(np.mean(per_core)),
            "operator": f"{'Roe' if flux else 'LF'} + {'BR2' if viscous else 'BR1'}",
            "sample": f"CPU oracle (C++ restatement of the reference's Navier_Stokes3d residual and LSRK14, "
                      f"FP64) on the TGV mesh: {ne} hexes at p={p} as {workers} periodic {n}x{n}x{nz} slabs, "
                      f"one process per core, {steps} timed LSRK14 step(s) each after one untimed step, "
                      f"{wall:.1f} s"}


def serving_sample(args, seed=1234, precision=dgx.DGX_MIXED):
    """A serving mix shaped like the ensemble workload: Poisson arrivals of 2D
    forward solves, each shallow water about rest (phi = 1) or Euler about a
    freestream (seeded choice), plus 8 seeded Fourier modes of amplitude 0.05;
    n x n periodic quads with n log-uniform in [8, nmax], p uniform in {2..8},
    200 LSRK14 steps each, FP32-mixed (FP32 element contractions, FP64 state)
    or FP64, continuous batching on one GPU (requests of one order, physics
    and precision share their launches unless --serve-unbatched).
    The shallow-water requests leave out the balance term, which would hold
    them at their initial state."""
    from dgm_amd import serve
    from dgm_amd.mesh import quad_nodes
    rng = np.random.default_rng(seed)
    jobs, dofs = [], []
    base = {dgx.DGX_EULER: [1.0, 0.5, 0.0, (1.0 / 1.4) / 0.4 + 0.125], dgx.DGX_SHALLOW: [1.0, 0.0, 0.0]}
    speed = {dgx.DGX_EULER: 0.5 + np.sqrt(1.4 * (1 / 1.4)), dgx.DGX_SHALLOW: 1.0}
    nshallow = 0
    for _ in range(args.serve_requests):
        nn = int(round(np.exp(rng.uniform(np.log(8), np.log(args.serve_nmax)))))
        pp = int(rng.integers(2, 9))
        eqn = dgx.DGX_SHALLOW if rng.integers(2) == 0 else dgx.DGX_EULER
        nshallow += eqn == dgx.DGX_SHALLOW
        m = meshgen.periodic_quads(nn, nn, 0.0, 10.0, 0.0, 10.0)
        x = quad_nodes(m, pp)   # nodal GLL coordinates without a device domain
        U = np.stack([b + ic.fourier_modes(x, (10.0, 10.0), 8, 0.05, rng) for b in base[eqn]])
        dt = 0.2 * (10.0 / nn) / ((pp + 1) ** 2 * speed[eqn])
        job = dict(mesh=m, p=pp, eqntype=eqn, U0=U, inttype=dgx.DGX_LSRK14, dt=dt, nsteps=args.serve_steps,
                   precision=precision)
        if eqn == dgx.DGX_SHALLOW:
            job["balance"] = False
        jobs.append(job)
        dofs.append(len(base[eqn]) * nn * nn * (pp + 1) ** 2)
    if args.serve_requests > 64:
        print(f"[serve] built {len(jobs)} requests", flush=True)
    srv = serve.Server(hbm_budget=200 * 2**30, max_active=args.serve_active, batch=not args.serve_unbatched,
                       steps_per_slice=args.serve_slice, persistent=args.serve_persistent)
    arrivals = serve.poisson_arrivals(len(jobs), rate=args.serve_rate, rng=rng)
    t0 = time.perf_counter()
    ids = serve.serve_timeline(srv, jobs, arrivals, timeout_s=1200.0,
                               report_every=32 if args.serve_requests > 64 else 0)
    wall = time.perf_counter() - t0
    lat, wait = [], []
    for rid in ids:
        _, (ts, ta, td) = srv.result(rid)
        lat.append(td - ts)
        wait.append(ta - ts)
    peak_bytes, peak_active = srv.peak()
    srv.close()
    lat_ms = np.array(lat) * 1e3
    return {"requests": len(jobs), "p50_latency_ms": float(np.percentile(lat_ms, 50)),
            "p99_latency_ms": float(np.percentile(lat_ms, 99)),
            "mean_queue_wait_ms": float(np.mean(wait) * 1e3),
            "throughput_dof_stage_per_s": float(sum(dofs) * 14 * args.serve_steps / wall),
            "makespan_s": wall, "peak_active": peak_active, "peak_device_bytes": peak_bytes,
            "precision": "mixed" if precision == dgx.DGX_MIXED else "fp64",
            "sample": f"{len(jobs)} 2D solves ({nshallow} shallow water, {len(jobs) - nshallow} Euler), n log-uniform in [8,{args.serve_nmax}], p in 2..8, "
                      f"{args.serve_steps} LSRK14 steps, Poisson arrivals at {args.serve_rate}/s, "
                      f"{'FP32-mixed' if precision == dgx.DGX_MIXED else 'FP64'}, "
                      f"{args.serve_active} slots, slices of {args.serve_slice} steps, "
                      + ("per-request launches replayed from hipGraphs" if args.serve_unbatched
                         else "requests of one order and physics batched into shared launches"
                         + (", each group slice one persistent cooperative kernel" if args.serve_persistent
                            else ", one launch per stage (traces of the next stage written by the residual)"))}


def adjoint_sample(args, seed=4096):
    """The acoustic inverse-source workload: linearized 2D Euler about a rest
    state on 64x64 periodic quads at p = 5, a source of one seeded Gaussian
    coefficient per element (4096) with a Ricker time signature, 16 seeded
    receivers, 2000 LSRK14 steps, ceil(sqrt(2000)) = 45 checkpoints besides
    the initial state.  One gradient = one checkpointed forward sweep and the
    adjoint sweep; checked against central differences of the objective in 8
    seeded directions (J is quadratic in theta, so they are exact up to
    rounding).  FP64."""
    from dgm_amd import acoustic
    rng = np.random.default_rng(seed)
    n, p, nsteps = args.adj_n, 5, args.adj_steps
    m = meshgen.periodic_quads(n, n, 0.0, 1.0, 0.0, 1.0)
    rec = rng.uniform(0.0, 1.0, size=(16, 2))
    dt = 0.3 * (1.0 / n) / ((p + 1) ** 2 * 1.0)
    nsnap = int(np.ceil(np.sqrt(nsteps))) + 1
    f0, t0 = 4.0 / (nsteps * dt), 0.3 * nsteps * dt
    prob = acoustic.AcousticProblem(m, p, (1.0, 0.0, 0.0, 1.0 / 1.4), rec, f0, t0, dt, nsteps, nsnap=nsnap)
    data = prob.forward(rng.standard_normal(m.ne))
    theta = rng.standard_normal(m.ne)
    t0w = time.perf_counter()
    prob.forward(theta)
    t_fwd = time.perf_counter() - t0w
    t0w = time.perf_counter()
    J, g = prob.gradient(theta, data)
    t_grad = time.perf_counter() - t0w
    st = prob.stats()
    errs = []
    h = 1e-1
    for _ in range(args.adj_dirs):
        d = rng.standard_normal(m.ne)
        fd = (prob.objective(theta + h * d, data) - prob.objective(theta - h * d, data)) / (2 * h)
        errs.append(abs(fd - g @ d) / max(abs(fd), 1e-300))
    prob.close()
    # the mixed variant (FP32 element contractions) on the same inputs
    pm = acoustic.AcousticProblem(m, p, (1.0, 0.0, 0.0, 1.0 / 1.4), rec, f0, t0, dt, nsteps, nsnap=nsnap,
                                  precision=dgx.DGX_MIXED)
    t0w = time.perf_counter()
    pm.forward(theta)
    tm_fwd = time.perf_counter() - t0w
    t0w = time.perf_counter()
    Jm, gm = pm.gradient(theta, data)
    tm_grad = time.perf_counter() - t0w
    pm.close()
    dofs = 4 * m.ne * (p + 1) ** 2
    adv, _, _ = acoustic.revolve_count(nsteps, nsnap)
    return {"workload": f"linearized 2D Euler acoustics, {n}x{n} quads, p={p}, {m.ne} source coefficients, "
                        f"16 receivers, {nsteps} LSRK14 steps, {nsnap - 1} checkpoints + the initial state, FP64",
            "forward_s": t_fwd, "gradient_s": t_grad, "J": J,
            "forward_steps_per_gradient": adv, "adjoint_steps_per_gradient": nsteps,
            "forward_dof_stage_per_s": dofs * 14 * nsteps / t_fwd,
            "fd_directions": len(errs), "fd_rel_err_max": float(max(errs)) if errs else None,
            "fd_rel_err": [float(e) for e in errs],
            "mixed": {"forward_s": tm_fwd, "gradient_s": tm_grad, "J_rel_vs_fp64": abs(Jm - J) / J,
                      "gradient_rel_l2_vs_fp64": float(np.linalg.norm(gm - g) / np.linalg.norm(g)),
                      "speedup_gradient_vs_fp64": t_grad / tm_grad}}


def euler_adjoint_sample(args, seed=77):
    """Initial-state gradient through nonlinear 2D Euler: 64x64 periodic quads
    on [0, 2 pi]^2 at p = 5, a smooth seeded initial flow, 16 seeded
    receivers observing the pressure after every step, 200 LSRK14 steps at
    CFL 0.5, ceil(sqrt(200)) = 15 checkpoints besides the initial state.  One
    gradient = one checkpointed forward sweep, each reversed step recomputing
    its 14 stage states, and the backward kernels (dgx_adjoint2d.hip) at every
    stage; checked against central differences of the objective along smooth
    seeded directions.  FP64."""
    from dgm_amd import euler_adjoint
    rng = np.random.default_rng(seed)
    n, p, nsteps = args.eadj_n, 5, args.eadj_steps
    L = 2 * np.pi
    m = meshgen.periodic_quads(n, n, 0.0, L, 0.0, L)
    dom = dgx.GPUDomain(m, p, dgx.DGX_EULER)
    x = dom.coords()

    def smooth(scale):
        return np.stack([sum(scale[c] * rng.uniform(0.2, 1.0) *
                             np.sin(rng.integers(0, 3) * x[0] + rng.integers(1, 3) * x[1] + rng.uniform(0, L))
                             for _ in range(3)) for c in range(4)])

    def modal(f):
        dom.upload(f)
        return dom.download(dgx.DGX_MODAL)

    rho = 1.0 + 0.1 * np.sin(x[0]) * np.cos(x[1])
    u, v = 0.4 + 0.1 * np.sin(x[0] + x[1]), 0.25 - 0.2 * np.cos(x[0]) * np.sin(2 * x[1])
    pr = 1.0 / 1.4 + 0.05 * np.cos(2 * x[0]) * np.cos(x[1])
    U0 = modal(np.stack([rho, rho * u, rho * v, pr / 0.4 + 0.5 * rho * (u * u + v * v)]))
    dom.upload(U0, dgx.DGX_MODAL)
    _, dt = dom.set_time_step(1e-3, 0.5)
    nsnap = int(np.ceil(np.sqrt(nsteps))) + 1
    rec = rng.uniform(0.0, L, size=(16, 2))
    prob = euler_adjoint.EulerAdjointProblem(m, p, rec, dt, nsteps, nsnap=nsnap)
    data, _ = prob.forward(U0 + 0.02 * modal(smooth((1.0, 0.4, 0.25, 2.5))))
    t0w = time.perf_counter()
    prob.forward(U0)
    t_fwd = time.perf_counter() - t0w
    t0w = time.perf_counter()
    J, g = prob.gradient(U0, data)
    t_grad = time.perf_counter() - t0w
    st = prob.stats()
    errs = []
    h = 1e-5
    for _ in range(args.adj_dirs // 4):
        d = modal(smooth((1.0, 0.4, 0.25, 2.5)))
        fd = (prob.objective(U0 + h * d, data) - prob.objective(U0 - h * d, data)) / (2 * h)
        errs.append(abs(fd - float(np.sum(g * d))) / max(abs(fd), 1e-300))
    prob.close()
    # the same gradient with the Roe flux (its transposed Jacobian from dual-number derivatives)
    dirs = [modal(smooth((1.0, 0.4, 0.25, 2.5))) for _ in range(args.adj_dirs // 4)]
    roe = euler_adjoint.EulerAdjointProblem(m, p, rec, dt, nsteps, nsnap=nsnap, flux=dgx.DGX_FLUX_ROE)
    data_r, _ = roe.forward(U0 + 0.02 * modal(smooth((1.0, 0.4, 0.25, 2.5))))
    t0w = time.perf_counter()
    roe.forward(U0)
    t_fwd_r = time.perf_counter() - t0w
    t0w = time.perf_counter()
    _, g_r = roe.gradient(U0, data_r)
    t_grad_r = time.perf_counter() - t0w
    errs_r = []
    for d in dirs:
        fd = (roe.objective(U0 + h * d, data_r) - roe.objective(U0 - h * d, data_r)) / (2 * h)
        errs_r.append(abs(fd - float(np.sum(g_r * d))) / max(abs(fd), 1e-300))
    roe.close()
    dom.close()
    dofs = 4 * m.ne * (p + 1) ** 2
    fwd_per_grad = st["forward_steps"] - nsteps * (1 + 2 * len(errs))
    return {"workload": f"nonlinear 2D Euler, {n}x{n} periodic quads, p={p}, initial-state gradient "
                        f"({dofs} unknowns), 16 pressure receivers, {nsteps} LSRK14 steps at CFL 0.5, "
                        f"{nsnap - 1} checkpoints + the initial state, FP64, Lax-Friedrichs flux",
            "forward_s": t_fwd, "gradient_s": t_grad, "gradient_over_forward": t_grad / t_fwd, "J": J,
            "forward_steps_per_gradient": fwd_per_grad, "adjoint_steps_per_gradient": nsteps,
            "forward_dof_stage_per_s": dofs * 14 * nsteps / t_fwd,
            "fd_directions": len(errs), "fd_rel_err_max": float(max(errs)) if errs else None,
            "fd_rel_err": [float(e) for e in errs],
            "roe": {"forward_s": t_fwd_r, "gradient_s": t_grad_r, "gradient_over_forward": t_grad_r / t_fwd_r,
                    "fd_rel_err_max": float(max(errs_r)) if errs_r else None}}


def curved_sample(args, seed=5):
    """The curved-bump workload: 3D Euler in an n^3 hex channel whose bottom wall
    carries a smooth bump of amplitude 0.05 (curved elements throughout,
    metrics at the nodes), p = 5, slip wall on the bump, the free stream imposed
    on the top boundary, periodic sides; free stream at Mach 0.5 plus seeded
    density noise of 1e-3.  LSRK14 in FP64 and in the mixed precision (FP32
    residual contractions, FP64 mass solve and state) from the same state;
    one warmup step, then `cv_steps` timed steps each, and the mixed state's
    deviation from the FP64 one."""
    n, p = args.cv_n, 5
    lo, hi = (0.0, 0.0, 0.0), (3.0, 3.0, 1.0)
    m = meshgen.channel_hexes(n, n, n, lo, hi, bc_bottom=dgx.DGX_BC_WALL, bc_top=dgx.DGX_BC_STATE)
    t0 = time.perf_counter()
    xn = meshgen.mapped_nodes(m, p, meshgen.bump_map(lo, hi, 0.05))
    rho0, u0, p0 = 1.0, 0.5, 1.0 / 1.4
    E0 = p0 / 0.4 + 0.5 * rho0 * u0 * u0
    rng = np.random.default_rng(seed)
    U = np.empty((5,) + xn.shape[1:])
    U[0] = rho0 + 1e-3 * rng.standard_normal(xn.shape[1:])
    U[1] = rho0 * u0
    U[2] = 0.0
    U[3] = 0.0
    U[4] = E0
    dt = 0.3 * (hi[2] / n) / ((p + 1) ** 2 * (u0 + 1.0))
    dofs = 5 * m.ne * (p + 1) ** 3
    res = {}
    for prec in (dgx.DGX_FP64, dgx.DGX_MIXED):
        t0 = time.perf_counter()
        gpu = dgx.GPUDomain(m, p, dgx.DGX_EULER3D, gamma=1.4, bc_state=[rho0, rho0 * u0, 0.0, 0.0, E0], xnodes=xn,
                            precision=prec)
        gpu.upload(U)
        gpu.synchronize()
        t_setup = time.perf_counter() - t0
        print(f"[bench] curved: {n}^3 p={p} precision {prec}: {dofs / 1e6:.1f} M DOF, setup {t_setup:.1f} s",
              file=sys.stderr, flush=True)
        gpu.advance(dgx.DGX_LSRK14, dt, 1)
        gpu.synchronize()
        t0 = time.perf_counter()
        done = 0
        while done < args.cv_steps:   # chunks of 20 steps, a progress line after each
            k = min(20, args.cv_steps - done)
            gpu.advance(dgx.DGX_LSRK14, dt, k)
            done += k
            if args.cv_steps > 20:
                gpu.synchronize()
                print(f"[bench] curved: {done} of {args.cv_steps} steps at {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
        gpu.synchronize()
        el = time.perf_counter() - t0
        finite = bool(np.all(np.isfinite(gpu.norms())))
        res[prec] = (el, t_setup, finite, gpu.solver_iterations(), gpu.download(dgx.DGX_MODAL))
        gpu.close()
    del xn, U
    el, t_setup, finite, cg_iters, U64 = res[dgx.DGX_FP64]
    elm, _, finite_m, _, U32 = res[dgx.DGX_MIXED]
    err_f = [float(np.linalg.norm(U32[v] - U64[v]) / np.linalg.norm(U64[v])) for v in range(5)]
    err_w = [float(np.linalg.norm(U32[v] - U64[v]) / np.linalg.norm(U64)) for v in range(5)]
    del res, U64, U32
    value = dofs * 14 * args.cv_steps / el
    # HBM bytes per DOF-stage: 32 for the 2N state streams, plus the per-node
    # metrics read once per stage (10 doubles per node over the 5 fields' modes)
    metric_b = 10 * 8 * (p + 2) ** 3 / (5 * (p + 1) ** 3)
    return {"workload": f"3D Euler, {n}^3 curved hexes (bump of amplitude 0.05 on a slip wall, free stream "
                        f"on top, periodic sides), p={p}, Mach 0.5 + 1e-3 seeded density noise, LSRK14, FP64",
            "value": value, "unit": "DOF-stage/s", "dofs": dofs, "steps": args.cv_steps, "dt": dt,
            "ms_per_step": el / args.cv_steps * 1e3, "setup_s": t_setup, "state_finite": finite,
            "mass_solve_cg_iterations_max": cg_iters,
            "mixed": {"value": dofs * 14 * args.cv_steps / elm, "speedup_vs_fp64": el / elm, "state_finite": finite_m,
                      "rel_l2_per_field_vs_fp64": err_f, "rel_l2_per_field_of_state_vs_fp64": err_w,
                      "after_steps": args.cv_steps + 1},
            "hbm_gbs_at_algorithmic_bytes": value * (32.0 + metric_b) / 1e9,
            "algorithmic_bytes_per_dof_stage": 32.0 + metric_b}


def curved_ns_sample(args, seed=6):
    """Navier-Stokes (BR1) on curved hexes: an n^3 hex channel whose bottom
    wall carries a bump of amplitude 0.05, isothermal walls at the bottom and
    top, periodic sides, p = 4, Re = 1000 on the channel height, Mach 0.5
    free stream plus seeded density noise of 1e-3, LSRK14, FP64: one warmup
    step, then `cvns_steps` timed steps."""
    n, p = args.cvns_n, 4
    lo, hi = (0.0, 0.0, 0.0), (3.0, 3.0, 1.0)
    m = meshgen.channel_hexes(n, n, n, lo, hi, bc_bottom=dgx.DGX_BC_ISOTHERMAL, bc_top=dgx.DGX_BC_ISOTHERMAL)
    xn = meshgen.mapped_nodes(m, p, meshgen.bump_map(lo, hi, 0.05))
    rho0, u0, p0 = 1.0, 0.5, 1.0 / 1.4
    E0 = p0 / 0.4 + 0.5 * rho0 * u0 * u0
    rng = np.random.default_rng(seed)
    U = np.empty((5,) + xn.shape[1:])
    U[0] = rho0 + 1e-3 * rng.standard_normal(xn.shape[1:])
    U[1], U[2], U[3], U[4] = rho0 * u0, 0.0, 0.0, E0
    mu = rho0 * u0 * (hi[2] - lo[2]) / 1000.0
    e0 = p0 / (rho0 * 0.4)
    wall = (e0, 1.5)   # Tw, Eref: the wall ghost's internal energy (Eref - 1/2) Tw is the free stream's e0
    dofs = 5 * m.ne * (p + 1) ** 3
    h = (hi[2] - lo[2]) / n
    dt = min(0.3 * h / ((p + 1) ** 2 * (u0 + 1.0)), 0.2 * rho0 * h * h / ((p + 1) ** 4 * mu))
    t0 = time.perf_counter()
    gpu = dgx.GPUDomain(m, p, dgx.DGX_NAVIER_STOKES3D, gamma=1.4, vis=mu, Pr=0.72, xnodes=xn, wall=wall,
                        bc_state=[rho0, rho0 * u0, 0.0, 0.0, E0])
    gpu.upload(U)
    gpu.synchronize()
    t_setup = time.perf_counter() - t0
    print(f"[bench] curved NS: {n}^3 p={p}: {dofs / 1e6:.1f} M DOF, setup {t_setup:.1f} s", file=sys.stderr, flush=True)
    gpu.advance(dgx.DGX_LSRK14, dt, 1)
    gpu.synchronize()
    t0 = time.perf_counter()
    gpu.advance(dgx.DGX_LSRK14, dt, args.cvns_steps)
    gpu.synchronize()
    el = time.perf_counter() - t0
    finite = bool(np.all(np.isfinite(gpu.norms())))
    cg = gpu.solver_iterations()
    gpu.close()
    return {"workload": f"3D Navier-Stokes (BR1, LF), {n}^3 curved hexes (bump of amplitude 0.05), isothermal "
                        f"walls at bottom and top, periodic sides, p={p}, Re=1000, Mach 0.5 + 1e-3 seeded density "
                        f"noise, LSRK14, FP64",
            "value": dofs * 14 * args.cvns_steps / el, "unit": "DOF-stage/s", "dofs": dofs,
            "steps": args.cvns_steps, "dt": dt, "ms_per_step": el / args.cvns_steps * 1e3, "setup_s": t_setup,
            "state_finite": finite, "mass_solve_cg_iterations_max": cg}


def element_block_sample(args, seed=11):
    """The element-block seam (GpGpu::hex_kernels: inner_product and
    backward_transform of a block of hexes sharing one basis) on the TGV mesh's
    element count and order: 32768 hexes, 8 modes and 9 GLL points per
    direction, device-resident slots (no PCIe), random basis and wJ.  HBM view:
    each transform reads and writes each element's slot once, (729 + 512) x 8 B
    per element."""
    from dgm_amd.hexblock import DeviceBuffer, HexKernels
    ne, nm, q, reps = args.hb_ne, 8, 9, args.hb_reps
    rng = np.random.default_rng(seed)
    hk = HexKernels().initialize(ne, nm, q, q, q)
    hk.Basis(*[rng.standard_normal((nm, q)) for _ in range(3)])
    hk.wJ(rng.uniform(0.5, 2.0, q ** 3))
    d = DeviceBuffer(ne * q ** 3)
    d.upload(rng.standard_normal(ne * q ** 3))
    out = {}
    for name, fn in (("inner_product", hk.inner_product), ("backward_transform", hk.backward_transform)):
        fn(d.address)
        hk.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn(d.address)
        hk.synchronize()
        el = (time.perf_counter() - t0) / reps
        nbytes = ne * (q ** 3 + nm ** 3) * 8
        out[name] = {"ms_per_call": el * 1e3, "hbm_gbs": nbytes / el / 1e9, "frac_of_hbm_peak": nbytes / el / 1e9 / HBM_PEAK_GBS}
    hk.close()
    d.free()
    out["workload"] = (f"{ne} hexes, {nm} modes and {q} GLL points per direction, FP64, device-resident slots, "
                       f"{reps} calls each; bytes per call = ne x (q^3 + nm^3) x 8")
    return out


def vortex_sample(args, seed=2024):
    """The 2D isentropic-vortex workload: periodic [0,10]^2 with 32x32 affine
    quads, p = 4, free stream rho = 1, u = 0.5, v = 0, p = 1/1.4 plus a
    Gaussian vortex of strength 5 centred at a seeded point, Lax-Friedrichs,
    RK4, dt from CFL 0.2 (dgx_set_time_step on the initial state), 100 steps;
    FP64, then mixed from the same state, and the mixed state's per-field
    relative L2 against the FP64 one."""
    n, p, steps = 32, 4, args.vortex_steps
    m = meshgen.periodic_quads(n, n, 0.0, 10.0, 0.0, 10.0)
    rng = np.random.default_rng(seed)
    x0, y0 = rng.uniform(0.0, 10.0, 2)
    res, states = {}, {}
    for prec in (dgx.DGX_FP64, dgx.DGX_MIXED):
        gpu = dgx.GPUDomain(m, p, dgx.DGX_EULER, gamma=1.4, precision=prec)
        x = gpu.coords()
        gpu.upload(ic.isentropic_vortex(x[0], x[1], x0, y0, period=(10.0, 10.0)))
        _, dt = gpu.set_time_step(1.0, CFLmax=0.2)
        gpu.advance(dgx.DGX_RK4, dt, 2)            # warmup, then the timed run from the same state
        gpu.upload(ic.isentropic_vortex(x[0], x[1], x0, y0, period=(10.0, 10.0)))
        gpu.synchronize()
        t0 = time.perf_counter()
        gpu.advance(dgx.DGX_RK4, dt, steps)
        gpu.synchronize()
        el = time.perf_counter() - t0
        dofs = 4 * m.ne * (p + 1) ** 2
        states[prec] = gpu.download(dgx.DGX_MODAL)
        res[prec] = {"value": dofs * 4 * steps / el, "ms_per_step": el / steps * 1e3, "dt": dt}
        gpu.close()
    a, b = states[dgx.DGX_MIXED], states[dgx.DGX_FP64]
    rel = [float(np.linalg.norm(a[v] - b[v]) / np.linalg.norm(b[v])) for v in range(4)]
    out = {"workload": f"2D Euler isentropic vortex, periodic [0,10]^2, {n}x{n} quads, p={p}, LF, RK4, dt from CFL 0.2, "
                       f"{steps} steps, {4 * m.ne * (p + 1) ** 2} DOF, FP64",
           "unit": "DOF-stage/s", **res[dgx.DGX_FP64]}
    out["mixed"] = {"value": res[dgx.DGX_MIXED]["value"], "speedup_vs_fp64": res[dgx.DGX_MIXED]["value"] /
                    res[dgx.DGX_FP64]["value"], "rel_l2_per_field_vs_fp64": rel}
    return out


def run_path(m, n, p, mu, precision, args, flux=None, viscous=None, compare=True):
    """Time one precision on the TGV workload; returns (result dict, final modal state)."""
    flux = flux or args.flux
    viscous = viscous or args.viscous
    gpu = dgx.GPUDomain(m, p, dgx.DGX_NAVIER_STOKES3D, gamma=1.4, vis=mu, Pr=0.71, precision=precision,
                        flux=FLUXES[flux], viscous=VISCOUS[viscous])
    x = gpu.coords()
    U = ic.taylor_green(x[0], x[1], x[2])
    del x
    gpu.upload(U)
    del U
    gpu.synchronize()
    dt = tgv_dt(n, p)
    dofs = 5 * m.ne * (p + 1) ** 3
    name = "mixed" if precision == dgx.DGX_MIXED else "fp64"
    print(f"[bench] {name}: TGV NS3d {n}^3 p={p}: {dofs / 1e6:.1f} M DOF, dt={dt:.3e}", file=sys.stderr,
          flush=True)
    gpu.advance(dgx.DGX_LSRK14, dt, args.warmup)
    gpu.synchronize()
    t0 = time.perf_counter()
    gpu.advance(dgx.DGX_LSRK14, dt, args.steps)
    gpu.synchronize()
    el = time.perf_counter() - t0
    value = dofs * 14 * args.steps / el
    print(f"[bench] {name}: timed {args.steps} steps in {el:.3f} s", file=sys.stderr, flush=True)

    # per-kernel times from HIP events on the library's stream (one more step)
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
    flops = FLOPS_PER_DOF_STAGE.get(p, {}).get(dom)
    if flops is not None and dom == "visc" and viscous == "br2":
        flops += BR2_VISC_FLOPS[p]
    roofline = None
    if flops is not None:
        algo_flops = flops * dofs
        achieved = algo_flops / mean_s / 1e12
        peak = FP64_PEAK_TFLOPS if precision == dgx.DGX_FP64 else FP32_MFMA_PEAK_TFLOPS
        # the FP64 kernels run every flop on the FP64 vector pipe (no MFMA); the
        # mixed kernels run their contractions on MFMA
        roofline = {"bound": "valu" if precision == dgx.DGX_FP64 else "mfma", "achieved": achieved, "peak": peak,
                    "unit": "TFLOP/s", "frac": achieved / peak,
                    "traffic": MEASURED_TRAFFIC.get((name, n, p, dom, flux, viscous)),
                    "traffic_source": "fixed calibration: rocprofv3 FETCH_SIZE + WRITE_SIZE of this kernel on this "
                                      "workload, scaled by a plain copy (profiles/tgv_ns3d_32cubed_p7_*.md), "
                                      "not measured in this run",
                    "kernel": dom, "arithmetic": ARITHMETIC[name],
                    "algorithmic_flops_per_launch": algo_flops, "ms_mean": kern[dom]["ms_mean"]}
    stage_s = sum(kern[k]["ms_total"] for k in kern) / 1e3 / (14 * prof_steps)
    algo_bytes = BYTES_PER_DOF_STAGE * dofs
    hbm = algo_bytes / stage_s / 1e9
    roofline_hbm = {"bound": "hbm", "achieved": hbm, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": hbm / HBM_PEAK_GBS, "algorithmic_bytes_per_stage": algo_bytes,
                    "ms_per_stage": stage_s * 1e3}
    # continue to the configuration's step count for the mixed-vs-FP64 comparison
    done = args.warmup + args.steps + prof_steps
    if compare and args.compare_steps > done:
        gpu.advance(dgx.DGX_LSRK14, dt, args.compare_steps - done)
        gpu.synchronize()
        done = args.compare_steps
    finite = bool(np.all(np.isfinite(gpu.norms())))
    state = gpu.download(dgx.DGX_MODAL)
    if args.copy_calibration:
        nbytes = 8 * dofs
        rc = gpu.lib.dgx_residual_download(gpu.h, ctypes.cast(gpu.device_state_ptr(),
                                           ctypes.POINTER(ctypes.c_double)), dgx.DGX_MODAL, 1)
        if rc != 0:
            raise RuntimeError(gpu.lib.dgx_last_error().decode())
        gpu.synchronize()
        print(f"[bench] calibration copy: {nbytes} bytes read and written", file=sys.stderr, flush=True)
    gpu.close()
    res = {"value": value, "ms_per_step": el / args.steps * 1e3, "dt": dt, "dofs": dofs,
           "state_finite": finite, "roofline": roofline, "roofline_hbm": roofline_hbm, "kernels": kern,
           "steps_run": done}
    return res, state


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=32, help="elements per direction")
    ap.add_argument("--p", type=int, default=7)
    ap.add_argument("--flux", choices=sorted(FLUXES), default="roe", help="Riemann flux")
    ap.add_argument("--viscous", choices=sorted(VISCOUS), default="br2", help="viscous face gradient")
    ap.add_argument("--precision", choices=["both", "fp64", "mixed"], default="both",
                    help="paths to run; with both, the mixed path is also compared with the FP64 one")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--compare-steps", type=int, default=500,
                    help="steps both paths run before their states are compared (the configuration's 500)")
    ap.add_argument("--no-reference-operator", action="store_true",
                    help="skip the FP64 timing of the reference's own operator (LF + BR1)")
    ap.add_argument("--cpu-steps", type=int, default=1, help="timed LSRK14 steps of the CPU baseline")
    ap.add_argument("--cpu-workers", type=int, default=0, help="CPU baseline workers (0: one per core, at most 16)")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop sample")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint (inverse-source) sample")
    ap.add_argument("--adj-n", type=int, default=64, help="quads per direction of the adjoint sample")
    ap.add_argument("--adj-steps", type=int, default=2000)
    ap.add_argument("--adj-dirs", type=int, default=8, help="finite-difference directions")
    ap.add_argument("--eadj-n", type=int, default=64, help="quads per direction of the Euler adjoint sample")
    ap.add_argument("--eadj-steps", type=int, default=200)
    ap.add_argument("--no-curved", action="store_true", help="skip the curved-bump sample")
    ap.add_argument("--cv-n", type=int, default=96, help="hexes per direction of the curved sample")
    ap.add_argument("--cv-steps", type=int, default=2)
    ap.add_argument("--no-hexblock", action="store_true", help="skip the element-block seam sample")
    ap.add_argument("--hb-ne", type=int, default=32768, help="elements of the element-block sample")
    ap.add_argument("--hb-reps", type=int, default=50)
    ap.add_argument("--serve-requests", type=int, default=64)
    ap.add_argument("--serve-nmax", type=int, default=64)
    ap.add_argument("--serve-steps", type=int, default=200)
    ap.add_argument("--serve-rate", type=float, default=32.0, help="arrivals per second")
    ap.add_argument("--serve-active", type=int, default=32)
    ap.add_argument("--serve-slice", type=int, default=20)
    ap.add_argument("--serve-unbatched", action="store_true", help="serve every request on its own launches")
    ap.add_argument("--serve-precision", choices=("mixed", "fp64"), default="mixed",
                    help="precision of the requests with --serve-only")
    ap.add_argument("--serve-persistent", action="store_true",
                    help="run each batched group slice as one persistent cooperative kernel")
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the serving sample and print its JSON (for profiling it alone)")
    ap.add_argument("--cvns-n", type=int, default=48, help="hexes per direction of the curved Navier-Stokes sample")
    ap.add_argument("--cvns-steps", type=int, default=5)
    ap.add_argument("--no-curved-ns", action="store_true", help="skip the curved Navier-Stokes sample")
    ap.add_argument("--vortex-steps", type=int, default=100, help="RK4 steps of the 2D vortex sample")
    ap.add_argument("--no-vortex", action="store_true", help="skip the 2D vortex sample")
    ap.add_argument("--only", choices=("serving", "adjoint", "euler_adjoint", "curved", "curved_ns", "element_block",
                                       "vortex"),
                    help="run only this sample and print its JSON (for profiling it alone)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with one device-to-device copy of the modal state (a plain copy "
                         "kernel of known size for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the engine has no multi-GPU path (a domain is one GPU's share; "
              "sharding it by element slabs with a trace halo exchange is not built); rerun with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.serve_only or args.only == "serving":
        print(json.dumps(serving_sample(args, precision=dgx.DGX_FP64 if args.serve_precision == "fp64"
                                        else dgx.DGX_MIXED)), flush=True)
        return
    if args.only:
        fn = {"adjoint": adjoint_sample, "euler_adjoint": euler_adjoint_sample, "curved": curved_sample,
              "curved_ns": curved_ns_sample, "element_block": element_block_sample, "vortex": vortex_sample}[args.only]
        print(json.dumps(fn(args)), flush=True)
        return

    n, p, mu = args.n, args.p, 1.0 / 1600.0
    # the CPU baseline runs the reference's own operator (LF + BR1)
    cpu = None if args.no_cpu else cpu_baseline(p, mu, n, args.cpu_workers, args.cpu_steps,
                                                dgx.DGX_FLUX_LF, dgx.DGX_VISCOUS_BR1)
    m = tgv_domain(n, p, mu)
    runs, states = {}, {}
    if args.precision in ("both", "fp64"):
        runs["fp64"], states["fp64"] = run_path(m, n, p, mu, dgx.DGX_FP64, args)
    if args.precision in ("both", "mixed"):
        runs["mixed"], states["mixed"] = run_path(m, n, p, mu, dgx.DGX_MIXED, args)

    head = runs.get("fp64") or runs["mixed"]
    head_name = "fp64" if "fp64" in runs else "mixed"
    line = {
        "metric": METRIC,
        "value": head["value"],
        "unit": "DOF-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head_name == "fp64" else "f32/f64",
        "data": "synthetic",
        "config": {"workload": f"3D Navier-Stokes Taylor-Green vortex, periodic [0,2pi]^3, {n}^3 hexes, "
                               f"p={p}, Re=1600, Ma=0.1, Pr=0.71, {OPNAME[args.flux]} + {args.viscous.upper()}, "
                               f"LSRK14, "
                               f"{'FP64' if head_name == 'fp64' else 'mixed'}",
                   "dofs": head["dofs"], "stages_per_step": 14, "dt": head["dt"], "path": head_name,
                   "state_finite": head["state_finite"]},
        "roofline": head["roofline"],
        "roofline_hbm": head["roofline_hbm"],
        "kernels": head["kernels"],
    }
    if "mixed" in runs and "fp64" in runs:
        mx = runs["mixed"]
        a, b = states["fp64"], states["mixed"]
        d = np.sqrt(np.sum((b - a) ** 2, axis=(1, 2)))
        r = np.sqrt(np.sum(a ** 2, axis=(1, 2)))
        mom = np.sqrt(np.sum(a[1:4] ** 2))
        rel = (d / r).tolist()
        ok = all(e <= 1e-6 for e in rel)
        line["mixed"] = {
            "value": mx["value"], "unit": "DOF-stage/s", "ms_per_step": mx["ms_per_step"],
            "speedup_vs_fp64": mx["value"] / runs["fp64"]["value"],
            "rel_err_l2_per_field": rel,
            "compared_after_steps": min(runs["fp64"]["steps_run"], mx["steps_run"]),
            "criterion": "relative L2 of every conserved field's modal state (the Domain::diff L2) <= 1e-6 "
                         "against the FP64 path after the configuration's step count",
            "criterion_met": ok,
            "diagnostic_rho_w_err_vs_momentum_norm": float(d[3] / mom),
            "roofline": mx["roofline"], "roofline_hbm": mx["roofline_hbm"], "kernels": mx["kernels"],
        }
    elif "mixed" in runs:
        line["mixed"] = {"value": runs["mixed"]["value"]}
    if not args.no_reference_operator and "fp64" in runs and (args.flux, args.viscous) != ("lf", "br1"):
        ref, _ = run_path(m, n, p, mu, dgx.DGX_FP64, args, flux="lf", viscous="br1", compare=False)
        line["reference_operator"] = {
            "operator": "LF + BR1 (the reference's Navier_Stokes3d residual), FP64",
            "value": ref["value"], "unit": "DOF-stage/s", "ms_per_step": ref["ms_per_step"],
            "vs_cpu_baseline": (ref["value"] / cpu["value"]) if cpu else None,
            "roofline": ref["roofline"], "kernels": ref["kernels"]}
    if not args.no_serve:
        line["serving"] = serving_sample(args, precision=dgx.DGX_MIXED)      # the ensemble mix is FP32-mixed
        line["serving_fp64"] = serving_sample(args, precision=dgx.DGX_FP64)
    if not args.no_adjoint:
        line["adjoint"] = adjoint_sample(args)
        line["euler_adjoint"] = euler_adjoint_sample(args)
    if not args.no_curved:
        line["curved"] = curved_sample(args)
    if not args.no_curved_ns:
        line["curved_ns"] = curved_ns_sample(args)
    if not args.no_hexblock:
        line["element_block"] = element_block_sample(args)
    if not args.no_vortex:
        line["vortex"] = vortex_sample(args)
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
