// Discrete adjoint of the nonlinear 2D Euler equations (dgx_euler_adjoint_*):
// gradients of a pressure-misfit objective with respect to the initial state,
// through LSRK14 time stepping of the DG residual (Euler::convection and the
// Lax-Friedrichs flux, src/Euler.cpp:296-757) on periodic affine quads.  The
// reverse of each step is LSRK14::adjoint_advance's recurrence
// (src/TimeInt.cpp:2064-2092) with the transposed residual Jacobian of the
// hand-written backward kernels (dgx_adjoint2d.hip), applied at the step's own
// stage states; forward states are recomputed from binomial checkpoints
// (revolve; cf. DTK::Griewank, src/DTK_Griewank.hpp:42-125).
//
// Forward stage s of step k, from the state U_s at its start:
//   Ut <- A_s Ut - dt R(U_s),  U_{s+1} = U_s + B_s Ut
// (R = -dU/dt, A_0 = 0).  With lamU, lamT the adjoints of U_{s+1} and Ut, the
// reversed stage is
//   g = lamT + B_s lamU;  lamU <- lamU - dt (dR/dU(U_s))^T g;  lamT = A_s g.
// The objective J = 1/2 dt sum_k sum_r (p_r(state_{k+1}) - d_{k,r})^2 adds
// its observation adjoint to lamU before step k is reversed; dJ/dU_0 is lamU
// after step 0.  Because the residual is nonlinear, reversing step k needs its
// 14 stage states: the forward advance of step k records them when step k is
// the next to be reversed (revolve always reverses a step right after
// advancing it).
#include "dgx_adjoint_common.hpp"

#include <algorithm>
#include <cmath>
#include <string>
#include <utility>
#include <vector>

namespace {

int fail(int code, const std::string &msg) { return dgx::set_error(code, msg); }

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

constexpr int kStages = 14;

}  // namespace

struct dgx_euler_adjoint {
  dgx_handle dom = nullptr;
  int p = 0, ne = 0, nm = 0, nrec = 0, nsteps = 0, nsnap = 0;
  double dt = 0.0;
  int *d_rec_e = nullptr, *d_grp = nullptr;   // d_grp: receivers grouped by element (ObsArgs)
  int ngrp = 0;
  double *d_rec_w = nullptr, *d_y = nullptr, *d_data = nullptr, *d_J = nullptr, *d_coef = nullptr;
  double *lamU = nullptr, *lamT = nullptr, *Lt = nullptr, *SB = nullptr;
  std::vector<double *> snaps;
  int rec_step = -1;       // the step whose stage states the next advance records
  int have_step = -1;      // the step whose stage states SB holds
  long long fwd_steps = 0, adj_steps = 0;
  int snaps_used = 0;

  size_t n() const { return (size_t)4 * ne * nm; }
};

namespace {

void release(dgx_euler_adjoint *a) {
  void *ptrs[] = {a->d_rec_e, a->d_grp, a->d_rec_w, a->d_y, a->d_data, a->d_J, a->d_coef,
                  a->lamU, a->lamT, a->Lt, a->SB};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (double *s : a->snaps)
    if (s) (void)hipFree(s);
  if (a->dom) dgx_domain_destroy(a->dom);
}

dgx::ObsArgs obs_args(dgx_euler_adjoint *a) {
  dgx::ObsArgs o;
  o.U = a->dom->U; o.rec_e = a->d_rec_e; o.rec_w = a->d_rec_w; o.nrec = a->nrec; o.ne = a->ne;
  o.L2 = (a->p + 1) * (a->p + 1); o.gamma = a->dom->gamma; o.stream = a->dom->stream;
  o.lam = a->lamU; o.J = a->d_J; o.coef = a->d_coef; o.dt = a->dt;
  o.grp_e = a->d_grp; o.grp_ptr = a->d_grp + a->ngrp; o.grp_r = a->d_grp + 2 * a->ngrp + 1; o.ngrp = a->ngrp;
  return o;
}

int traces(dgx_euler_adjoint *a, const double *U) {
  dgx_domain *d = a->dom;
  dgx::LaunchArgs la;
  la.op = dgx::OP_TRACES; la.p = a->p; la.ne = a->ne; la.nfield = 4; la.U = U; la.T = d->T;
  la.tab = d->tab; la.geo = d->geo; la.phys = d->phys; la.stream = d->stream;
  HIPCHK(dgx::launch_2d(la));
  return DGX_OK;
}

// one LSRK14 stage of the domain's state: the residual fused with the
// register update Ut = A Ut - dt R, U += B Ut and with the traces of the new
// state for the next stage (k_step2); stage 0 evaluates its own traces
int stage(dgx_euler_adjoint *a, int s) {
  dgx_domain *d = a->dom;
  int rc;
  if (s == 0 && (rc = traces(a, d->U))) return rc;
  dgx::LaunchArgs la;
  la.op = dgx::OP_STEP; la.p = a->p; la.ne = a->ne; la.nfield = 4; la.eqn = DGX_EULER;
  la.U = d->U; la.Us = d->U; la.Ut = d->Ut; la.T = d->T; la.Tn = d->T2; la.flag = d->d_flag;
  la.ca = dgx::kLsrkA[s]; la.cb = dgx::kLsrkB[s]; la.dt = a->dt;
  la.tab = d->tab; la.geo = d->geo; la.phys = d->phys; la.stream = d->stream;
  HIPCHK(dgx::launch_2d(la));
  std::swap(d->T, d->T2);
  return DGX_OK;
}

// out = (dR/dU(U))^T lam; or, with `s` >= 0, reversed LSRK stage s at the
// stage state U: g = lam + B_s lamU, lamU -= dt (dR/dU(U))^T g, lam = A_s g
// (lam = lamT, out = lamU).  Overwrites the domain's traces with those of U.
int vjp(dgx_euler_adjoint *a, const double *U, double *lam, double *out, int s) {
  dgx_domain *d = a->dom;
  int rc = traces(a, U);
  if (rc) return rc;
  dgx::AdjArgs aa;
  aa.p = a->p; aa.ne = a->ne; aa.U = U; aa.T = d->T; aa.lam = lam; aa.Lt = a->Lt; aa.out = out;
  aa.gamma = d->gamma; aa.flux = d->phys.flux; aa.geo = d->geo; aa.stream = d->stream;
  if (s >= 0) {
    aa.lamU = out; aa.B = dgx::kLsrkB[s]; aa.A = dgx::kLsrkA[s]; aa.stage = true;
    aa.scale = -a->dt; aa.accumulate = true;
  }
  aa.op = 0;
  HIPCHK(dgx::launch_adj2(aa));
  aa.op = 1;
  HIPCHK(dgx::launch_adj2(aa));
  return DGX_OK;
}

// forward step k: state_k (in the domain's U) -> state_{k+1}
int step_forward(dgx_euler_adjoint *a, int k) {
  dgx_domain *d = a->dom;
  const size_t n = a->n();
  const bool record = k == a->rec_step;
  for (int s = 0; s < kStages; ++s) {
    if (record) HIPCHK(hipMemcpyAsync(a->SB + s * n, d->U, n * 8, hipMemcpyDeviceToDevice, d->stream));
    int rc = stage(a, s);
    if (rc) return rc;
  }
  if (record) a->have_step = k;
  ++a->fwd_steps;
  return DGX_OK;
}

// reverse of step k, with state_{k+1} in the domain's U and step k's stage
// states in SB
int step_adjoint(dgx_euler_adjoint *a, int k) {
  dgx_domain *d = a->dom;
  if (a->have_step != k) return fail(DGX_ERR_ARGUMENT, "adjoint step " + std::to_string(k) + " has no stage states");
  const size_t n = a->n();
  if (a->nrec > 0) {
    dgx::ObsArgs o = obs_args(a);
    o.d = a->d_data + (size_t)k * a->nrec;
    HIPCHK(dgx::launch_pobs2_adj(o));
  }
  for (int s = kStages - 1; s >= 0; --s) {
    int rc = vjp(a, a->SB + s * n, a->lamT, a->lamU, s);
    if (rc) return rc;
  }
  a->have_step = -1;
  a->rec_step = k - 1;
  ++a->adj_steps;
  return DGX_OK;
}

int upload_state(dgx_euler_adjoint *a, const double *U0) {
  if (!U0) return fail(DGX_ERR_ARGUMENT, "null initial state");
  dgx_domain *d = a->dom;
  const size_t n = a->n();
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(U0[i])) return fail(DGX_ERR_ARGUMENT, "initial state is not finite");
  HIPCHK(hipMemcpyAsync(d->U, U0, n * 8, hipMemcpyHostToDevice, d->stream));
  HIPCHK(hipMemsetAsync(d->Ut, 0, n * 8, d->stream));
  HIPCHK(hipMemsetAsync(d->d_flag, 0, sizeof(int), d->stream));
  return DGX_OK;
}

// the residual kernels flag a density or pressure at or below eps^(2/3)
int check_flag(dgx_euler_adjoint *a) {
  dgx_domain *d = a->dom;
  int flag = 0;
  HIPCHK(hipMemcpyAsync(&flag, d->d_flag, sizeof(int), hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  if (flag) {
    HIPCHK(hipMemset(d->d_flag, 0, sizeof(int)));
    return fail(DGX_ERR_NONREALIZABLE, "density or pressure fell below aepsM in the forward solve");
  }
  return DGX_OK;
}

}  // namespace

extern "C" {

int dgx_euler_adjoint_create(const dgx_domain_desc *mesh, const dgx_euler_adjoint_desc *ad, void *stream,
                             dgx_euler_adjoint_handle *out) {
  if (!mesh || !ad || !out) return fail(DGX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (mesh->nsd != 2 || mesh->eqntype != DGX_EULER || mesh->vis != 0.0)
    return fail(DGX_ERR_ARGUMENT, "the Euler adjoint needs a 2D mesh with eqntype DGX_EULER and vis = 0");
  if (mesh->precision != DGX_FP64) return fail(DGX_ERR_ARGUMENT, "the Euler adjoint runs in DGX_FP64");
  if (mesh->flux != DGX_FLUX_LAX_FRIEDRICHS && mesh->flux != DGX_FLUX_ROE)
    return fail(DGX_ERR_ARGUMENT, "the Euler adjoint differentiates the Lax-Friedrichs and Roe fluxes");
  if (mesh->xnodes || mesh->bc_function || mesh->source || mesh->porder)
    return fail(DGX_ERR_ARGUMENT, "the Euler adjoint runs on affine quads of one order (no xnodes, functions "
                                  "or porder)");
  for (int r = 0; r < mesh->ncon; ++r)
    if (mesh->con && mesh->con[4 * r + 2] == -1)
      return fail(DGX_ERR_ARGUMENT, "the Euler adjoint runs on periodic meshes (no boundary rows)");
  if (!(ad->dt > 0) || !std::isfinite(ad->dt) || ad->nsteps < 1)
    return fail(DGX_ERR_ARGUMENT, "dt > 0 and nsteps >= 1 required");
  if (ad->nrec < 0 || (ad->nrec > 0 && !ad->xrec)) return fail(DGX_ERR_ARGUMENT, "bad receivers");
  if (ad->nsnap < 0) return fail(DGX_ERR_ARGUMENT, "nsnap must be >= 0");
  dgx_euler_adjoint *a = new dgx_euler_adjoint();
  int rc = dgx_domain_create(mesh, stream, &a->dom);
  if (rc) { delete a; return rc; }
  dgx_domain *d = a->dom;
  if (d->curved) {
    dgx_domain_destroy(d);
    delete a;
    return fail(DGX_ERR_ARGUMENT, "the Euler adjoint needs parallelogram quads (constant Jacobians)");
  }
  a->p = d->p; a->ne = d->ne; a->nm = d->nm; a->nrec = ad->nrec; a->nsteps = ad->nsteps; a->dt = ad->dt;
  std::vector<int> rec_e;
  std::vector<double> rec_w;
  const int outside = dgx::place_receivers(mesh, d, ad->xrec, a->nrec, rec_e, rec_w);
  if (outside >= 0) {
    release(a); delete a;
    return fail(DGX_ERR_ARGUMENT, "receiver " + std::to_string(outside) + " lies outside the mesh");
  }
  a->nsnap = ad->nsnap > 0 ? ad->nsnap : (int)std::ceil(std::sqrt((double)ad->nsteps)) + 1;
  const size_t n = a->n(), nobs = std::max<size_t>(1, (size_t)a->nsteps * a->nrec);
  const size_t nlt = (size_t)a->ne * 4 * 4 * (a->p + 1);
  auto fail_mem = [&](const char *what) {
    release(a); delete a;
    return fail(DGX_ERR_MEMORY, std::string("hipMalloc failed: ") + what);
  };
  // receivers grouped by element, in increasing order within each element:
  // [grp_e (ngrp)][grp_ptr (ngrp + 1)][grp_r (nrec)]
  std::vector<int> order(a->nrec);
  for (int r = 0; r < a->nrec; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return rec_e[x] < rec_e[y]; });
  std::vector<int> ge, gp;
  for (int k = 0; k < a->nrec; ++k)
    if (k == 0 || rec_e[order[k]] != rec_e[order[k - 1]]) {
      ge.push_back(rec_e[order[k]]);
      gp.push_back(k);
    }
  gp.push_back(a->nrec);
  a->ngrp = (int)ge.size();
  std::vector<int> grp(ge);
  grp.insert(grp.end(), gp.begin(), gp.end());
  grp.insert(grp.end(), order.begin(), order.end());
  if (hipMalloc((void **)&a->d_grp, std::max<size_t>(1, grp.size()) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&a->d_coef, std::max<size_t>(1, 5 * (size_t)a->nrec) * 8) != hipSuccess)
    return fail_mem("receiver groups");
  if (hipMalloc((void **)&a->d_rec_e, std::max(1, a->nrec) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&a->d_rec_w, std::max<size_t>(1, rec_w.size()) * 8) != hipSuccess ||
      hipMalloc((void **)&a->d_y, nobs * 8) != hipSuccess || hipMalloc((void **)&a->d_data, nobs * 8) != hipSuccess ||
      hipMalloc((void **)&a->d_J, 8) != hipSuccess || hipMalloc((void **)&a->lamU, n * 8) != hipSuccess ||
      hipMalloc((void **)&a->lamT, n * 8) != hipSuccess || hipMalloc((void **)&a->Lt, nlt * 8) != hipSuccess ||
      hipMalloc((void **)&a->SB, kStages * n * 8) != hipSuccess)
    return fail_mem("Euler adjoint buffers");
  a->snaps.assign(a->nsnap, nullptr);
  for (int s = 0; s < a->nsnap; ++s)
    if (hipMalloc((void **)&a->snaps[s], n * 8) != hipSuccess) return fail_mem("checkpoint slots");
  hipStream_t st = d->stream;
  if (a->nrec > 0) {
    (void)hipMemcpyAsync(a->d_rec_e, rec_e.data(), a->nrec * sizeof(int), hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(a->d_rec_w, rec_w.data(), rec_w.size() * 8, hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(a->d_grp, grp.data(), grp.size() * sizeof(int), hipMemcpyHostToDevice, st);
  }
  if (hipStreamSynchronize(st) != hipSuccess) { release(a); delete a; return fail(DGX_ERR_DEVICE, "upload"); }
  *out = a;
  return DGX_OK;
}

int dgx_euler_adjoint_destroy(dgx_euler_adjoint_handle a) {
  if (!a) return DGX_OK;
  (void)hipStreamSynchronize(a->dom->stream);
  release(a);
  delete a;
  return DGX_OK;
}

int dgx_euler_adjoint_forward(dgx_euler_adjoint_handle a, const double *U0, double *y, double *UT) {
  if (!a) return fail(DGX_ERR_ARGUMENT, "null handle");
  int rc;
  if ((rc = upload_state(a, U0))) return rc;
  dgx_domain *d = a->dom;
  a->rec_step = -1;
  for (int k = 0; k < a->nsteps; ++k) {
    if ((rc = step_forward(a, k))) return rc;
    if (a->nrec > 0) {
      dgx::ObsArgs o = obs_args(a);
      o.y = a->d_y + (size_t)k * a->nrec;
      HIPCHK(dgx::launch_pobs2(o));
    }
  }
  if (y && a->nrec > 0)
    HIPCHK(hipMemcpyAsync(y, a->d_y, (size_t)a->nsteps * a->nrec * 8, hipMemcpyDeviceToHost, d->stream));
  if (UT) HIPCHK(hipMemcpyAsync(UT, d->U, a->n() * 8, hipMemcpyDeviceToHost, d->stream));
  return check_flag(a);
}

int dgx_euler_adjoint_gradient(dgx_euler_adjoint_handle a, const double *U0, const double *data, double *J,
                               double *grad) {
  if (!a) return fail(DGX_ERR_ARGUMENT, "null handle");
  if (a->nrec > 0 && !data) return fail(DGX_ERR_ARGUMENT, "null data");
  int rc;
  if ((rc = upload_state(a, U0))) return rc;
  dgx_domain *d = a->dom;
  hipStream_t st = d->stream;
  const size_t n = a->n();
  if (a->nrec > 0)
    HIPCHK(hipMemcpyAsync(a->d_data, data, (size_t)a->nsteps * a->nrec * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(a->lamU, 0, n * 8, st));
  HIPCHK(hipMemsetAsync(a->lamT, 0, n * 8, st));
  HIPCHK(hipMemsetAsync(a->d_J, 0, 8, st));
  HIPCHK(hipMemcpyAsync(a->snaps[0], d->U, n * 8, hipMemcpyDeviceToDevice, st));
  a->rec_step = a->nsteps - 1;
  a->have_step = -1;
  dgx::Revolve rv;
  rv.restore = [&](int s) {
    if (hipMemcpyAsync(d->U, a->snaps[s], n * 8, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemsetAsync(d->Ut, 0, n * 8, st) != hipSuccess)
      return fail(DGX_ERR_DEVICE, "checkpoint restore");
    return (int)DGX_OK;
  };
  rv.store = [&](int s) {
    if (hipMemcpyAsync(a->snaps[s], d->U, n * 8, hipMemcpyDeviceToDevice, st) != hipSuccess)
      return fail(DGX_ERR_DEVICE, "checkpoint store");
    return (int)DGX_OK;
  };
  rv.advance = [&](int k) { return step_forward(a, k); };
  rv.reverse = [&](int k) { return step_adjoint(a, k); };
  if ((rc = rv.run(a->nsteps, a->nsnap))) return rc;
  a->snaps_used = rv.max_used;
  if (J) HIPCHK(hipMemcpyAsync(J, a->d_J, 8, hipMemcpyDeviceToHost, st));
  if (grad) HIPCHK(hipMemcpyAsync(grad, a->lamU, n * 8, hipMemcpyDeviceToHost, st));
  return check_flag(a);
}

int dgx_euler_adjoint_vjp(dgx_euler_adjoint_handle a, const double *U, const double *lam, double *out) {
  if (!a) return fail(DGX_ERR_ARGUMENT, "null handle");
  if (!U || !lam || !out) return fail(DGX_ERR_ARGUMENT, "null argument");
  dgx_domain *d = a->dom;
  hipStream_t st = d->stream;
  const size_t n = a->n();
  HIPCHK(hipMemcpyAsync(d->U, U, n * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(a->lamT, lam, n * 8, hipMemcpyHostToDevice, st));
  int rc = vjp(a, d->U, a->lamT, a->lamU, -1);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(out, a->lamU, n * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return DGX_OK;
}

int dgx_euler_adjoint_stats(dgx_euler_adjoint_handle a, long long *forward_steps, long long *adjoint_steps,
                            int *snapshots) {
  if (!a) return fail(DGX_ERR_ARGUMENT, "null handle");
  if (forward_steps) *forward_steps = a->fwd_steps;
  if (adjoint_steps) *adjoint_steps = a->adj_steps;
  if (snapshots) *snapshots = a->snaps_used;
  return DGX_OK;
}

}  // extern "C"
