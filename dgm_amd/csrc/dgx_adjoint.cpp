// Acoustic inverse-source problem on quads (dgx_acoustic_*): forward LSRK14
// solves of the linearized Euler equations about a uniform mean flow (the
// reference's LinEuler_quasi, src/LinEuler_quasi.cpp:151-495) and the discrete
// adjoint of the whole time loop, with forward states recomputed from
// binomial checkpoints (revolve; the reference keeps the same kind of record
// database in DTK::Griewank, src/DTK_Griewank.hpp:42-125, and reverses
// LSRK14 in LSRK14::adjoint_advance, src/TimeInt.cpp:2064-2092).
//
// The forward step k (k = 0 .. nsteps-1) maps state_k to state_{k+1} with
// F(U) = L U - g(t) b(theta) in each of the 14 stages; the objective is
// J = 1/2 dt sum_k sum_r (y_r(state_{k+1}) - d_{k,r})^2.  The reverse of step
// k needs state_{k+1} only (for the observation residual): the residual is
// linear, so its transpose does not depend on the forward state.
#include "dgx_adjoint_common.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <string>
#include <vector>

namespace {

using dgx::Revolve;

int fail(int code, const std::string &msg) { return dgx::set_error(code, msg); }

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(_e)); \
  } while (0)

double ricker(double t, double f0, double t0) {
  const double a = (M_PI * f0 * (t - t0)) * (M_PI * f0 * (t - t0));
  return (1.0 - 2.0 * a) * std::exp(-a);
}

}  // namespace

struct dgx_acoustic {
  dgx_handle dom = nullptr;
  int p = 0, ne = 0, nm = 0, nrec = 0, nsteps = 0, nsnap = 0;
  bool mixed = false;   // DGX_MIXED: FP32 contractions in k_lin2 and the traces
  double dt = 0, f0 = 0, t0 = 0, u0 = 0, v0 = 0, c0 = 0;
  double Ax[16], Ay[16], cobs[4];
  int *d_rec_e = nullptr;
  double *d_rec_w = nullptr, *d_cobs = nullptr, *d_theta = nullptr, *d_gtheta = nullptr, *d_y = nullptr;
  double *d_data = nullptr, *d_J = nullptr, *Ub = nullptr, *Ubt = nullptr;
  std::vector<double *> snaps;
  long long fwd_steps = 0, adj_steps = 0;
  int snaps_used = 0;

  size_t n() const { return (size_t)4 * ne * nm; }
};

namespace {

void release(dgx_acoustic *a) {
  void *ptrs[] = {a->d_rec_e, a->d_rec_w, a->d_cobs, a->d_theta, a->d_gtheta, a->d_y, a->d_data, a->d_J,
                  a->Ub, a->Ubt};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (double *s : a->snaps)
    if (s) (void)hipFree(s);
  if (a->dom) dgx_domain_destroy(a->dom);
}

dgx::LinArgs lin_args(dgx_acoustic *a, bool adjoint) {
  dgx_domain *d = a->dom;
  dgx::LinArgs la;
  la.ne = a->ne; la.p = a->p; la.adjoint = adjoint; la.mixed = a->mixed; la.tab = d->tab; la.geo = d->geo;
  la.stream = d->stream;
  la.u0 = a->u0; la.v0 = a->v0; la.c0 = a->c0; la.dt = a->dt;
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      la.Ax[4 * i + j] = adjoint ? -a->Ax[4 * j + i] : a->Ax[4 * i + j];
      la.Ay[4 * i + j] = adjoint ? -a->Ay[4 * j + i] : a->Ay[4 * i + j];
    }
  return la;
}

int traces(dgx_acoustic *a, const double *X) {
  dgx_domain *d = a->dom;
  dgx::LaunchArgs la;
  la.op = dgx::OP_TRACES; la.p = a->p; la.ne = a->ne; la.nfield = 4; la.U = X; la.T = d->T;
  la.tab = d->tab; la.geo = d->geo; la.stream = d->stream; la.mixed = a->mixed;
  HIPCHK(dgx::launch_2d(la));
  return DGX_OK;
}

// forward step k: state_k (in the domain's U) -> state_{k+1}
int step_forward(dgx_acoustic *a, int k) {
  dgx_domain *d = a->dom;
  dgx::LinArgs la = lin_args(a, false);
  la.X = d->U; la.T = d->T; la.Y = d->U; la.Yt = d->Ut; la.theta = a->d_theta;
  const double tk = k * a->dt;
  for (int s = 0; s < 14; ++s) {
    int rc = traces(a, d->U);
    if (rc) return rc;
    la.ca = dgx::kLsrkA[s]; la.cb = dgx::kLsrkB[s];
    la.g = ricker(tk + dgx::kLsrkC[s] * a->dt, a->f0, a->t0);
    HIPCHK(dgx::launch_lin2(la));
  }
  ++a->fwd_steps;
  return DGX_OK;
}

// reverse of step k, with state_{k+1} in the domain's U
int step_adjoint(dgx_acoustic *a, int k) {
  dgx_domain *d = a->dom;
  const int L2 = (a->p + 1) * (a->p + 1);
  const size_t n = a->n();
  if (a->nrec > 0)
    HIPCHK(dgx::launch_observe_adjoint(d->U, a->Ub, a->d_rec_e, a->d_rec_w, a->d_cobs, a->nrec, a->ne, L2,
                                       a->d_data + (size_t)k * a->nrec, a->dt, a->d_J, d->stream));
  dgx::LinArgs la = lin_args(a, true);
  la.X = a->Ubt; la.T = d->T; la.Y = a->Ub; la.Yt = a->Ubt; la.gtheta = a->d_gtheta;
  const double tk = k * a->dt;
  for (int s = 13; s >= 0; --s) {
    HIPCHK(dgx::blas_axpy(dgx::kLsrkB[s], a->Ub, a->Ubt, a->Ubt, n, d->stream));   // Ubar_t += B_s Ubar
    int rc = traces(a, a->Ubt);
    if (rc) return rc;
    la.ca = dgx::kLsrkA[s];
    la.g = ricker(tk + dgx::kLsrkC[s] * a->dt, a->f0, a->t0);
    HIPCHK(dgx::launch_lin2(la));
  }
  ++a->adj_steps;
  return DGX_OK;
}

int zero_state(dgx_acoustic *a) {
  dgx_domain *d = a->dom;
  HIPCHK(hipMemsetAsync(d->U, 0, a->n() * 8, d->stream));
  HIPCHK(hipMemsetAsync(d->Ut, 0, a->n() * 8, d->stream));
  return DGX_OK;
}

int upload_theta(dgx_acoustic *a, const double *theta) {
  if (!theta) return fail(DGX_ERR_ARGUMENT, "null theta");
  for (int e = 0; e < a->ne; ++e)
    if (!std::isfinite(theta[e])) return fail(DGX_ERR_ARGUMENT, "theta is not finite");
  HIPCHK(hipMemcpyAsync(a->d_theta, theta, (size_t)a->ne * 8, hipMemcpyHostToDevice, a->dom->stream));
  return DGX_OK;
}

}  // namespace

extern "C" {

int dgx_acoustic_create(const dgx_domain_desc *mesh, const dgx_acoustic_desc *ad, void *stream,
                        dgx_acoustic_handle *out) {
  if (!mesh || !ad || !out) return fail(DGX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (mesh->nsd != 2 || mesh->eqntype != DGX_LINEULER)
    return fail(DGX_ERR_ARGUMENT, "the acoustic problem needs a 2D mesh with eqntype DGX_LINEULER");
  if (mesh->precision != DGX_FP64 && mesh->precision != DGX_MIXED)
    return fail(DGX_ERR_ARGUMENT, "the acoustic problem runs in DGX_FP64 or DGX_MIXED");
  if (!(ad->rho0 > 0) || !(ad->p0 > 0) || !std::isfinite(ad->u0) || !std::isfinite(ad->v0))
    return fail(DGX_ERR_ARGUMENT, "mean flow needs rho0 > 0, p0 > 0 and finite velocity");
  if (!(ad->dt > 0) || ad->nsteps < 1) return fail(DGX_ERR_ARGUMENT, "dt > 0 and nsteps >= 1 required");
  if (ad->nrec < 0 || (ad->nrec > 0 && !ad->xrec)) return fail(DGX_ERR_ARGUMENT, "bad receivers");
  if (ad->nsnap < 0) return fail(DGX_ERR_ARGUMENT, "nsnap must be >= 0");
  dgx_domain_desc dd = *mesh;
  dd.eqntype = DGX_EULER;   // geometry, links, tables and work buffers of a 4-field quad domain
  dd.flux = DGX_FLUX_LAX_FRIEDRICHS;
  dd.vis = 0.0;
  dd.precision = DGX_FP64;   // the acoustic kernels take the precision from the problem
  if (mesh->xnodes || mesh->bc_function || mesh->source)
    return fail(DGX_ERR_ARGUMENT, "the acoustic problem runs on affine quads (no xnodes or functions)");
  for (int r = 0; r < mesh->ncon; ++r)
    if (mesh->con && mesh->con[4 * r + 2] == -1)
      return fail(DGX_ERR_ARGUMENT, "the acoustic problem runs on periodic meshes (no boundary rows)");
  dgx_acoustic *a = new dgx_acoustic();
  int rc = dgx_domain_create(&dd, stream, &a->dom);
  if (rc) { delete a; return rc; }
  dgx_domain *d = a->dom;
  if (d->curved) {   // a non-parallelogram quad: k_lin2 uses constant per-element metrics
    dgx_domain_destroy(d);
    delete a;
    return fail(DGX_ERR_ARGUMENT, "the acoustic problem needs parallelogram quads (constant Jacobians)");
  }
  a->p = d->p; a->ne = d->ne; a->nm = d->nm; a->nrec = ad->nrec; a->nsteps = ad->nsteps;
  a->mixed = mesh->precision == DGX_MIXED;
  a->dt = ad->dt; a->f0 = ad->f0; a->t0 = ad->t0; a->u0 = ad->u0; a->v0 = ad->v0;
  const double g = d->gamma, g1 = g - 1.0, rho0 = ad->rho0, u0 = ad->u0, v0 = ad->v0, p0 = ad->p0;
  a->c0 = std::sqrt(g * p0 / rho0);
  // Jacobians of the 2D Euler flux at the uniform state (conservative variables)
  const double q2 = u0 * u0 + v0 * v0, H = (p0 / g1 + 0.5 * rho0 * q2 + p0) / rho0;
  const double Ax[16] = {0.0, 1.0, 0.0, 0.0,
                         0.5 * g1 * q2 - u0 * u0, (3.0 - g) * u0, -g1 * v0, g1,
                         -u0 * v0, v0, u0, 0.0,
                         u0 * (0.5 * g1 * q2 - H), H - g1 * u0 * u0, -g1 * u0 * v0, g * u0};
  const double Ay[16] = {0.0, 0.0, 1.0, 0.0,
                         -u0 * v0, v0, u0, 0.0,
                         0.5 * g1 * q2 - v0 * v0, -g1 * u0, (3.0 - g) * v0, g1,
                         v0 * (0.5 * g1 * q2 - H), -g1 * u0 * v0, H - g1 * v0 * v0, g * v0};
  std::copy(Ax, Ax + 16, a->Ax);
  std::copy(Ay, Ay + 16, a->Ay);
  // pressure perturbation p' = (gamma-1) (E' - u0 m' - v0 n' + |u0|^2/2 rho')
  a->cobs[0] = g1 * 0.5 * q2; a->cobs[1] = -g1 * u0; a->cobs[2] = -g1 * v0; a->cobs[3] = g1;
  // receivers: element and modal weights phi_l(xi) phi_m(eta)
  std::vector<int> rec_e;
  std::vector<double> rec_w;
  const int outside = dgx::place_receivers(mesh, d, ad->xrec, a->nrec, rec_e, rec_w);
  if (outside >= 0) {
    release(a); delete a;
    return fail(DGX_ERR_ARGUMENT, "receiver " + std::to_string(outside) + " lies outside the mesh");
  }
  a->nsnap = ad->nsnap > 0 ? ad->nsnap : (int)std::ceil(std::sqrt((double)ad->nsteps)) + 1;
  const size_t n = a->n();
  auto fail_mem = [&](const char *what) {
    release(a); delete a;
    return fail(DGX_ERR_MEMORY, std::string("hipMalloc failed: ") + what);
  };
  if (hipMalloc((void **)&a->d_rec_e, std::max(1, a->nrec) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&a->d_rec_w, std::max<size_t>(1, rec_w.size()) * 8) != hipSuccess ||
      hipMalloc((void **)&a->d_cobs, 4 * 8) != hipSuccess || hipMalloc((void **)&a->d_theta, a->ne * 8) != hipSuccess ||
      hipMalloc((void **)&a->d_gtheta, a->ne * 8) != hipSuccess ||
      hipMalloc((void **)&a->d_y, std::max<size_t>(1, (size_t)a->nsteps * a->nrec) * 8) != hipSuccess ||
      hipMalloc((void **)&a->d_data, std::max<size_t>(1, (size_t)a->nsteps * a->nrec) * 8) != hipSuccess ||
      hipMalloc((void **)&a->d_J, 8) != hipSuccess || hipMalloc((void **)&a->Ub, n * 8) != hipSuccess ||
      hipMalloc((void **)&a->Ubt, n * 8) != hipSuccess)
    return fail_mem("acoustic buffers");
  a->snaps.assign(a->nsnap, nullptr);
  for (int s = 0; s < a->nsnap; ++s)
    if (hipMalloc((void **)&a->snaps[s], n * 8) != hipSuccess) return fail_mem("checkpoint slots");
  hipStream_t st = d->stream;
  if (a->nrec > 0) {
    (void)hipMemcpyAsync(a->d_rec_e, rec_e.data(), a->nrec * sizeof(int), hipMemcpyHostToDevice, st);
    (void)hipMemcpyAsync(a->d_rec_w, rec_w.data(), rec_w.size() * 8, hipMemcpyHostToDevice, st);
  }
  (void)hipMemcpyAsync(a->d_cobs, a->cobs, 4 * 8, hipMemcpyHostToDevice, st);
  if (hipStreamSynchronize(st) != hipSuccess) { release(a); delete a; return fail(DGX_ERR_DEVICE, "upload"); }
  *out = a;
  return DGX_OK;
}

int dgx_acoustic_destroy(dgx_acoustic_handle a) {
  if (!a) return DGX_OK;
  (void)hipStreamSynchronize(a->dom->stream);
  release(a);
  delete a;
  return DGX_OK;
}

int dgx_acoustic_forward(dgx_acoustic_handle a, const double *theta, double *y) {
  if (!a) return fail(DGX_ERR_ARGUMENT, "null handle");
  int rc;
  if ((rc = upload_theta(a, theta)) || (rc = zero_state(a))) return rc;
  dgx_domain *d = a->dom;
  const int L2 = (a->p + 1) * (a->p + 1);
  for (int k = 0; k < a->nsteps; ++k) {
    if ((rc = step_forward(a, k))) return rc;
    if (a->nrec > 0)
      HIPCHK(dgx::launch_observe(d->U, a->d_rec_e, a->d_rec_w, a->d_cobs, a->nrec, a->ne, L2,
                                 a->d_y + (size_t)k * a->nrec, d->stream));
  }
  if (y && a->nrec > 0)
    HIPCHK(hipMemcpyAsync(y, a->d_y, (size_t)a->nsteps * a->nrec * 8, hipMemcpyDeviceToHost, d->stream));
  HIPCHK(hipStreamSynchronize(d->stream));
  return DGX_OK;
}

int dgx_acoustic_gradient(dgx_acoustic_handle a, const double *theta, const double *data, double *J,
                          double *grad) {
  if (!a) return fail(DGX_ERR_ARGUMENT, "null handle");
  if (a->nrec > 0 && !data) return fail(DGX_ERR_ARGUMENT, "null data");
  int rc;
  if ((rc = upload_theta(a, theta)) || (rc = zero_state(a))) return rc;
  dgx_domain *d = a->dom;
  hipStream_t st = d->stream;
  const size_t n = a->n();
  if (a->nrec > 0)
    HIPCHK(hipMemcpyAsync(a->d_data, data, (size_t)a->nsteps * a->nrec * 8, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(a->Ub, 0, n * 8, st));
  HIPCHK(hipMemsetAsync(a->Ubt, 0, n * 8, st));
  HIPCHK(hipMemsetAsync(a->d_gtheta, 0, (size_t)a->ne * 8, st));
  HIPCHK(hipMemsetAsync(a->d_J, 0, 8, st));
  HIPCHK(hipMemcpyAsync(a->snaps[0], d->U, n * 8, hipMemcpyDeviceToDevice, st));
  Revolve rv;
  rv.restore = [&](int s) {
    if (hipMemcpyAsync(d->U, a->snaps[s], n * 8, hipMemcpyDeviceToDevice, st) != hipSuccess)
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
  if (grad) HIPCHK(hipMemcpyAsync(grad, a->d_gtheta, (size_t)a->ne * 8, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return DGX_OK;
}

int dgx_acoustic_stats(dgx_acoustic_handle a, long long *forward_steps, long long *adjoint_steps, int *snapshots) {
  if (!a) return fail(DGX_ERR_ARGUMENT, "null handle");
  if (forward_steps) *forward_steps = a->fwd_steps;
  if (adjoint_steps) *adjoint_steps = a->adj_steps;
  if (snapshots) *snapshots = a->snaps_used;
  return DGX_OK;
}

int dgx_revolve_count(int nsteps, int nsnap, long long *advances, int *max_slots, int *order_ok) {
  if (nsteps < 1 || nsnap < 1) return fail(DGX_ERR_ARGUMENT, "nsteps and nsnap must be >= 1");
  // simulate the schedule on step indices: the "state" is the index it holds
  std::vector<int> slot(nsnap, -1);
  int state = 0, expect = nsteps - 1, bad = 0;
  slot[0] = 0;
  Revolve rv;
  rv.restore = [&](int s) { state = slot[s]; return 0; };
  rv.store = [&](int s) { slot[s] = state; return 0; };
  rv.advance = [&](int k) { if (state != k) bad = 1; state = k + 1; return 0; };
  rv.reverse = [&](int k) { if (k != expect || state != k + 1) bad = 1; --expect; return 0; };
  rv.run(nsteps, nsnap);
  if (advances) *advances = rv.advances;
  if (max_slots) *max_slots = rv.max_used;
  if (order_ok) *order_ok = (!bad && expect == -1) ? 1 : 0;
  return DGX_OK;
}

}  // extern "C"
