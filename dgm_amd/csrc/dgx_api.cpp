// Host engine and C ABI (include/dgx.h): mesh tables, affine geometry, side
// links and orientation flips, basis tables, device buffers, and the explicit
// Runge-Kutta drivers that chain the residual kernels on one HIP stream.
#include "../../include/dgx.h"
#include "dgx_domain.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <limits>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <thread>
#include <vector>

namespace {
thread_local std::string g_err;
}  // namespace

namespace dgx {
// Error message of the calling thread's last failure (dgx_last_error).
int set_error(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK_X(expr)                                                                \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      return set_error(DGX_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)
// Largest idt of Euler::set_time_step over the domain's nodes (dgx_cfl.hip),
// with `part` as the kernels' scratch (ne + 1 doubles); waits for the stream.
int cfl_max(dgx_domain *d, double *part, double *m) {
  auto fail = [](int c, const std::string &msg) { return set_error(c, msg); };
  if (!d->d_cgeo)
    return fail(DGX_ERR_ARGUMENT, "the CFL time step is available for 2D Euler, Navier_Stokes and Shallow domains");
  dgx::CflArgs a;
  a.U = d->U; a.cgeo = d->d_cgeo; a.part = part; a.out = part + d->ne;
  for (int i = 0; i < d->Q; ++i) {   // GLL spacings (src/Element.cpp:513-527)
    const std::vector<double> &z = d->hb.z;
    a.dz[i] = i == 0 ? z[1] - z[0] : (i == d->Q - 1 ? z[i] - z[i - 1] : 0.5 * (z[i + 1] - z[i - 1]));
  }
  a.gamma = d->gamma; a.ne = d->ne; a.p = d->p;
  a.shallow = d->eqn == DGX_SHALLOW; a.curved = d->curved; a.stream = d->stream;
  hipError_t e = hipSuccess;
  if (d->varp) {   // each order group with its own GLL spacings, then one reduction
    for (size_t g = 0; g < d->vgrp_p.size() && e == hipSuccess; ++g) {
      const int P = d->vgrp_p[g];
      for (int i = 0; i < P + 2; ++i) a.dz[i] = d->vgrp_dz[g * (dgx::kMaxOrder + 2) + i];
      a.p = P; a.elems = d->d_velems + d->vgrp_off[g]; a.nel = d->vgrp_n[g]; a.NM = d->nm;
      e = dgx::launch_cfl2(a);
    }
    if (e == hipSuccess) e = dgx::launch_cfl_max1(part, d->ne, a.out, d->stream);
  } else {
    e = dgx::launch_cfl2(a);
  }
  if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("CFL kernels: ") + hipGetErrorString(e));
  HIPCHK_X(hipMemcpyAsync(m, a.out, sizeof(double), hipMemcpyDeviceToHost, d->stream));
  HIPCHK_X(hipStreamSynchronize(d->stream));
  if (!(*m == *m)) return fail(DGX_ERR_NONREALIZABLE, "non-realizable state in the CFL evaluation");
  return DGX_OK;
}

// The time step rule of Euler::set_time_step (src/Euler.cpp:861-874) from the
// largest idt: CFL of dt, new dt under CFLmax (and dtmax), CFL of the new dt.
int cfl_finish(double dt, double CFLmax, double dtmax, double m, double *cfl, double *dt_new) {
  const double c = dt * m;
  double dn = dt;
  if (CFLmax != 0 && dtmax != 0) dn = std::min(dtmax, CFLmax * dt / c);
  else if (CFLmax != 0) dn = CFLmax * dt / c;
  *dt_new = dn;
  *cfl = c * dn / dt;
  return DGX_OK;
}

#undef HIPCHK_X
}  // namespace dgx

namespace {

int fail(int code, const std::string &msg) { return dgx::set_error(code, msg); }

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess)                                                             \
      return fail(DGX_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

typedef long double ld;

// ---------------------------------------------------------------------------
// 1D quadrature and basis tables from Legendre recurrences in long double.
// GLL nodes: -1, +1 and the roots of P'_{q-1}; Gauss nodes: roots of P_{qf}
// (Karniadakis & Sherwin, Appendix B; src/Polylib.tpp restates the same).
void leg(int n, ld z, ld &P, ld &dP) {
  ld p0 = 1, p1 = z;
  if (n == 0) { P = 1; dP = 0; return; }
  for (int k = 1; k < n; ++k) {
    ld p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
    p0 = p1; p1 = p2;
  }
  P = p1;
  if (fabsl(z * z - 1) < 1e-30L) dP = (z > 0 ? 1 : ((n % 2) ? 1 : -1)) * (ld)n * (n + 1) / 2;
  else dP = n * (z * p1 - p0) / (z * z - 1);
}

void gll_nodes(int q, std::vector<ld> &z, std::vector<ld> &w) {
  const int N = q - 1;
  z.assign(q, 0); w.assign(q, 0);
  z[0] = -1; z[N] = 1;
  for (int i = 1; i < N; ++i) {
    ld x = -cosl(M_PIl * i / N);
    for (int it = 0; it < 100; ++it) {
      ld P, dP; leg(N, x, P, dP);
      const ld d2 = (2 * x * dP - N * (N + 1) * P) / (1 - x * x);
      const ld dx = dP / d2;
      x -= dx;
      if (fabsl(dx) < 1e-30L) break;
    }
    z[i] = x;
  }
  for (int i = 0; i < q; ++i) { ld P, dP; leg(N, z[i], P, dP); w[i] = 2.0L / (N * (N + 1) * P * P); }
}

void gauss_nodes(int n, std::vector<ld> &z, std::vector<ld> &w) {
  z.assign(n, 0); w.assign(n, 0);
  for (int i = 0; i < n; ++i) {
    ld x = -cosl(M_PIl * (4 * i + 3) / (4 * n + 2));
    for (int it = 0; it < 100; ++it) {
      ld P, dP; leg(n, x, P, dP);
      const ld dx = P / dP;
      x -= dx;
      if (fabsl(dx) < 1e-30L) break;
    }
    ld P, dP; leg(n, x, P, dP);
    z[i] = x; w[i] = 2.0L / ((1 - x * x) * dP * dP);
  }
}

ld lagrange(const std::vector<ld> &nodes, int j, ld x) {
  ld h = 1;
  for (size_t m = 0; m < nodes.size(); ++m)
    if ((int)m != j) h *= (x - nodes[m]) / (nodes[j] - nodes[m]);
  return h;
}

ld normleg(int l, ld x) { ld P, dP; leg(l, x, P, dP); return P * sqrtl((2 * l + 1) / 2.0L); }

using dgx::HostBasis;

HostBasis make_basis(int p) {
  HostBasis b;
  b.p = p; b.L = p + 1; b.Q = p + 2; b.QF = p + 1;
  const int L = b.L, Q = b.Q, QF = b.QF;
  std::vector<ld> z, w, zg, wg;
  gll_nodes(Q, z, w);
  gauss_nodes(QF, zg, wg);
  const int N = Q - 1;
  std::vector<ld> PN(Q), Dl(Q * Q);
  for (int i = 0; i < Q; ++i) { ld dP; leg(N, z[i], PN[i], dP); }
  for (int i = 0; i < Q; ++i)
    for (int j = 0; j < Q; ++j) {
      ld v = 0;
      if (i != j) v = PN[i] / (PN[j] * (z[i] - z[j]));
      else if (i == 0) v = -(ld)N * (N + 1) / 4;
      else if (i == N) v = (ld)N * (N + 1) / 4;
      Dl[i * Q + j] = v;
    }
  b.z.resize(Q); b.w.resize(Q); b.D.resize(Q * Q);
  for (int i = 0; i < Q; ++i) { b.z[i] = (double)z[i]; b.w[i] = (double)w[i]; }
  for (int i = 0; i < Q * Q; ++i) b.D[i] = (double)Dl[i];
  b.BT.resize(Q * L); b.BW.resize(L * Q); b.QD.resize(L * Q);
  b.Bg.resize(QF * L); b.Bend.resize(2 * L); b.PW.resize(L * QF); b.Ig.resize(QF * Q);
  for (int l = 0; l < L; ++l) {
    for (int i = 0; i < Q; ++i) {
      b.BT[i * L + l] = (double)normleg(l, z[i]);
      b.BW[l * Q + i] = (double)(normleg(l, z[i]) * w[i]);
    }
    for (int k = 0; k < Q; ++k) {
      ld s = 0;
      for (int i = 0; i < Q; ++i) s += normleg(l, z[i]) * w[i] * Dl[i * Q + k];
      b.QD[l * Q + k] = (double)s;
    }
    for (int a = 0; a < QF; ++a) {
      b.Bg[a * L + l] = (double)normleg(l, zg[a]);
      b.PW[l * QF + a] = (double)(wg[a] * normleg(l, zg[a]));
    }
    b.Bend[l] = (double)normleg(l, -1.0L);
    b.Bend[L + l] = (double)normleg(l, 1.0L);
  }
  for (int a = 0; a < QF; ++a)
    for (int i = 0; i < Q; ++i) b.Ig[a * Q + i] = (double)lagrange(z, i, zg[a]);
  b.PI.resize(L * Q); b.KE.resize(2 * Q); b.BL.resize(L * Q); b.GL.resize(Q * QF);
  for (int l = 0; l < L; ++l)
    for (int i = 0; i < Q; ++i) b.BL[l * Q + i] = (double)normleg(l, z[i]);
  for (int i = 0; i < Q; ++i)
    for (int a = 0; a < QF; ++a) {
      ld s = 0;
      for (int l = 0; l < L; ++l) s += normleg(l, z[i]) * wg[a] * normleg(l, zg[a]);
      b.GL[i * QF + a] = (double)s;
    }
  for (int l = 0; l < L; ++l)
    for (int i = 0; i < Q; ++i) {
      ld s = 0;
      for (int a = 0; a < QF; ++a) s += wg[a] * normleg(l, zg[a]) * lagrange(z, i, zg[a]);
      b.PI[l * Q + i] = (double)s;
    }
  for (int i = 0; i < Q; ++i) {
    ld lo = 0, hi = 0;
    for (int l = 0; l < L; ++l) {
      lo += normleg(l, -1.0L) * normleg(l, z[i]);
      hi += normleg(l, 1.0L) * normleg(l, z[i]);
    }
    b.KE[i] = (double)lo;
    b.KE[Q + i] = (double)hi;
  }
  // NP1 = BT BW and NPD = BT QD: projection (and projected derivative) from
  // GLL values back to GLL values, in long double
  b.NP1.resize(Q * Q); b.NPD.resize(Q * Q);
  for (int i = 0; i < Q; ++i)
    for (int m = 0; m < Q; ++m) {
      ld sp = 0, sd = 0;
      for (int l = 0; l < L; ++l) {
        ld qd = 0;
        for (int j = 0; j < Q; ++j) qd += normleg(l, z[j]) * w[j] * Dl[j * Q + m];
        sp += normleg(l, z[i]) * normleg(l, z[m]) * w[m];
        sd += normleg(l, z[i]) * qd;
      }
      b.NP1[i * Q + m] = (double)sp;
      b.NPD[i * Q + m] = (double)sd;
    }
  return b;
}

// Basis tables of all orders, packed as the kernels' constant memory expects
// (dgx_device.hpp: tab_offset), uploaded once per process.
int upload_all_tables() {
  static int status = 1;
  if (status == 0) return DGX_OK;
  std::vector<double> packed(dgx::kTabTotal, 0.0);
  for (int p = 1; p <= dgx::kMaxOrder; ++p) {
    HostBasis b = make_basis(p);
    size_t off = dgx::tab_offset(p);
    for (const std::vector<double> *t : {&b.BT, &b.BW, &b.QD, &b.Bg, &b.Bend, &b.PW, &b.Ig, &b.PI, &b.KE, &b.BL, &b.GL,
                                         &b.D, &b.NP1, &b.NPD}) {
      std::copy(t->begin(), t->end(), packed.begin() + off);
      off += t->size();
    }
    if (off != (size_t)dgx::tab_offset(p) + dgx::tab_size(p)) return fail(DGX_ERR_ARGUMENT, "table layout");
  }
  hipError_t e;
  if ((e = dgx::upload_tables_2d(packed.data())) != hipSuccess ||
      (e = dgx::upload_tables_3d(packed.data())) != hipSuccess ||
      (e = dgx::upload_tables_s3d(packed.data())) != hipSuccess ||
      (e = dgx::upload_tables_m3d(packed.data())) != hipSuccess ||
      (e = dgx::upload_tables_c3d(packed.data())) != hipSuccess ||
      (e = dgx::upload_tables_c2d(packed.data())) != hipSuccess ||
      (e = dgx::upload_tables_cfl(packed.data())) != hipSuccess ||
      (e = dgx::upload_tables_v2d(packed.data())) != hipSuccess ||
      (e = dgx::upload_tables_a2d(packed.data())) != hipSuccess)
    return fail(DGX_ERR_DEVICE, std::string("basis table upload: ") + hipGetErrorString(e));
  status = 0;
  return DGX_OK;
}

// ---------------------------------------------------------------------------
struct Side {
  bool master = false;
  int le = -1, ls = -1;
  int bc = -1;   // boundary side: DGX_BC_WALL or DGX_BC_STATE
  bool fa = false, fb = false, tr = false;
};

}  // namespace



namespace {

int alloc(double **p, size_t n) {
  if (hipMalloc((void **)p, n * sizeof(double) + 64) != hipSuccess)
    return fail(DGX_ERR_MEMORY, "hipMalloc of " + std::to_string(n * 8) + " bytes failed");
  return DGX_OK;
}

// Record a start event before and a stop event after `fn` when profiling.
template <typename Fn>
hipError_t timed(dgx_domain *d, int kind, Fn fn) {
  if (!d->profiling) return fn();
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return hipErrorOutOfMemory;
  (void)hipEventRecord(a, d->stream);
  hipError_t e = fn();
  (void)hipEventRecord(b, d->stream);
  d->ev.push_back(a);
  d->ev.push_back(b);
  d->ev_kind.push_back(kind);
  return e;
}

hipError_t launch_varp(const dgx_domain *d, int op, const double *in, double *out);

dgx::LaunchArgs launch_args(dgx_domain *d, int op, const double *U, double *out) {
  dgx::LaunchArgs a;
  a.op = op; a.p = d->p; a.ne = d->ne; a.nfield = d->nf; a.ns = d->visc;
  a.U = U; a.T = d->T; a.R = out; a.G = d->G; a.V = d->V; a.TV = d->TV;
  a.flag = d->d_flag; a.tab = d->tab; a.geo = d->geo; a.phys = d->phys; a.stream = d->stream;
  a.eqn = d->eqn; a.Bal = d->Bal; a.bal = (op == dgx::OP_RHS || op == dgx::OP_STEP) ? d->bal : 0;
  a.mixed = d->mixed && d->nsd == 2;
  return a;
}

int launch(dgx_domain *d, int op, const double *U, double *out) {
  dgx::LaunchArgs a = launch_args(d, op, U, out);
  hipError_t e = timed(d, op, [&] {
    if (d->varp) return launch_varp(d, op, U, out);
    return (d->nsd == 2) ? dgx::launch_2d(a) : dgx::launch_3d(a);
  });
  if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  if (op == dgx::OP_RHS && d->bal == 1) d->bal = 2;   // the first residual is stored
  return DGX_OK;
}

dgx::CurvedArgs curved_args(dgx_domain *d) {
  dgx::CurvedArgs ca;
  ca.gm = d->d_gm; ca.fgeo = d->d_fgeo; ca.jav = d->d_jav; ca.pre = d->d_pre; ca.geo = d->geo;
  ca.chol = d->nsd == 2 ? d->d_chol : nullptr;
  ca.phys = d->phys; ca.ne = d->ne; ca.p = d->p; ca.flag = d->d_flag; ca.iters = d->d_iters;
  ca.shallow = d->eqn == DGX_SHALLOW; ca.stream = d->stream;
  ca.ns = d->visc; ca.G = d->G; ca.V = d->V; ca.TV = d->TV;
  ca.fn = d->fn; ca.t = d->time;
  ca.mixed = d->mixed && d->nsd == 3;
  return ca;
}

hipError_t launch_curved(const dgx_domain *d, const dgx::CurvedArgs &ca) {
  return d->nsd == 2 ? dgx::launch_curved2(ca) : dgx::launch_curved3(ca);
}

int transform(dgx_domain *d, int op, const double *x, double *y) {
  if (d->curved && op == dgx::OP_FORWARD) {   // projection through the curved mass matrix
    dgx::CurvedArgs ca = curved_args(d);
    ca.U = x; ca.R = y; ca.project = true;
    hipError_t e = launch_curved(d, ca);
    if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("curved projection: ") + hipGetErrorString(e));
    return DGX_OK;
  }
  dgx::LaunchArgs a;
  a.op = op; a.p = d->p; a.ne = d->ne; a.nfield = d->nf;
  a.X = x; a.Y = y; a.tab = d->tab; a.stream = d->stream;
  hipError_t e = d->varp ? launch_varp(d, op, x, y) : (d->nsd == 2) ? dgx::launch_2d(a) : dgx::launch_3d(a);
  if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("transform: ") + hipGetErrorString(e));
  return DGX_OK;
}

dgx::StageLaunch stage_args(dgx_domain *d) {
  dgx::StageLaunch s;
  s.p = d->p; s.ne = d->ne; s.ns = d->visc; s.flag = d->d_flag;
  s.tab = d->tab; s.geo = d->geo; s.phys = d->phys; s.stream = d->stream;
  return s;
}

int launch_stage(dgx_domain *d, const dgx::StageLaunch &s) {
  hipError_t e = timed(d, s.op == 0 ? dgx::OP_VISC : dgx::OP_RHS,
                       [&] { return d->mixed ? dgx::launch_mixed3(s) : dgx::launch_stage3(s); });
  if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("stage launch: ") + hipGetErrorString(e));
  return DGX_OK;
}

// BR1 viscous terms of state U from traces T into V, TV (fused hex path).
int vterm(dgx_domain *d, const double *U, const double *T) {
  dgx::StageLaunch s = stage_args(d);
  s.op = 0; s.U = const_cast<double *>(U); s.T = T; s.Vw = d->V; s.TVw = d->TV;
  s.Vfw = d->Vf; s.TVfw = d->TVf;
  return launch_stage(d, s);
}

int residual_core(dgx_domain *d, const double *U, double *R);

// Residual of state `U` into `R` (Domain::time_derivative), then the channel's
// bulk forcing where the domain has it (src/Navier_Stokes3d.cpp:160).
int residual(dgx_domain *d, const double *U, double *R) {
  int rc = residual_core(d, U, R);
  if (rc || !d->bulk) return rc;
  hipError_t e = timed(d, 6, [&] { return dgx::bulk_hold(R, d->d_J, d->ne, d->nm, d->nf, d->stream); });
  if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("bulk forcing: ") + hipGetErrorString(e));
  return DGX_OK;
}

int residual_core(dgx_domain *d, const double *U, double *R) {
  int rc;
  if (d->fused) {
    if ((rc = launch(d, dgx::OP_TRACES, U, nullptr))) return rc;
    if (d->visc && (rc = vterm(d, U, d->T))) return rc;
    dgx::StageLaunch s = stage_args(d);
    s.op = 1; s.U = const_cast<double *>(U); s.T = d->T; s.V = d->V; s.TV = d->TV; s.R = R;
    s.Vf = d->Vf; s.TVf = d->TVf;
    if ((rc = launch_stage(d, s))) return rc;
    ++d->stages;
    return DGX_OK;
  }
  if ((rc = launch(d, dgx::OP_TRACES, U, nullptr))) return rc;
  if (d->curved) {
    dgx::CurvedArgs ca = curved_args(d);
    ca.U = U; ca.T = d->T; ca.R = R;
    for (int op : {1, 2}) {   // Navier-Stokes: lifted gradients, then the viscous terms
      if (!d->visc) break;
      ca.op = op;
      hipError_t e = timed(d, op == 1 ? dgx::OP_GRAD : dgx::OP_VISC, [&] { return launch_curved(d, ca); });
      if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("curved viscous terms: ") + hipGetErrorString(e));
    }
    ca.op = 0;
    ca.Bal = d->Bal; ca.bal = d->bal;
    hipError_t e = timed(d, 1, [&] { return launch_curved(d, ca); });
    if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("curved residual: ") + hipGetErrorString(e));
    if (d->bal == 1) d->bal = 2;   // the first residual is stored
    ++d->stages;
    return DGX_OK;
  }
  if (d->visc) {
    if ((rc = launch(d, dgx::OP_GRAD, U, nullptr))) return rc;
    if ((rc = launch(d, dgx::OP_VISC, U, nullptr))) return rc;
  }
  if ((rc = launch(d, dgx::OP_RHS, U, R))) return rc;
  ++d->stages;
  return DGX_OK;
}

// One LSRK stage of an affine 2D domain (not curved, one order): the traces
// (unless the previous stage of this call wrote them) and the viscous terms,
// then the residual fused with the register update Ut = A Ut - dt R,
// U += B Ut and with the traces of the new state (k_step2), bitwise the same
// as the residual followed by the axpby and axpy.
int step2(dgx_domain *d, double A, double B, double dt, bool traces_ready) {
  int rc;
  if (!traces_ready && (rc = launch(d, dgx::OP_TRACES, d->U, nullptr))) return rc;
  if (d->visc) {
    if ((rc = launch(d, dgx::OP_GRAD, d->U, nullptr))) return rc;
    if ((rc = launch(d, dgx::OP_VISC, d->U, nullptr))) return rc;
  }
  dgx::LaunchArgs a = launch_args(d, dgx::OP_STEP, d->U, nullptr);
  a.Us = d->U; a.Ut = d->Ut; a.ca = A; a.cb = B; a.dt = dt; a.Tn = d->T2;
  hipError_t e = timed(d, dgx::OP_RHS, [&] { return dgx::launch_2d(a); });
  if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("stage launch: ") + hipGetErrorString(e));
  std::swap(d->T, d->T2);        // the next stage's traces
  if (d->bal == 1) d->bal = 2;   // the first residual is stored
  ++d->stages;
  return DGX_OK;
}

// One LSRK stage of a curved hex domain: the traces, on Navier-Stokes the
// lifted gradients and viscous terms, then k_rhs3c with the register update
// fused into its store (bitwise the same as the residual followed by the
// axpby and axpy).
int step_curved3(dgx_domain *d, double A, double B, double dt) {
  int rc;
  if ((rc = launch(d, dgx::OP_TRACES, d->U, nullptr))) return rc;
  dgx::CurvedArgs ca = curved_args(d);
  ca.U = d->U; ca.T = d->T; ca.R = nullptr;
  for (int op : {1, 2}) {   // Navier-Stokes: lifted gradients, then the viscous terms
    if (!d->visc) break;
    ca.op = op;
    hipError_t e = timed(d, op == 1 ? dgx::OP_GRAD : dgx::OP_VISC, [&] { return launch_curved(d, ca); });
    if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("curved viscous terms: ") + hipGetErrorString(e));
  }
  ca.op = 0;
  ca.update = true; ca.Us = d->U; ca.Ut = d->Ut; ca.ca = A; ca.cb = B; ca.dt = dt;
  hipError_t e = timed(d, dgx::OP_RHS, [&] { return launch_curved(d, ca); });
  if (e != hipSuccess) return fail(DGX_ERR_DEVICE, std::string("curved stage: ") + hipGetErrorString(e));
  ++d->stages;
  return DGX_OK;
}

int ensure_nodal(dgx_domain *d) {
  if (d->nodal) return DGX_OK;
  return alloc(&d->nodal, (size_t)d->nf * d->ne * d->nq);
}

// Curved hexes from nodal coordinates xn[3][ne][Q^3]: metrics and J at the
// nodes as Hex::compute_curved_geometry (src/Hex.cpp:3992-4080), the face
// normals and J_side/J_vol at the Gauss points from them as
// Hex::compute_side_geometry (src/Hex.cpp:1385-1494), and per element the mean
// J that starts the kernel's mass-matrix solve.
// Curved quads from nodal coordinates xn[2][ne][Q^2]: metrics and J at the
// nodes as Quad::compute_curved_geometry (src/Quad.cpp:1309-1363), the edge
// normals and J_side/J_vol at the Gauss points from them as
// Quad::compute_side_geometry (src/Quad.cpp:1367-1452), and per element the
// mean J that starts the kernel's mass-matrix solve.
// Quad::is_affine (src/Quad.cpp:1184-1205): each half edge vector equals its
// opposite within REL = sqrt(eps) or ABS = eps^(2/3) (CMC::approx_equal,
// src/Vector.hpp:1253-1264).  Other quads are curved in the reference
// (Field::read_curve_pc, src/Field.cpp:1226-1240).
bool quad_affine(const double *vx, const double *vy) {
  const double eps = std::numeric_limits<double>::epsilon();
  const double REL = std::sqrt(eps), ABS = std::pow(eps, 0.6666666666666666667);
  auto close = [&](double u, double v) {
    if (u == v) return true;
    const double diff = std::fabs(u - v);
    if (diff <= ABS) return true;
    return diff <= REL * std::max(std::fabs(u), std::fabs(v));
  };
  return close((vx[1] - vx[0]) * 0.5, (vx[2] - vx[3]) * 0.5) && close((vx[3] - vx[0]) * 0.5, (vx[2] - vx[1]) * 0.5) &&
         close((vy[1] - vy[0]) * 0.5, (vy[2] - vy[3]) * 0.5) && close((vy[3] - vy[0]) * 0.5, (vy[2] - vy[1]) * 0.5);
}

bool all_quads_affine(const dgx_domain_desc *desc) {
  if (desc->nsd != 2 || !desc->vert) return true;
  for (int e = 0; e < desc->ne; ++e) {
    const double *vx = desc->vert + (size_t)e * 8;
    if (!quad_affine(vx, vx + 4)) return false;
  }
  return true;
}

// A dgx_function with ncomp components whose programs all run within the
// stack and leave one value (the device evaluator relies on it).
bool check_function(const dgx_function *f, int ncomp, std::string *why) {
  if (f->ncomp != ncomp) { *why = "must have " + std::to_string(ncomp) + " components"; return false; }
  if (!f->start || !f->code || (f->nconst > 0 && !f->consts) || f->nconst < 0) { *why = "null arrays"; return false; }
  if (f->start[0] != 0) { *why = "start[0] must be 0"; return false; }
  for (int c = 0; c < ncomp; ++c) {
    if (f->start[c + 1] <= f->start[c] || f->start[c + 1] > (1 << 24)) {
      *why = "component " + std::to_string(c) + " is empty or too long"; return false;
    }
    int sp = 0;
    for (int i = f->start[c]; i < f->start[c + 1]; ++i) {
      const int ins = f->code[i], op = ins & 255;
      if (ins < 0 || op > DGX_OP_ATAN2 || (op != DGX_OP_CONST && (ins >> 8) != 0)) {
        *why = "bad instruction " + std::to_string(ins); return false;
      }
      if (op == DGX_OP_CONST && (ins >> 8) >= f->nconst) { *why = "constant index out of range"; return false; }
      const int need = (op <= DGX_OP_T) ? 0 : ((op <= DGX_OP_POW || op == DGX_OP_ATAN2) ? 2 : 1);
      if (sp < need) { *why = "stack underflow in component " + std::to_string(c); return false; }
      sp += (op <= DGX_OP_T) ? 1 : ((need == 2) ? -1 : 0);
      if (sp > DGX_FN_STACK) { *why = "stack deeper than DGX_FN_STACK"; return false; }
    }
    if (sp != 1) { *why = "component " + std::to_string(c) + " does not leave one value"; return false; }
  }
  return true;
}

// Device copies of the boundary and source programs (one code and constant
// table, the source's indices shifted behind the boundary's) and the
// coordinates they are evaluated at: every side's Gauss points (the GLL edge
// coordinates interpolated, as Element::get_side_coordinates does) and the nodes.
int setup_functions(dgx_domain *d, const dgx_domain_desc *desc) {
  const dgx_function *fs[2] = {desc->bc_function, desc->source};
  std::vector<int> code;
  std::vector<double> consts;
  for (int k = 0; k < 2; ++k) {
    if (!fs[k]) continue;
    const dgx_function *f = fs[k];
    int *bounds = k == 0 ? d->fn.bc : d->fn.src;
    const int base = (int)code.size(), cbase = (int)consts.size();
    for (int c = 0; c <= 3; ++c) bounds[c] = base + f->start[c];
    for (int i = 0; i < f->start[3]; ++i) {
      const int ins = f->code[i];
      code.push_back((ins & 255) == DGX_OP_CONST ? (((ins >> 8) + cbase) << 8) : ins);
    }
    consts.insert(consts.end(), f->consts, f->consts + f->nconst);
    (k == 0 ? d->fn.has_bc : d->fn.has_src) = 1;
  }
  if (code.empty()) return DGX_OK;
  if (consts.empty()) consts.push_back(0.0);
  const int ne = d->ne, Q = d->Q, QF = d->QF, Q2 = d->nq;
  std::vector<double> fxy((size_t)ne * 4 * QF * 2);
  const std::vector<double> &Ig = d->hb.Ig;
  for (int e = 0; e < ne; ++e)
    for (int s = 0; s < 4; ++s)
      for (int a = 0; a < QF; ++a) {
        double x = 0.0, y = 0.0;
        for (int t = 0; t < Q; ++t) {
          const int n = s == 0 ? t * Q : (s == 1 ? (Q - 1) * Q + t : (s == 2 ? t * Q + Q - 1 : t));
          x += Ig[a * Q + t] * d->coords[(size_t)e * Q2 + n];
          y += Ig[a * Q + t] * d->coords[((size_t)ne + e) * Q2 + n];
        }
        double *o = fxy.data() + (((size_t)e * 4 + s) * QF + a) * 2;
        o[0] = x; o[1] = y;
      }
  int rc;
  HIPCHK(hipMalloc((void **)&d->d_fcode, code.size() * sizeof(int)));
  HIPCHK(hipMemcpy(d->d_fcode, code.data(), code.size() * sizeof(int), hipMemcpyHostToDevice));
  if ((rc = alloc(&d->d_fconst, consts.size())) || (rc = alloc(&d->d_fxy, fxy.size())) ||
      (rc = alloc(&d->d_xn, d->coords.size())))
    return rc;
  HIPCHK(hipMemcpy(d->d_fconst, consts.data(), consts.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_fxy, fxy.data(), fxy.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_xn, d->coords.data(), d->coords.size() * 8, hipMemcpyHostToDevice));
  d->fn.code = d->d_fcode; d->fn.consts = d->d_fconst; d->fn.fxy = d->d_fxy; d->fn.xn = d->d_xn;
  return DGX_OK;
}

int curved_geometry2(dgx_domain *d, const double *xn) {
  const int ne = d->ne, Q = d->Q, QF = d->QF, Q2 = d->nq;
  const HostBasis &b = d->hb;
  const std::vector<double> &D = b.D, &w = b.w, &Ig = b.Ig;
  std::vector<double> gm((size_t)ne * 5 * Q2), fgeo((size_t)ne * 4 * QF * 4), jav(ne);
  std::vector<double> xr(Q2), xs(Q2), yr(Q2), ys(Q2), J(Q2), mr[2] = {std::vector<double>(Q2), std::vector<double>(Q2)},
      ms[2] = {std::vector<double>(Q2), std::vector<double>(Q2)};
  std::vector<double> ex(Q), ey(Q), en0(Q), en1(Q), ej(Q), dl(Q);
  auto gauss = [&](const std::vector<double> &f, double *to) {
    for (int a = 0; a < QF; ++a) {
      double t = 0.0;
      for (int m = 0; m < Q; ++m) t += Ig[a * Q + m] * f[m];
      to[a] = t;
    }
  };
  for (int e = 0; e < ne; ++e) {
    const double *X = xn + (size_t)e * Q2, *Y = xn + ((size_t)ne + e) * Q2;
    for (int i = 0; i < Q; ++i)
      for (int j = 0; j < Q; ++j) {
        double a0 = 0, a1 = 0, c0 = 0, c1 = 0;
        for (int m = 0; m < Q; ++m) {
          a0 += D[i * Q + m] * X[m * Q + j]; a1 += D[j * Q + m] * X[i * Q + m];
          c0 += D[i * Q + m] * Y[m * Q + j]; c1 += D[j * Q + m] * Y[i * Q + m];
        }
        const int n = i * Q + j;
        xr[n] = a0; xs[n] = a1; yr[n] = c0; ys[n] = c1;
      }
    double sw = 0.0, swJ = 0.0;
    for (int n = 0; n < Q2; ++n) {
      const double Jn = xr[n] * ys[n] - xs[n] * yr[n];
      if (!(Jn > 0.0)) return fail(DGX_ERR_ARGUMENT, "element " + std::to_string(e) + " has a nonpositive Jacobian");
      J[n] = Jn;
      mr[0][n] = ys[n] / Jn; mr[1][n] = -(xs[n] / Jn);
      ms[0][n] = -(yr[n] / Jn); ms[1][n] = xr[n] / Jn;
      const int i = n / Q, j = n % Q;
      const double wJ = J[n] * w[i] * w[j];
      sw += w[i] * w[j]; swJ += wJ;
      double *g = gm.data() + (size_t)e * 5 * Q2;
      g[0 * Q2 + n] = wJ * mr[0][n]; g[1 * Q2 + n] = wJ * mr[1][n];
      g[2 * Q2 + n] = wJ * ms[0][n]; g[3 * Q2 + n] = wJ * ms[1][n];
      g[4 * Q2 + n] = wJ;
    }
    jav[e] = swJ / sw;
    for (int s = 0; s < 4; ++s) {
      const std::vector<double> *M = (s == 0 || s == 2) ? ms : mr;
      for (int t = 0; t < Q; ++t) {
        const int n = s == 0 ? t * Q : (s == 1 ? (Q - 1) * Q + t : (s == 2 ? t * Q + Q - 1 : t));
        en0[t] = M[0][n]; en1[t] = M[1][n]; ej[t] = J[n]; ex[t] = X[n]; ey[t] = Y[n];
      }
      double gx[16], gy[16], gj[16], gl[16];
      gauss(en0, gx); gauss(en1, gy); gauss(ej, gj);
      for (int t = 0; t < Q; ++t) {
        double dx = 0, dy = 0;
        for (int m = 0; m < Q; ++m) { dx += D[t * Q + m] * ex[m]; dy += D[t * Q + m] * ey[m]; }
        dl[t] = std::sqrt(dx * dx + dy * dy);
      }
      gauss(dl, gl);
      const double sg = (s == 0 || s == 3) ? -1.0 : 1.0;
      for (int a = 0; a < QF; ++a) {
        const double nx = sg * gx[a], ny = sg * gy[a], len = std::sqrt(nx * nx + ny * ny);
        double *o = fgeo.data() + (((size_t)e * 4 + s) * QF + a) * 4;
        o[0] = nx / len; o[1] = ny / len; o[2] = 0.0; o[3] = gl[a] / gj[a];
      }
    }
  }
  int rc;
  if ((rc = alloc(&d->d_gm, gm.size())) || (rc = alloc(&d->d_fgeo, fgeo.size())) || (rc = alloc(&d->d_jav, jav.size())))
    return rc;
  HIPCHK(hipMemcpy(d->d_gm, gm.data(), gm.size() * 8, hipMemcpyHostToDevice));
  d->gm_host = gm;   // the norms integrate with the nodal metrics
  HIPCHK(hipMemcpy(d->d_fgeo, fgeo.data(), fgeo.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_jav, jav.data(), jav.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMalloc((void **)&d->d_iters, 64));
  HIPCHK(hipMemset(d->d_iters, 0, 64));
  // Cholesky factors of the element mass matrices M[a][b] = sum_n phi_a wJ phi_b
  // over the GLL nodes (Quad::get_mass_matrix and its factorization with
  // DGM_CHOLESKY_MASS_MATRIX, src/Quad.cpp:263-300), rows of the lower factor
  // packed; the projection then solves with them instead of iterating.  Held
  // while they take at most 1 GiB (p = 10: 59 KB per element); beyond that the
  // kernels keep the in-workgroup conjugate gradients.
  const int L = d->L, nm = L * L;
  const size_t pk = (size_t)nm * (nm + 1) / 2;
  if ((size_t)ne * pk * 8 <= (size_t(1) << 30)) {
    std::vector<double> cf((size_t)ne * pk);
    const std::vector<double> &BT = b.BT;   // [Q][L]
    auto factor = [&](int e0, int e1) {
      std::vector<double> Phi((size_t)nm * Q2), M((size_t)nm * nm), C((size_t)nm * nm);
      for (int a = 0; a < L; ++a)
        for (int c = 0; c < L; ++c)
          for (int i = 0; i < Q; ++i)
            for (int j = 0; j < Q; ++j) Phi[(size_t)(a * L + c) * Q2 + i * Q + j] = BT[i * L + a] * BT[j * L + c];
      for (int e = e0; e < e1; ++e) {
        const double *wJ = gm.data() + (size_t)e * 5 * Q2 + 4 * Q2;
        for (int r = 0; r < nm; ++r)
          for (int c = 0; c <= r; ++c) {
            double t = 0.0;
            for (int n = 0; n < Q2; ++n) t += Phi[(size_t)r * Q2 + n] * wJ[n] * Phi[(size_t)c * Q2 + n];
            M[(size_t)r * nm + c] = t;
          }
        for (int r = 0; r < nm; ++r)
          for (int c = 0; c <= r; ++c) {
            double t = M[(size_t)r * nm + c];
            for (int k = 0; k < c; ++k) t -= C[(size_t)r * nm + k] * C[(size_t)c * nm + k];
            C[(size_t)r * nm + c] = (r == c) ? std::sqrt(t) : t / C[(size_t)c * nm + c];
          }
        double *o = cf.data() + (size_t)e * pk;
        for (int r = 0; r < nm; ++r)
          for (int c = 0; c <= r; ++c) o[(size_t)r * (r + 1) / 2 + c] = C[(size_t)r * nm + c];
      }
    };
    const int nth = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    const int chunk = (ne + nth - 1) / nth;
    for (int t = 0; t < nth; ++t) {
      const int e0 = t * chunk, e1 = std::min(ne, e0 + chunk);
      if (e0 < e1) pool.emplace_back(factor, e0, e1);
    }
    for (auto &th : pool) th.join();
    for (size_t k = 0; k < cf.size(); k += pk)
      if (!(cf[k] > 0.0) || !std::isfinite(cf[k + pk - 1]) || !(cf[k + pk - 1] > 0.0))
        return fail(DGX_ERR_ARGUMENT, "element " + std::to_string(k / pk) + ": mass matrix not positive definite");
    int rc2;
    if ((rc2 = alloc(&d->d_chol, cf.size()))) return rc2;
    HIPCHK(hipMemcpy(d->d_chol, cf.data(), cf.size() * 8, hipMemcpyHostToDevice));
  }
  std::copy(xn, xn + (size_t)2 * ne * Q2, d->coords.begin());
  d->curved = true;
  return DGX_OK;
}

// Inverse of a small symmetric positive definite n x n matrix (row-major) by
// Cholesky factorization; false if it is not positive definite.
bool spd_inverse(const double *A, int n, double *out) {
  std::vector<double> C((size_t)n * n, 0.0), y(n);
  for (int r = 0; r < n; ++r)
    for (int c = 0; c <= r; ++c) {
      double t = A[r * n + c];
      for (int k = 0; k < c; ++k) t -= C[r * n + k] * C[c * n + k];
      if (r == c) {
        if (!(t > 0.0)) return false;
        C[r * n + r] = std::sqrt(t);
      } else {
        C[r * n + c] = t / C[c * n + c];
      }
    }
  for (int col = 0; col < n; ++col) {
    for (int r = 0; r < n; ++r) {   // C y = e_col
      double t = r == col ? 1.0 : 0.0;
      for (int k = 0; k < r; ++k) t -= C[r * n + k] * y[k];
      y[r] = t / C[r * n + r];
    }
    for (int r = n - 1; r >= 0; --r) {   // C^T x = y
      double t = y[r];
      for (int k = r + 1; k < n; ++k) t -= C[k * n + r] * out[k * n + col];
      out[r * n + col] = t / C[r * n + r];
    }
  }
  return true;
}

int curved_geometry(dgx_domain *d, const double *xn) {
  const int ne = d->ne, Q = d->Q, QF = d->QF, Q3 = d->nq, QF2 = QF * QF, Q2 = Q * Q;
  const HostBasis &b = d->hb;
  const std::vector<double> &D = b.D, &w = b.w, &Ig = b.Ig;
  std::vector<double> gm((size_t)ne * 10 * Q3), fgeo((size_t)ne * 6 * QF2 * 4), jav(ne);
  const int L = d->L, LL = L * L;
  std::vector<double> pre((size_t)ne * 3 * LL);
  std::atomic<int> bad_elem(-1);
  auto work = [&](int e0, int e1) {
  std::vector<double> der(9 * (size_t)Q3), mn(9 * (size_t)Q3), J(Q3), wJ(Q3);
  auto face = [&](const double *f, int s, double *to) {   // GLL face values [Q][Q] (get_side)
    for (int u = 0; u < Q; ++u)
      for (int v = 0; v < Q; ++v) {
        int i, j, k;
        switch (s) {
          case 0: i = u; j = v; k = 0; break;
          case 1: i = u; j = 0; k = v; break;
          case 2: i = Q - 1; j = u; k = v; break;
          case 3: i = u; j = Q - 1; k = v; break;
          case 4: i = 0; j = u; k = v; break;
          default: i = u; j = v; k = Q - 1; break;
        }
        to[u * Q + v] = f[(i * Q + j) * Q + k];
      }
  };
  auto gauss = [&](const double *f, double *to) {   // GLL face -> Gauss points (Ig (x) Ig)
    std::vector<double> t(QF * Q);
    for (int a2 = 0; a2 < QF; ++a2)
      for (int v = 0; v < Q; ++v) {
        double s = 0.0;
        for (int u = 0; u < Q; ++u) s += Ig[a2 * Q + u] * f[u * Q + v];
        t[a2 * Q + v] = s;
      }
    for (int a2 = 0; a2 < QF; ++a2)
      for (int b2 = 0; b2 < QF; ++b2) {
        double s = 0.0;
        for (int v = 0; v < Q; ++v) s += t[a2 * Q + v] * Ig[b2 * Q + v];
        to[a2 * QF + b2] = s;
      }
  };
  std::vector<double> fv(Q2), fg(QF2), X(Q2), Y(Q2), Z(Q2), nrm(3 * QF2), jf(QF2), cr(Q2);
  for (int e = e0; e < e1; ++e) {
    const double *C[3] = {xn + ((size_t)0 * ne + e) * Q3, xn + ((size_t)1 * ne + e) * Q3,
                          xn + ((size_t)2 * ne + e) * Q3};
    for (int c = 0; c < 3; ++c)
      for (int i = 0; i < Q; ++i)
        for (int j = 0; j < Q; ++j)
          for (int k = 0; k < Q; ++k) {
            double r = 0.0, s = 0.0, t = 0.0;
            for (int m = 0; m < Q; ++m) {
              r += D[i * Q + m] * C[c][(m * Q + j) * Q + k];
              s += D[j * Q + m] * C[c][(i * Q + m) * Q + k];
              t += D[k * Q + m] * C[c][(i * Q + j) * Q + m];
            }
            const int n = (i * Q + j) * Q + k;
            der[(3 * c + 0) * Q3 + n] = r; der[(3 * c + 1) * Q3 + n] = s; der[(3 * c + 2) * Q3 + n] = t;
          }
    double sw = 0.0, swJ = 0.0;
    for (int n = 0; n < Q3; ++n) {
      const double xr = der[n], xs = der[Q3 + n], xt = der[2 * Q3 + n];
      const double yr = der[3 * Q3 + n], ys = der[4 * Q3 + n], yt = der[5 * Q3 + n];
      const double zr = der[6 * Q3 + n], zs = der[7 * Q3 + n], zt = der[8 * Q3 + n];
      const double Jn = xr * ys * zt + yr * zs * xt + zr * xs * yt - xt * ys * zr - zt * xs * yr - xr * zs * yt;
      if (!(Jn > 0.0)) { int none = -1; bad_elem.compare_exchange_strong(none, e); return; }
      J[n] = Jn;
      mn[0 * Q3 + n] = (ys * zt - yt * zs) / Jn;
      mn[1 * Q3 + n] = -(xs * zt - xt * zs) / Jn;
      mn[2 * Q3 + n] = (xs * yt - xt * ys) / Jn;
      mn[3 * Q3 + n] = -(yr * zt - yt * zr) / Jn;
      mn[4 * Q3 + n] = (xr * zt - xt * zr) / Jn;
      mn[5 * Q3 + n] = -(xr * yt - xt * yr) / Jn;
      mn[6 * Q3 + n] = (yr * zs - ys * zr) / Jn;
      mn[7 * Q3 + n] = -(xr * zs - xs * zr) / Jn;
      mn[8 * Q3 + n] = (xr * ys - xs * yr) / Jn;
      const int i = n / Q2, j = (n / Q) % Q, k = n % Q;
      const double ww = w[i] * w[j] * w[k];
      wJ[n] = ww * Jn;
      sw += ww; swJ += wJ[n];
      for (int c = 0; c < 9; ++c) gm[((size_t)e * 10 + c) * Q3 + n] = wJ[n] * mn[c * Q3 + n];
      gm[((size_t)e * 10 + 9) * Q3 + n] = wJ[n];
    }
    jav[e] = swJ / sw;
    // Preconditioner of the mass solve: J ~ ja(i) jb(j) jc(k), the log-mean
    // separable fit, whose mass matrix is the tensor product of the 1D
    // matrices BT^T diag(w jd) BT; their inverses, [3][L][L] per element
    {
      std::vector<double> lg(3 * (size_t)Q, 0.0);
      double lm = 0.0;
      for (int n = 0; n < Q3; ++n) {
        const double l = std::log(J[n]);
        const int i = n / Q2, j = (n / Q) % Q, k = n % Q;
        lg[i] += l; lg[Q + j] += l; lg[2 * Q + k] += l;
        lm += l;
      }
      lm /= Q3;
      std::vector<double> A(LL), Ai(LL);
      for (int dd = 0; dd < 3; ++dd) {
        for (int a = 0; a < L; ++a)
          for (int c = 0; c < L; ++c) {
            double t = 0.0;
            for (int i = 0; i < Q; ++i)
              t += b.BT[i * L + a] * w[i] * std::exp(lg[dd * Q + i] / Q2 - 2.0 * lm / 3.0) * b.BT[i * L + c];
            A[a * L + c] = t;
          }
        if (!spd_inverse(A.data(), L, Ai.data())) { int none = -1; bad_elem.compare_exchange_strong(none, e); return; }
        std::copy(Ai.begin(), Ai.end(), pre.begin() + ((size_t)e * 3 + dd) * LL);
      }
    }
    for (int s = 0; s < 6; ++s) {
      const int row = (s == 0 || s == 5) ? 6 : ((s == 1 || s == 3) ? 3 : 0);
      const double sg = (s == 0 || s == 1 || s == 4) ? -1.0 : 1.0;
      for (int c = 0; c < 3; ++c) {
        face(mn.data() + (size_t)(row + c) * Q3, s, fv.data());
        gauss(fv.data(), nrm.data() + c * QF2);
      }
      face(J.data(), s, fv.data());
      gauss(fv.data(), jf.data());
      face(C[0], s, X.data()); face(C[1], s, Y.data()); face(C[2], s, Z.data());
      for (int u = 0; u < Q; ++u)
        for (int v = 0; v < Q; ++v) {
          double XR = 0, XS = 0, YR = 0, YS = 0, ZR = 0, ZS = 0;
          for (int m = 0; m < Q; ++m) {
            XR += D[u * Q + m] * X[m * Q + v]; XS += X[u * Q + m] * D[v * Q + m];
            YR += D[u * Q + m] * Y[m * Q + v]; YS += Y[u * Q + m] * D[v * Q + m];
            ZR += D[u * Q + m] * Z[m * Q + v]; ZS += Z[u * Q + m] * D[v * Q + m];
          }
          const double cx = YR * ZS - ZR * YS, cy = ZR * XS - XR * ZS, cz = XR * YS - XS * YR;
          cr[u * Q + v] = std::sqrt(cx * cx + cy * cy + cz * cz);
        }
      gauss(cr.data(), fg.data());
      for (int g = 0; g < QF2; ++g) {
        double nx = sg * nrm[g], ny = sg * nrm[QF2 + g], nz = sg * nrm[2 * QF2 + g];
        const double len = std::sqrt(nx * nx + ny * ny + nz * nz);
        double *o = fgeo.data() + (((size_t)e * 6 + s) * QF2 + g) * 4;
        o[0] = nx / len; o[1] = ny / len; o[2] = nz / len; o[3] = fg[g] / jf[g];
      }
    }
  }
  };
  {
    const int nth = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
    const int chunk = (ne + nth - 1) / nth;
    std::vector<std::thread> pool;
    for (int t = 0; t < nth && t * chunk < ne; ++t) pool.emplace_back(work, t * chunk, std::min(ne, (t + 1) * chunk));
    for (auto &th : pool) th.join();
  }
  if (bad_elem.load() >= 0)
    return fail(DGX_ERR_ARGUMENT, "element " + std::to_string(bad_elem.load()) + " has a nonpositive Jacobian");

  int rc;
  if ((rc = alloc(&d->d_gm, gm.size())) || (rc = alloc(&d->d_fgeo, fgeo.size())) || (rc = alloc(&d->d_jav, jav.size())) ||
      (rc = alloc(&d->d_pre, pre.size())))
    return rc;
  HIPCHK(hipMemcpy(d->d_gm, gm.data(), gm.size() * 8, hipMemcpyHostToDevice));
  d->gm_host = gm;   // the norms integrate with the nodal metrics
  HIPCHK(hipMemcpy(d->d_fgeo, fgeo.data(), fgeo.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_jav, jav.data(), jav.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_pre, pre.data(), pre.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMalloc((void **)&d->d_iters, 64));
  HIPCHK(hipMemset(d->d_iters, 0, 64));
  std::copy(xn, xn + (size_t)3 * ne * Q3, d->coords.begin());   // the nodes are the element's coordinates
  d->curved = true;
  return DGX_OK;
}

// Variable order (root.ord): elements grouped by order, every side's Gauss
// points max(L_e, L_neighbour) (Side::match, src/Side.cpp:106-117) and flip,
// and the side tables of each (order, point count): Bgv = the normalized
// Legendre modes at the side's Gauss points, PWv = (B diag(w)) G2GL, the
// projected lift of Quad::local_lift (src/Quad.cpp:1971-2038).  Quads, Euler
// (or Navier_Stokes without viscosity, whose viscous terms are then zero),
// affine, FP64.
// The orders were checked by dgx_domain_create.
int varp_setup(dgx_domain *d, const dgx_domain_desc *desc, const std::vector<Side> &sides) {
  const int ne = d->ne, p = d->p;
  d->porder.assign(desc->porder, desc->porder + ne);
  d->varp = true;
  d->visc = false;   // Navier_Stokes with vis = 0: the diffusion term adds exact zeros
  std::vector<int> elems;
  for (int P = 1; P <= p; ++P) {
    const int off = (int)elems.size();
    for (int e = 0; e < ne; ++e)
      if (d->porder[e] == P) elems.push_back(e);
    if ((int)elems.size() > off) {
      d->vgrp_p.push_back(P); d->vgrp_off.push_back(off); d->vgrp_n.push_back((int)elems.size() - off);
      // GLL node spacings of the order, for the CFL metrics (src/Element.cpp:513-527)
      const HostBasis hb = make_basis(P);
      const int Q = P + 2;
      for (int i = 0; i < dgx::kMaxOrder + 2; ++i)
        d->vgrp_dz.push_back(i >= Q ? 0.0 : i == 0 ? hb.z[1] - hb.z[0]
                                          : (i == Q - 1 ? hb.z[i] - hb.z[i - 1] : 0.5 * (hb.z[i + 1] - hb.z[i - 1])));
    }
  }
  std::vector<int> qs((size_t)ne * 4);
  std::vector<uint8_t> flip((size_t)ne * 4, 0);
  for (int e = 0; e < ne; ++e)
    for (int s = 0; s < 4; ++s) {
      const Side &S = sides[(size_t)e * 4 + s];
      int n = d->porder[e] + 1;
      if (S.bc < 0) {
        n = std::max(n, d->porder[S.le] + 1);
        // 2D flips reverse the side; the master holds the flag (src/Field.cpp:416-668)
        flip[(size_t)e * 4 + s] = S.master ? S.fa : sides[(size_t)S.le * 4 + S.ls].fa;
      }
      qs[(size_t)e * 4 + s] = n;
    }
  std::vector<double> vt(dgx::kVpTab, 0.0);
  for (int P = 1; P <= dgx::kMaxOrder; ++P) {
    const int L = P + 1, Q = P + 2;
    std::vector<ld> z, w;
    gll_nodes(Q, z, w);
    for (int n = L; n <= dgx::kMaxOrder + 1; ++n) {
      std::vector<ld> zg, wg;
      gauss_nodes(n, zg, wg);
      double *Bg = &vt[dgx::vp_block(P, n, 0)], *PW = &vt[dgx::vp_block(P, n, 1)];
      for (int g = 0; g < n; ++g)
        for (int l = 0; l < L; ++l) Bg[g * L + l] = (double)normleg(l, zg[g]);
      for (int l = 0; l < L; ++l)
        for (int g = 0; g < n; ++g) {
          ld acc = 0;
          for (int i = 0; i < Q; ++i) acc += normleg(l, z[i]) * w[i] * lagrange(zg, g, z[i]);
          PW[l * n + g] = (double)acc;
        }
    }
  }
  HIPCHK(hipMalloc((void **)&d->d_velems, elems.size() * sizeof(int)));
  HIPCHK(hipMalloc((void **)&d->d_vqs, qs.size() * sizeof(int)));
  HIPCHK(hipMalloc((void **)&d->d_vflip, flip.size()));
  int rc;
  if ((rc = alloc(&d->d_vtab, vt.size()))) return rc;
  HIPCHK(hipMemcpy(d->d_velems, elems.data(), elems.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_vqs, qs.data(), qs.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_vflip, flip.data(), flip.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_vtab, vt.data(), vt.size() * sizeof(double), hipMemcpyHostToDevice));
  // nodal coordinates at each element's own GLL points (padded slots)
  std::fill(d->coords.begin(), d->coords.end(), 0.0);
  for (int e = 0; e < ne; ++e) {
    const int Q = d->porder[e] + 2;
    std::vector<ld> z, w;
    gll_nodes(Q, z, w);
    const double *vx = desc->vert + (size_t)e * 2 * 4, *vy = vx + 4;
    static const int ai[4] = {0, 1, 1, 0}, bi[4] = {0, 0, 1, 1};
    double *cx = &d->coords[(size_t)e * d->nq], *cy = &d->coords[((size_t)ne + e) * d->nq];
    for (int n = 0; n < 4; ++n)
      for (int i = 0; i < Q; ++i)
        for (int j = 0; j < Q; ++j) {
          const double za = (double)z[i], zb = (double)z[j];
          const double phi = (ai[n] ? 0.5 * (1.0 + za) : 0.5 * (1.0 - za)) * (bi[n] ? 0.5 * (1.0 + zb) : 0.5 * (1.0 - zb));
          cx[i * Q + j] += vx[n] * phi;
          cy[i * Q + j] += vy[n] * phi;
        }
  }
  // traces in slots of kMaxOrder + 1 points per side and field
  (void)hipFree(d->T);
  d->T = nullptr;
  if ((rc = alloc(&d->T, (size_t)ne * 4 * d->nf * (dgx::kMaxOrder + 1)))) return rc;
  return DGX_OK;
}

// Launch `op` (OP_TRACES, OP_RHS, OP_FORWARD, OP_BACKWARD) of a variable-order
// domain, one launch per order group.
hipError_t launch_varp(const dgx_domain *d, int op, const double *in, double *out) {
  for (size_t g = 0; g < d->vgrp_p.size(); ++g) {
    dgx::VarpArgs a;
    a.op = op; a.p = d->vgrp_p[g];
    a.elems = d->d_velems + d->vgrp_off[g]; a.nel = d->vgrp_n[g]; a.ne = d->ne;
    a.NM = d->nm; a.QM = d->nq; a.QFM = dgx::kMaxOrder + 1;
    a.T = d->T; a.qs = d->d_vqs; a.flip = d->d_vflip; a.vtab = d->d_vtab; a.flag = d->d_flag;
    a.geo = d->geo; a.phys = d->phys; a.stream = d->stream;
    if (op == dgx::OP_TRACES || op == dgx::OP_RHS) { a.U = in; a.R = out; }
    else { a.X = in; a.Y = out; }
    const hipError_t e = dgx::launch_varp2(a);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

int build(dgx_domain *d, const dgx_domain_desc *desc) {
  const int nsd = desc->nsd, ne = desc->ne, p = desc->p;
  d->nsd = nsd; d->ne = ne; d->p = p;
  d->L = p + 1; d->Q = p + 2; d->QF = p + 1;
  d->nf = desc->eqntype == DGX_SHALLOW ? 3 : nsd + 2;
  d->nm = nsd == 2 ? d->L * d->L : d->L * d->L * d->L;
  d->nq = nsd == 2 ? d->Q * d->Q : d->Q * d->Q * d->Q;
  d->ns = nsd == 2 ? 4 : 6;
  d->qs = nsd == 2 ? d->QF : d->QF * d->QF;
  d->eqn = desc->eqntype; d->flux = desc->flux; d->prec = desc->precision;
  d->visc = (desc->eqntype == DGX_NAVIER_STOKES || desc->eqntype == DGX_NAVIER_STOKES3D);
  d->gamma = desc->gamma > 0 ? desc->gamma : 1.4;
  d->vis = desc->vis;
  d->Pr = desc->Pr > 0 ? desc->Pr : 0.72;
  d->hb = make_basis(p);
  (void)hipGetDevice(&d->device);
  const HostBasis &b = d->hb;
  const int Q = d->Q, nvert = nsd == 2 ? 4 : 8;

  // geometry (affine elements: constant Jacobian, src/Quad.cpp:1227-1292,
  // src/Hex.cpp:1124-1286) and nodal coordinates (vertex-mode interpolation)
  d->coords.assign((size_t)nsd * ne * d->nq, 0.0);
  d->metric.assign((size_t)ne * 9, 0.0);
  d->J.assign(ne, 0.0);
  std::vector<double> fnorm((size_t)ne * d->ns * 3, 0.0), fjr((size_t)ne * d->ns, 0.0);
  auto vmode = [](int which, double x) { return which ? 0.5 * (1.0 + x) : 0.5 * (1.0 - x); };
  bool nonaffine = false;
  for (int e = 0; e < ne; ++e) {
    const double *vx = desc->vert + (size_t)e * nsd * nvert, *vy = vx + nvert;
    const double *vz = nsd == 3 ? vy + nvert : nullptr;
    double *cx = &d->coords[((size_t)0 * ne + e) * d->nq];
    double *cy = &d->coords[((size_t)1 * ne + e) * d->nq];
    double *mt = &d->metric[(size_t)e * 9];
    if (nsd == 2) {
      static const int ai[4] = {0, 1, 1, 0}, bi[4] = {0, 0, 1, 1};
      for (int n = 0; n < 4; ++n)
        for (int i = 0; i < Q; ++i)
          for (int j = 0; j < Q; ++j) {
            const double phi = vmode(ai[n], b.z[i]) * vmode(bi[n], b.z[j]);
            cx[i * Q + j] += vx[n] * phi;
            cy[i * Q + j] += vy[n] * phi;
          }
      const double xr = (vx[1] - vx[0]) * 0.5, xs = (vx[3] - vx[0]) * 0.5;
      const double yr = (vy[1] - vy[0]) * 0.5, ys = (vy[3] - vy[0]) * 0.5;
      if (!quad_affine(vx, vy)) {   // the curved path replaces this element's geometry
        nonaffine = true;
        continue;
      }
      const double det = xr * ys - xs * yr;
      if (!(det > 0)) return fail(DGX_ERR_ARGUMENT, "element " + std::to_string(e) + " has J <= 0");
      d->J[e] = det;
      mt[0] = ys / det; mt[1] = -xs / det; mt[3] = -yr / det; mt[4] = xr / det;
      const double nrm[4][2] = {{-mt[3], -mt[4]}, {mt[0], mt[1]}, {mt[3], mt[4]}, {-mt[0], -mt[1]}};
      const double elen[4] = {std::sqrt(xr * xr + yr * yr), std::sqrt(xs * xs + ys * ys),
                              std::sqrt(xr * xr + yr * yr), std::sqrt(xs * xs + ys * ys)};
      for (int s = 0; s < 4; ++s) {
        const double len = std::sqrt(nrm[s][0] * nrm[s][0] + nrm[s][1] * nrm[s][1]);
        fnorm[((size_t)e * 4 + s) * 3 + 0] = nrm[s][0] / len;
        fnorm[((size_t)e * 4 + s) * 3 + 1] = nrm[s][1] / len;
        fjr[(size_t)e * 4 + s] = elen[s] / det;
      }
    } else {
      double *cz = &d->coords[((size_t)2 * ne + e) * d->nq];
      static const int ai[8] = {0, 1, 1, 0, 0, 1, 1, 0}, bi[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                       ci[8] = {0, 0, 0, 0, 1, 1, 1, 1};
      for (int n = 0; n < 8; ++n)
        for (int i = 0; i < Q; ++i)
          for (int j = 0; j < Q; ++j)
            for (int k = 0; k < Q; ++k) {
              const double phi = vmode(ai[n], b.z[i]) * vmode(bi[n], b.z[j]) * vmode(ci[n], b.z[k]);
              const size_t m = (size_t)(i * Q + j) * Q + k;
              cx[m] += vx[n] * phi; cy[m] += vy[n] * phi; cz[m] += vz[n] * phi;
            }
      const double xr = (vx[1] - vx[0]) * 0.5, xs = (vx[3] - vx[0]) * 0.5, xt = (vx[4] - vx[0]) * 0.5;
      const double yr = (vy[1] - vy[0]) * 0.5, ys = (vy[3] - vy[0]) * 0.5, yt = (vy[4] - vy[0]) * 0.5;
      const double zr = (vz[1] - vz[0]) * 0.5, zs = (vz[3] - vz[0]) * 0.5, zt = (vz[4] - vz[0]) * 0.5;
      // affine check: opposite edges equal (src/Hex.cpp:1139-1163)
      const double scale = std::fabs(xr) + std::fabs(ys) + std::fabs(zt) + std::fabs(xs) + std::fabs(xt) +
                           std::fabs(yr) + std::fabs(yt) + std::fabs(zr) + std::fabs(zs);
      const double tol = 1e-10 * scale;
      auto nq_ = [&](double a, double c) { return std::fabs(a - c) > tol; };
      if (nq_(xr, (vx[6] - vx[7]) * 0.5) || nq_(xs, (vx[6] - vx[5]) * 0.5) || nq_(xt, (vx[6] - vx[2]) * 0.5) ||
          nq_(yr, (vy[6] - vy[7]) * 0.5) || nq_(ys, (vy[6] - vy[5]) * 0.5) || nq_(yt, (vy[6] - vy[2]) * 0.5) ||
          nq_(zr, (vz[6] - vz[7]) * 0.5) || nq_(zs, (vz[6] - vz[5]) * 0.5) || nq_(zt, (vz[6] - vz[2]) * 0.5))
        return fail(DGX_ERR_ARGUMENT, "element " + std::to_string(e) + " is not affine");
      const double det = xr * ys * zt + yr * zs * xt + zr * xs * yt - xt * ys * zr - zt * xs * yr - xr * zs * yt;
      if (!(det > 0)) return fail(DGX_ERR_ARGUMENT, "element " + std::to_string(e) + " has J <= 0");
      d->J[e] = det;
      mt[0] = (ys * zt - yt * zs) / det;  mt[1] = -(xs * zt - xt * zs) / det; mt[2] = (xs * yt - xt * ys) / det;
      mt[3] = -(yr * zt - yt * zr) / det; mt[4] = (xr * zt - xt * zr) / det;  mt[5] = -(xr * yt - xt * yr) / det;
      mt[6] = (yr * zs - ys * zr) / det;  mt[7] = -(xr * zs - xs * zr) / det; mt[8] = (xr * ys - xs * yr) / det;
      auto cross = [](const double a[3], const double c[3]) {
        const double x = a[1] * c[2] - a[2] * c[1], y = a[2] * c[0] - a[0] * c[2], z = a[0] * c[1] - a[1] * c[0];
        return std::sqrt(x * x + y * y + z * z);
      };
      const double Xr[3] = {xr, yr, zr}, Xs[3] = {xs, ys, zs}, Xt[3] = {xt, yt, zt};
      // faces: 0 (t=-1), 1 (s=-1), 2 (r=+1), 3 (s=+1), 4 (r=-1), 5 (t=+1)
      const int dirm[6] = {2, 1, 0, 1, 0, 2};
      const double sgn[6] = {-1, -1, 1, 1, -1, 1};
      const double fj[6] = {cross(Xr, Xs), cross(Xr, Xt), cross(Xs, Xt), cross(Xr, Xt), cross(Xs, Xt), cross(Xr, Xs)};
      for (int s = 0; s < 6; ++s) {
        const double *g = mt + 3 * dirm[s];
        const double len = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
        for (int c = 0; c < 3; ++c) fnorm[((size_t)e * 6 + s) * 3 + c] = sgn[s] * g[c] / len;
        fjr[(size_t)e * 6 + s] = fj[s] / det;
      }
    }
  }

  // side links: the first .con row naming a pair makes its side the master
  // (src/Field.cpp:3478-3556); flips are set on masters (src/Field.cpp:416-668)
  std::vector<Side> sides((size_t)ne * d->ns);
  for (int r = 0; r < desc->ncon; ++r) {
    const int e = desc->con[4 * r], s = desc->con[4 * r + 1], le = desc->con[4 * r + 2], ls = desc->con[4 * r + 3];
    if (e >= 0 && e < ne && s >= 0 && s < d->ns && le == -1) {   // boundary side (e, s, -1, bc type)
      sides[(size_t)e * d->ns + s].bc = ls;   // types checked in dgx_domain_create
      continue;
    }
    if (e < 0 || e >= ne || le < 0 || le >= ne || s < 0 || s >= d->ns || ls < 0 || ls >= d->ns)
      return fail(DGX_ERR_ARGUMENT, "connectivity row " + std::to_string(r) + " out of range");
    Side &me = sides[(size_t)e * d->ns + s], &nb = sides[(size_t)le * d->ns + ls];
    if (nb.master) {
      if (nb.le != e || nb.ls != s) return fail(DGX_ERR_ARGUMENT, "contradictory side links");
      continue;
    }
    me.master = true; me.le = le; me.ls = ls;
    nb.master = false; nb.le = e; nb.ls = s;
  }
  for (size_t i = 0; i < sides.size(); ++i) {
    if (sides[i].le < 0 && sides[i].bc < 0)
      return fail(DGX_ERR_ARGUMENT, "element " + std::to_string(i / d->ns) + " side " +
                                        std::to_string(i % d->ns) + " has neither a neighbour nor a boundary row");
    if (sides[i].le >= 0 && sides[i].bc >= 0)
      return fail(DGX_ERR_ARGUMENT, "element " + std::to_string(i / d->ns) + " side " +
                                        std::to_string(i % d->ns) + " has both a neighbour and a boundary row");
  }
  static const int con2d[4][4] = {{1, 1, 0, 0}, {1, 1, 0, 0}, {0, 0, 1, 1}, {0, 0, 1, 1}};
  static const int svm[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5}, {3, 2, 6, 7}, {0, 3, 7, 4}, {4, 5, 6, 7}};
  for (int e = 0; e < ne; ++e)
    for (int s = 0; s < d->ns; ++s) {
      Side &S = sides[(size_t)e * d->ns + s];
      if (!S.master) continue;
      if (nsd == 2) { S.fa = con2d[s][S.ls]; continue; }
      auto vec = [&](int el, int v1, int v0, double o[3]) {
        const double *bb = desc->vert + (size_t)el * 3 * 8;
        double n = 0;
        for (int c = 0; c < 3; ++c) { o[c] = bb[c * 8 + v1] - bb[c * 8 + v0]; n += o[c] * o[c]; }
        n = std::sqrt(n);
        for (int c = 0; c < 3; ++c) o[c] /= n;
      };
      auto dot = [](const double *a, const double *c) { return a[0] * c[0] + a[1] * c[1] + a[2] * c[2]; };
      double a0[3], b0[3], a1[3], b1[3], a1p[3];
      vec(e, svm[s][1], svm[s][0], a0);
      vec(e, svm[s][3], svm[s][0], b0);
      vec(S.le, svm[S.ls][1], svm[S.ls][0], a1);
      vec(S.le, svm[S.ls][3], svm[S.ls][0], b1);
      vec(S.le, svm[S.ls][2], svm[S.ls][3], a1p);
      const double x1 = dot(a0, a1), x2 = dot(a0, a1p);
      if (std::fabs(std::fabs(x1) - 1) < 1e-8 || std::fabs(std::fabs(x2) - 1) < 1e-8) {
        S.tr = false; S.fa = x1 < 0; S.fb = dot(b0, b1) < 0;
      } else {
        S.tr = true; S.fa = dot(b0, a1) < 0; S.fb = dot(a0, b1) < 0;
      }
    }
  // own face point -> neighbour face point permutations (src/BaseQuad.hpp:53-57,
  // src/BaseHex.hpp:291-313): master point sideMap(g) pairs slave point g
  const int QF = d->QF, qs = d->qs;
  auto sidemap = [&](const Side &S, int q) {
    if (nsd == 2) return S.fa ? QF - 1 - q : q;
    const int il = q / QF, jl = q % QF;
    int i, j;
    if (S.tr) { j = il; i = jl; if (S.fa) j = QF - 1 - j; if (S.fb) i = QF - 1 - i; }
    else { i = il; j = jl; if (S.fa) i = QF - 1 - i; if (S.fb) j = QF - 1 - j; }
    return j + QF * i;
  };
  std::map<std::vector<uint16_t>, int> perms;
  std::vector<uint16_t> permtab;
  std::vector<uint8_t> permid((size_t)ne * d->ns), smaster((size_t)ne * d->ns);
  std::vector<int> nbr((size_t)ne * d->ns), nside((size_t)ne * d->ns);
  for (int e = 0; e < ne; ++e)
    for (int s = 0; s < d->ns; ++s) {
      const Side &S = sides[(size_t)e * d->ns + s];
      std::vector<uint16_t> m(qs);
      if (S.bc >= 0) {
        for (int g = 0; g < qs; ++g) m[g] = (uint16_t)g;
      } else if (S.master) {
        for (int g = 0; g < qs; ++g) m[sidemap(S, g)] = (uint16_t)g;
      } else {
        const Side &M = sides[(size_t)S.le * d->ns + S.ls];
        for (int g = 0; g < qs; ++g) m[g] = (uint16_t)sidemap(M, g);
      }
      auto it = perms.find(m);
      int id;
      if (it == perms.end()) {
        id = (int)perms.size();
        perms[m] = id;
        permtab.insert(permtab.end(), m.begin(), m.end());
      } else id = it->second;
      permid[(size_t)e * d->ns + s] = (uint8_t)id;
      smaster[(size_t)e * d->ns + s] = (uint8_t)(S.bc >= 0 || S.master);
      nbr[(size_t)e * d->ns + s] = S.bc >= 0 ? -1 - S.bc : S.le;   // boundary: -1 - type
      nside[(size_t)e * d->ns + s] = S.bc >= 0 ? 0 : S.ls;
    }

  // device buffers
  int rc0 = upload_all_tables();
  if (rc0) return rc0;
  const std::vector<const std::vector<double> *> tabs = {&b.BT, &b.BW, &b.QD, &b.Bg, &b.Bend, &b.PW, &b.Ig};
  size_t ntab = 0;
  for (auto t : tabs) ntab += t->size();
  int rc;
  if ((rc = alloc(&d->d_tab, ntab))) return rc;
  {
    std::vector<double> packed;
    for (auto t : tabs) packed.insert(packed.end(), t->begin(), t->end());
    HIPCHK(hipMemcpy(d->d_tab, packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice));
    const double *ptr = d->d_tab;
    const double **dst[7] = {&d->tab.BT, &d->tab.BW, &d->tab.QD, &d->tab.Bg, &d->tab.Bend, &d->tab.PW, &d->tab.Ig};
    for (int i = 0; i < 7; ++i) { *dst[i] = ptr; ptr += tabs[i]->size(); }
  }
  if ((rc = alloc(&d->d_metric, d->metric.size()))) return rc;
  if ((rc = alloc(&d->d_fnorm, fnorm.size()))) return rc;
  if ((rc = alloc(&d->d_fjr, fjr.size()))) return rc;
  HIPCHK(hipMemcpy(d->d_metric, d->metric.data(), d->metric.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_fnorm, fnorm.data(), fnorm.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_fjr, fjr.data(), fjr.size() * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMalloc((void **)&d->d_nbr, nbr.size() * sizeof(int)));
  HIPCHK(hipMalloc((void **)&d->d_nside, nside.size() * sizeof(int)));
  HIPCHK(hipMalloc((void **)&d->d_permid, permid.size()));
  HIPCHK(hipMalloc((void **)&d->d_master, smaster.size()));
  HIPCHK(hipMalloc((void **)&d->d_perm, permtab.size() * sizeof(uint16_t)));
  HIPCHK(hipMalloc((void **)&d->d_flag, 64));
  HIPCHK(hipMemcpy(d->d_nbr, nbr.data(), nbr.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_nside, nside.data(), nside.size() * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_permid, permid.data(), permid.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_master, smaster.data(), smaster.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d->d_perm, permtab.data(), permtab.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemset(d->d_flag, 0, 64));
  d->geo.metric = d->d_metric; d->geo.fnorm = d->d_fnorm; d->geo.fjr = d->d_fjr;
  d->geo.nbr = d->d_nbr; d->geo.nside = d->d_nside; d->geo.permid = d->d_permid; d->geo.perm = d->d_perm;
  d->geo.master = d->d_master;
  d->phys.gamma = d->gamma; d->phys.vis = d->vis; d->phys.Pr = d->Pr; d->phys.flux = d->flux;
  d->phys.visc = desc->viscous;
  for (int v = 0; v < 5; ++v) d->phys.bcU[v] = desc->bc_state[v];
  d->phys.wallT = desc->wall_T;
  d->phys.wallE = desc->wall_Eref;
  d->phys.eta = desc->br2_eta > 0.0 ? desc->br2_eta : 6.0;

  const size_t n = d->state_len();
  if ((rc = alloc(&d->U, n)) || (rc = alloc(&d->Ut, n)) || (rc = alloc(&d->R0, n)) || (rc = alloc(&d->R1, n)))
    return rc;
  HIPCHK(hipMemset(d->U, 0, n * 8));
  HIPCHK(hipMemset(d->Ut, 0, n * 8));
  HIPCHK(hipMemset(d->R0, 0, n * 8));
  HIPCHK(hipMemset(d->R1, 0, n * 8));
  if ((rc = alloc(&d->T, (size_t)ne * d->ns * d->nf * qs))) return rc;
  if (d->eqn == DGX_SHALLOW && !desc->no_balance) {
    if ((rc = alloc(&d->Bal, n))) return rc;
    d->bal = 1;
  }
  if (desc->xnodes && (rc = (nsd == 2 ? curved_geometry2(d, desc->xnodes) : curved_geometry(d, desc->xnodes))))
    return rc;
  if (!desc->xnodes && (nonaffine || desc->bc_function || desc->source)) {
    // the curved path on the vertices' bilinear map (the element coordinates),
    // as the reference gives non-parallelogram quads a straight-sided curve
    const std::vector<double> xn = d->coords;
    if ((rc = curved_geometry2(d, xn.data()))) return rc;
  }
  if ((desc->bc_function || desc->source) && (rc = setup_functions(d, desc))) return rc;
  d->fused = (nsd == 3 && dgx::stage3_supported(p) && !d->curved);
  if (nsd == 2) {
    // CFL geometry: the coordinate derivatives x_r, x_s, y_r, y_s and J of
    // Element::compute_metrics_2d_tensor_prod (src/Element.cpp:485-555), from
    // the collocation derivative of the nodal coordinates as local_gradient
    // takes them; constant on parallelograms, per node on curved quads
    const int Q2 = d->nq, S = d->curved ? Q2 : 1;
    std::vector<double> cg((size_t)ne * 5 * S);
    const std::vector<double> &D = b.D;
    for (int e = 0; e < ne; ++e) {
      const double *X = &d->coords[(size_t)e * Q2], *Y = &d->coords[((size_t)ne + e) * Q2];
      for (int n = 0; n < S; ++n) {
        const int i = n / Q, j = n % Q;
        double xr = 0, xs = 0, yr = 0, ys = 0;
        for (int m = 0; m < Q; ++m) {
          xr += D[i * Q + m] * X[m * Q + j]; xs += D[j * Q + m] * X[i * Q + m];
          yr += D[i * Q + m] * Y[m * Q + j]; ys += D[j * Q + m] * Y[i * Q + m];
        }
        double *o = &cg[(size_t)e * 5 * S + n];
        o[0] = xr; o[S] = xs; o[2 * S] = yr; o[3 * S] = ys; o[4 * S] = xr * ys - xs * yr;
      }
    }
    if ((rc = alloc(&d->d_cgeo, cg.size()))) return rc;
    HIPCHK(hipMemcpy(d->d_cgeo, cg.data(), cg.size() * 8, hipMemcpyHostToDevice));
  }
  if (desc->bulk_forcing) {
    d->bulk = true;
    if ((rc = alloc(&d->d_J, ne))) return rc;
    HIPCHK(hipMemcpy(d->d_J, d->J.data(), ne * sizeof(double), hipMemcpyHostToDevice));
  }
  // a second trace buffer: the fused hex stages write the next stage's traces
  // into it; affine quads keep it for the serving loop's persistent slices
  if ((d->fused || (d->nsd == 2 && !d->curved && !d->varp)) &&
      (rc = alloc(&d->T2, (size_t)ne * d->ns * d->nf * qs)))
    return rc;
  d->mixed = d->prec == DGX_MIXED;
  if (d->visc && d->mixed) {
    const size_t nv = (size_t)ne * (d->nf - 1) * d->nm, ntv = (size_t)ne * d->ns * (d->nf - 1) * qs;
    if (hipMalloc((void **)&d->Vf, nv * sizeof(float) + 64) != hipSuccess ||
        hipMalloc((void **)&d->TVf, ntv * sizeof(float) + 64) != hipSuccess)
      return fail(DGX_ERR_MEMORY, "hipMalloc of the FP32 viscous terms failed");
    HIPCHK(hipMemset(d->Vf, 0, nv * sizeof(float)));
    HIPCHK(hipMemset(d->TVf, 0, ntv * sizeof(float)));
  } else if (d->visc) {
    const int ng = nsd == 2 ? 6 : 12;
    if (!d->fused && (rc = alloc(&d->G, (size_t)ne * ng * d->nm))) return rc;
    // curved elements keep the weighted nodal viscous divergence (Q^nsd per field)
    if ((rc = alloc(&d->V, (size_t)ne * (d->nf - 1) * (d->curved ? d->nq : d->nm)))) return rc;
    if ((rc = alloc(&d->TV, (size_t)ne * d->ns * (d->nf - 1) * qs))) return rc;
  }
  if (desc->porder && (rc = varp_setup(d, desc, sides))) return rc;
  return DGX_OK;
}

void release(dgx_domain *d) {
  void *ptrs[] = {d->d_tab, d->d_metric, d->d_fnorm, d->d_fjr, d->d_nbr, d->d_nside, d->d_flag,
                  d->d_permid, d->d_perm, d->U, d->Ut, d->R0, d->R1, d->T, d->G, d->V, d->TV, d->nodal,
                  d->T2, d->Vf, d->TVf, d->d_gm, d->d_fgeo, d->d_jav, d->d_pre, d->d_iters, d->Bal, d->d_master,
                  d->d_fcode, d->d_fconst, d->d_fxy, d->d_xn, d->d_J, d->d_cgeo, d->d_part, d->d_chol,
                  d->d_velems, d->d_vqs, d->d_vflip, d->d_vtab};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (d->own_stream && d->stream) (void)hipStreamDestroy(d->stream);
}

}  // namespace

extern "C" {

const char *dgx_version(void) { return DGX_VERSION; }
const char *dgx_last_error(void) { return g_err.c_str(); }

int dgx_domain_bytes(const dgx_domain_desc *desc, size_t *bytes) {
  if (!desc || !bytes) return fail(DGX_ERR_ARGUMENT, "null argument");
  if ((desc->nsd != 2 && desc->nsd != 3) || desc->p < 1 || desc->p > 10 || desc->ne <= 0)
    return fail(DGX_ERR_ARGUMENT, "bad domain description");
  const size_t nsd = desc->nsd, ne = desc->ne, L = desc->p + 1, Q = desc->p + 2, QF = desc->p + 1;
  const bool shallow = desc->eqntype == DGX_SHALLOW;
  const size_t nf = shallow ? 3 : nsd + 2, nm = nsd == 2 ? L * L : L * L * L, nq = nsd == 2 ? Q * Q : Q * Q * Q;
  const size_t ns = nsd == 2 ? 4 : 6, qs = nsd == 2 ? QF : QF * QF;
  const bool visc = desc->eqntype == DGX_NAVIER_STOKES || desc->eqntype == DGX_NAVIER_STOKES3D;
  const bool fused = nsd == 3 && dgx::stage3_supported(desc->p) && !desc->xnodes;
  const bool mixed = desc->precision == DGX_MIXED;
  size_t b = 4 * nf * ne * nm * 8;                          // U, Ut, R0, R1
  b += (fused || (nsd == 2 && !desc->porder) ? 2 : 1) * ne * ns * nf * qs * 8;   // face traces
  if (visc) {
    const size_t w = mixed ? 4 : 8;
    if (!fused && !mixed) b += ne * (nsd == 2 ? 6 : 12) * nm * 8;
    b += ne * (nf - 1) * (desc->xnodes ? nq : nm) * w + ne * ns * (nf - 1) * qs * w;
  }
  b += nf * ne * nq * 8;                                    // nodal transfer buffer
  b += ne * (9 + 4 * ns) * 8 + ne * ns * (4 + 4 + 1);       // geometry and links
  b += 8 * qs * 2 + (size_t)(1 << 20);                      // orientation tables, tables, rounding
  const bool fns = desc->bc_function || desc->source;
  if (desc->xnodes || fns || !all_quads_affine(desc))   // curved: nodal metrics, face geometry
    b += nsd == 2 ? ne * (5 * nq + 16 * qs + 1) * 8 : ne * (10 * nq + 24 * qs + 1) * 8;
  if (fns) {                 // programs, side Gauss-point and nodal coordinates
    for (const dgx_function *f : {desc->bc_function, desc->source})
      if (f && f->start && f->ncomp > 0 && f->ncomp <= 16 && f->start[f->ncomp] > 0 && f->nconst >= 0)
        b += (size_t)f->start[f->ncomp] * 4 + (size_t)f->nconst * 8;
    b += ne * (8 * qs + 2 * nq) * 8;
  }
  if (shallow && !desc->no_balance) b += nf * ne * nm * 8;    // balance term
  if (nsd == 2) b += ne * 5 * ((desc->xnodes || fns || !all_quads_affine(desc)) ? nq : 1) * 8;   // CFL geometry
  if (nsd == 2 && (desc->xnodes || fns || !all_quads_affine(desc)) && ne * nm * (nm + 1) / 2 * 8 <= (size_t(1) << 30))
    b += ne * nm * (nm + 1) / 2 * 8;                        // curved quads: mass-matrix factors
  if (desc->bulk_forcing) b += ne * 8;
  *bytes = b;
  return DGX_OK;
}

int dgx_domain_create(const dgx_domain_desc *desc, void *stream, dgx_handle *out) {
  if (!desc || !out) return fail(DGX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (desc->nsd != 2 && desc->nsd != 3) return fail(DGX_ERR_ARGUMENT, "nsd must be 2 or 3");
  if (desc->p < 1 || desc->p > 10) return fail(DGX_ERR_ARGUMENT, "p must be in [1, 10]");
  if (desc->ne <= 0 || !desc->vert || !desc->con) return fail(DGX_ERR_ARGUMENT, "empty mesh");
  const bool euler = desc->eqntype == DGX_EULER || desc->eqntype == DGX_NAVIER_STOKES ||
                     desc->eqntype == DGX_SHALLOW;
  const bool euler3 = desc->eqntype == DGX_EULER3D || desc->eqntype == DGX_NAVIER_STOKES3D;
  if ((desc->nsd == 2 && !euler) || (desc->nsd == 3 && !euler3))
    return fail(DGX_ERR_ARGUMENT, "eqntype does not match nsd");
  if (desc->eqntype == DGX_SHALLOW && desc->flux != DGX_FLUX_LAX_FRIEDRICHS)
    return fail(DGX_ERR_ARGUMENT, "Shallow has the Lax-Friedrichs flux only");
  if (desc->flux != DGX_FLUX_LAX_FRIEDRICHS && desc->flux != DGX_FLUX_ROE)
    return fail(DGX_ERR_ARGUMENT, "flux must be DGX_FLUX_LAX_FRIEDRICHS or DGX_FLUX_ROE");
  if (desc->precision != DGX_FP64 && desc->precision != DGX_MIXED)
    return fail(DGX_ERR_ARGUMENT, "precision must be DGX_FP64 or DGX_MIXED");
  if (desc->viscous != DGX_VISCOUS_BR1 && desc->viscous != DGX_VISCOUS_BR2)
    return fail(DGX_ERR_ARGUMENT, "viscous must be DGX_VISCOUS_BR1 or DGX_VISCOUS_BR2");
  if (desc->viscous == DGX_VISCOUS_BR2 && desc->eqntype != DGX_NAVIER_STOKES3D)
    return fail(DGX_ERR_ARGUMENT, "DGX_VISCOUS_BR2 is available for Navier_Stokes3d");
  if (!(desc->br2_eta >= 0.0) || !std::isfinite(desc->br2_eta))
    return fail(DGX_ERR_ARGUMENT, "br2_eta must be finite and >= 0");
  if (desc->precision == DGX_MIXED &&
      !((desc->nsd == 3 && dgx::mixed3_supported(desc->p)) ||
        (desc->nsd == 2 && (desc->eqntype == DGX_EULER || desc->eqntype == DGX_SHALLOW))))
    return fail(DGX_ERR_ARGUMENT, "DGX_MIXED is available for hexes (Euler3d, Navier_Stokes3d) with p <= 8 and for "
                                  "affine quads (Euler, Shallow)");
  const bool curved_hex_mixed = desc->precision == DGX_MIXED && desc->nsd == 3 && desc->xnodes &&
                                desc->eqntype == DGX_EULER3D && dgx::curved3_supported(desc->p);
  if (desc->precision == DGX_MIXED && (desc->bc_function || desc->source))
    return fail(DGX_ERR_ARGUMENT, "DGX_MIXED: boundary and source functions take the curved path, which is DGX_FP64");
  if (!desc->xnodes && !all_quads_affine(desc) && desc->precision != DGX_FP64)
    return fail(DGX_ERR_ARGUMENT, "non-parallelogram quads take the curved path, which is DGX_FP64");
  const bool curved_hex_ns = desc->nsd == 3 && desc->eqntype == DGX_NAVIER_STOKES3D &&
                             desc->viscous == DGX_VISCOUS_BR1 && dgx::curved3_ns_supported(desc->p);
  if (desc->xnodes && !((desc->precision == DGX_FP64 || curved_hex_mixed) &&
                        ((desc->nsd == 3 && desc->eqntype == DGX_EULER3D && dgx::curved3_supported(desc->p)) ||
                         curved_hex_ns ||
                         (desc->nsd == 2 && (desc->eqntype == DGX_EULER || desc->eqntype == DGX_NAVIER_STOKES ||
                                             desc->eqntype == DGX_SHALLOW)))))
    return fail(DGX_ERR_ARGUMENT,
                "curved elements (xnodes) are available for Euler3d (p <= 8, DGX_FP64 or DGX_MIXED), "
                "Navier_Stokes3d (p <= 7, DGX_VISCOUS_BR1, DGX_FP64), Euler, Navier_Stokes and Shallow (DGX_FP64)");
  const bool ns3 = desc->eqntype == DGX_NAVIER_STOKES3D;
  for (int r = 0; r < desc->ncon; ++r)
    if (desc->con[4 * r + 2] == -1) {
      if (desc->eqntype != DGX_EULER3D && desc->eqntype != DGX_SHALLOW && desc->eqntype != DGX_EULER && !ns3)
        return fail(DGX_ERR_ARGUMENT, "boundary sides are available for Euler, Euler3d, Navier_Stokes3d and Shallow");
      const int t = desc->con[4 * r + 3];
      if (t != DGX_BC_WALL && t != DGX_BC_STATE && t != DGX_BC_FUNCTION && t != DGX_BC_ISOTHERMAL)
        return fail(DGX_ERR_ARGUMENT, "connectivity row " + std::to_string(r) + ": unknown boundary type");
      if (t == DGX_BC_ISOTHERMAL && !ns3)
        return fail(DGX_ERR_ARGUMENT, "connectivity row " + std::to_string(r) +
                                          ": DGX_BC_ISOTHERMAL is a Navier_Stokes3d boundary");
      if (ns3 && t == DGX_BC_WALL)
        return fail(DGX_ERR_ARGUMENT, "connectivity row " + std::to_string(r) +
                                          ": Navier_Stokes3d has no slip-wall viscous flux "
                                          "(Euler3d::WallBC::apply_flux); use DGX_BC_ISOTHERMAL");
      if (ns3 && (desc->precision != DGX_FP64 || desc->viscous != DGX_VISCOUS_BR1 ||
                  !(desc->xnodes ? dgx::curved3_ns_supported(desc->p) : dgx::stage3_supported(desc->p))))
        return fail(DGX_ERR_ARGUMENT, "Navier_Stokes3d boundaries are available for affine hexes (p <= 9) and "
                                      "curved hexes (p <= 7), DGX_FP64 and DGX_VISCOUS_BR1");
      if (t == DGX_BC_ISOTHERMAL && !(std::isfinite(desc->wall_T) && std::isfinite(desc->wall_Eref)))
        return fail(DGX_ERR_ARGUMENT, "wall_T and wall_Eref must be finite");
      if (t == DGX_BC_FUNCTION && !desc->bc_function)
        return fail(DGX_ERR_ARGUMENT, "connectivity row " + std::to_string(r) +
                                          ": DGX_BC_FUNCTION boundary without bc_function");
    }
  if (desc->bulk_forcing && !(desc->nsd == 3 && desc->precision == DGX_FP64 && !desc->xnodes))
    return fail(DGX_ERR_ARGUMENT, "bulk_forcing is available for affine hexes, DGX_FP64");
  if ((desc->bc_function || desc->source) && desc->eqntype != DGX_SHALLOW)
    return fail(DGX_ERR_ARGUMENT, "boundary and source functions are available for Shallow");
  if (desc->porder) {   // variable order (root.ord)
    if (desc->nsd != 2 || !(desc->eqntype == DGX_EULER || (desc->eqntype == DGX_NAVIER_STOKES && desc->vis == 0.0)))
      return fail(DGX_ERR_ARGUMENT, "porder: variable order is built for quads, Euler (and Navier_Stokes with vis = 0)");
    if (desc->precision != DGX_FP64 || desc->xnodes || desc->bc_function || desc->source || !all_quads_affine(desc))
      return fail(DGX_ERR_ARGUMENT, "porder: variable order needs affine quads without functions, DGX_FP64");
    int pmax = 0;
    for (int e = 0; e < desc->ne; ++e) {
      if (desc->porder[e] < 1 || desc->porder[e] > desc->p)
        return fail(DGX_ERR_ARGUMENT, "porder: element " + std::to_string(e) + " has order " +
                                          std::to_string(desc->porder[e]) + " outside 1..p");
      pmax = std::max(pmax, desc->porder[e]);
    }
    if (pmax != desc->p) return fail(DGX_ERR_ARGUMENT, "porder: p must be the largest element order");
  }
  {
    std::string why;
    if (desc->bc_function && !check_function(desc->bc_function, 3, &why))
      return fail(DGX_ERR_ARGUMENT, "bc_function: " + why);
    if (desc->source && !check_function(desc->source, 3, &why))
      return fail(DGX_ERR_ARGUMENT, "source: " + why);
  }
  if (desc->eqntype == DGX_NAVIER_STOKES3D && !dgx::stage3_supported(desc->p) &&
      dgx::lds_bytes3(desc->p, 3) > 160 * 1024)
    return fail(DGX_ERR_ARGUMENT, "Navier_Stokes3d supports p <= 9 (LDS)");
  dgx_domain *d = new dgx_domain();
  if (stream) d->stream = (hipStream_t)stream;
  else {
    if (hipStreamCreate(&d->stream) != hipSuccess) { delete d; return fail(DGX_ERR_DEVICE, "hipStreamCreate failed"); }
    d->own_stream = true;
  }
  int rc = build(d, desc);
  if (rc) { release(d); delete d; return rc; }
  *out = d;
  return DGX_OK;
}

int dgx_domain_destroy(dgx_handle h) {
  if (!h) return DGX_OK;
  (void)hipStreamSynchronize(h->stream);
  release(h);
  delete h;
  return DGX_OK;
}

int dgx_set_time(dgx_handle h, double t) {
  if (!h) return fail(DGX_ERR_ARGUMENT, "null handle");
  if (!std::isfinite(t)) return fail(DGX_ERR_ARGUMENT, "time must be finite");
  h->time = t;
  return DGX_OK;
}

int dgx_get_time(dgx_handle h, double *t) {
  if (!h || !t) return fail(DGX_ERR_ARGUMENT, "null argument");
  *t = h->time;
  return DGX_OK;
}

int dgx_sizes(dgx_handle h, int *qtot, int *nmodes, int *nfields, int *ne) {
  if (!h) return fail(DGX_ERR_ARGUMENT, "null handle");
  if (qtot) *qtot = h->nq;
  if (nmodes) *nmodes = h->nm;
  if (nfields) *nfields = h->nf;
  if (ne) *ne = h->ne;
  return DGX_OK;
}

int dgx_vfield_sizes(dgx_handle h, int *size, long long *qtot, long long *nmodes, long long *sdof, int *nsd) {
  if (!h) return fail(DGX_ERR_ARGUMENT, "null handle");
  const long long nf = h->nf, ne = h->ne;
  if (size) *size = h->nf;
  if (qtot) *qtot = nf * ne * h->nq;
  if (nmodes) *nmodes = nf * ne * h->nm;
  if (sdof) *sdof = nf * ne * h->ns * h->qs;
  if (nsd) *nsd = h->nsd;
  return DGX_OK;
}

int dgx_coordinates(dgx_handle h, double *xyz) {
  if (!h || !xyz) return fail(DGX_ERR_ARGUMENT, "null argument");
  std::memcpy(xyz, h->coords.data(), h->coords.size() * sizeof(double));
  return DGX_OK;
}

int dgx_state_device_ptr(dgx_handle h, double **U) {
  if (!h || !U) return fail(DGX_ERR_ARGUMENT, "null argument");
  *U = h->U;
  return DGX_OK;
}

int dgx_state_upload(dgx_handle h, const double *U, int layout, int on_device) {
  if (!h || !U) return fail(DGX_ERR_ARGUMENT, "null argument");
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  if (layout == DGX_MODAL) {
    HIPCHK(hipMemcpyAsync(h->U, U, h->state_len() * 8, kind, h->stream));
    return DGX_OK;
  }
  int rc = ensure_nodal(h);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(h->nodal, U, (size_t)h->nf * h->ne * h->nq * 8, kind, h->stream));
  return transform(h, dgx::OP_FORWARD, h->nodal, h->U);
}

static int download(dgx_handle h, const double *src, double *dst, int layout, int on_device) {
  const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (layout == DGX_MODAL) {
    HIPCHK(hipMemcpyAsync(dst, src, h->state_len() * 8, kind, h->stream));
  } else {
    int rc = ensure_nodal(h);
    if (rc) return rc;
    if ((rc = transform(h, dgx::OP_BACKWARD, src, h->nodal))) return rc;
    HIPCHK(hipMemcpyAsync(dst, h->nodal, (size_t)h->nf * h->ne * h->nq * 8, kind, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return DGX_OK;
}

int dgx_state_download(dgx_handle h, double *U, int layout, int on_device) {
  if (!h || !U) return fail(DGX_ERR_ARGUMENT, "null argument");
  return download(h, h->U, U, layout, on_device);
}

int dgx_rhs(dgx_handle h) {
  if (!h) return fail(DGX_ERR_ARGUMENT, "null handle");
  return residual(h, h->U, h->R0);
}

int dgx_residual_download(dgx_handle h, double *F, int layout, int on_device) {
  if (!h || !F) return fail(DGX_ERR_ARGUMENT, "null argument");
  return download(h, h->R0, F, layout, on_device);
}

}  // extern "C"

namespace {

int nstages(int inttype) {
  return inttype == DGX_RK4 ? 4 : (inttype == DGX_TVD ? 3 : (inttype == DGX_LSRK14 ? 14 : 0));
}

// Stage `st` of one step of integrator `inttype`, as the reference's
// TimeInt::advance runs it (RungeKutta src/TimeInt.cpp:267-302, TVD :824-856,
// LSRK14 :1989-2013); the step starts at the domain time of stage 0.
// `traces_ready`: the fused LSRK14 path may take the face traces the previous
// stage wrote instead of evaluating them.
int run_stage(dgx_domain *h, int inttype, int st, double dt, bool traces_ready) {
  const size_t n = h->state_len();
  hipStream_t s = h->stream;
  int rc;
#define RUN(x) do { if ((rc = (x))) return rc; } while (0)
#define BL(x) do { hipError_t _e = timed(h, 6, [&] { return (x); }); if (_e != hipSuccess) return fail(DGX_ERR_DEVICE, hipGetErrorString(_e)); } while (0)
  if (st == 0) h->t_step = h->time;
  const double tn = h->t_step;
  if (inttype == DGX_RK4) {
    static const double a[3] = {0.5, 0.5, 1.0};
    static const double b[4] = {0.16666666666666667, 0.33333333333333333, 0.33333333333333333,
                                0.16666666666666667};
    if (st == 0) {
      RUN(residual(h, h->U, h->R0));
      BL(dgx::blas_axpy(-a[0] * dt, h->R0, h->U, h->Ut, n, s));
      BL(dgx::blas_scale(-b[0], h->R0, n, s));
      h->time = tn + 0.5 * dt;
    } else if (st < 3) {
      RUN(residual(h, h->Ut, h->R1));
      BL(dgx::blas_axpy(-a[st] * dt, h->R1, h->U, h->Ut, n, s));
      BL(dgx::blas_axpy(-b[st], h->R1, h->R0, h->R0, n, s));
      if (st == 2) h->time = tn + dt;
    } else {
      RUN(residual(h, h->Ut, h->R1));
      BL(dgx::blas_axpy(-b[3], h->R1, h->R0, h->R0, n, s));
      BL(dgx::blas_axpy(dt, h->R0, h->U, h->U, n, s));
    }
  } else if (inttype == DGX_TVD) {
    // stage times c = 0, 1, 1/2 (src/TimeInt.cpp:740)
    if (st == 0) {
      RUN(residual(h, h->U, h->R0));
      BL(dgx::blas_axpy(-dt, h->R0, h->U, h->Ut, n, s));
      h->time = tn + dt;
    } else if (st == 1) {
      RUN(residual(h, h->Ut, h->R0));
      BL(dgx::blas_axpy(-dt, h->R0, h->Ut, h->R0, n, s));
      BL(dgx::blas_axpby(3.0 / 4.0, h->U, 1.0 / 4.0, h->R0, h->Ut, n, s));
      h->time = tn + 0.5 * dt;
    } else {
      RUN(residual(h, h->Ut, h->R0));
      BL(dgx::blas_axpy(-dt, h->R0, h->Ut, h->R0, n, s));
      BL(dgx::blas_axpby(2.0 / 3.0, h->R0, 1.0 / 3.0, h->U, h->U, n, s));
      h->time = tn + dt;
    }
  } else {
    const double A = dgx::kLsrkA[st], B = dgx::kLsrkB[st];
    h->time = tn + dgx::kLsrkC[st] * dt;
    if (h->fused && !h->bulk) {
      // one fused kernel per stage: residual, register update, next traces
      // (the bulk forcing needs the whole residual before any update)
      if (!traces_ready) RUN(launch(h, dgx::OP_TRACES, h->U, nullptr));
      if (h->visc) RUN(vterm(h, h->U, h->T));
      dgx::StageLaunch sa = stage_args(h);
      sa.op = 1; sa.update = true; sa.U = h->U; sa.Ut = h->Ut; sa.T = h->T; sa.Tn = h->T2;
      sa.V = h->V; sa.TV = h->TV; sa.Vf = h->Vf; sa.TVf = h->TVf;
      sa.ca = A; sa.cb = B; sa.dt = dt;
      RUN(launch_stage(h, sa));
      std::swap(h->T, h->T2);
      ++h->stages;
    } else if (h->nsd == 2 && !h->curved && !h->varp && !h->bulk) {
      RUN(step2(h, A, B, dt, traces_ready));
    } else if (h->nsd == 3 && h->curved && !h->bulk) {
      RUN(step_curved3(h, A, B, dt));
    } else {
      RUN(residual(h, h->U, h->R0));
      BL(dgx::blas_axpby(-dt, h->R0, A, h->Ut, h->Ut, n, s));
      BL(dgx::blas_axpy(B, h->Ut, h->U, h->U, n, s));
    }
    if (st == 13) h->time = tn + dt;
  }
#undef RUN
#undef BL
  return DGX_OK;
}

}  // namespace

extern "C" {

int dgx_advance(dgx_handle h, int inttype, double dt, int nsteps) {
  if (!h) return fail(DGX_ERR_ARGUMENT, "null handle");
  const int ns = nstages(inttype);
  if (!ns) return fail(DGX_ERR_ARGUMENT, "unsupported inttype " + std::to_string(inttype));
  if (h->next_stage != 0)
    return fail(DGX_ERR_ARGUMENT, "a step begun with dgx_stage is unfinished (next stage " +
                                      std::to_string(h->next_stage) + ")");
  for (int step = 0; step < nsteps; ++step)
    for (int st = 0; st < ns; ++st) {
      int rc = run_stage(h, inttype, st, dt, st > 0 || step > 0);
      if (rc) return rc;
    }
  return DGX_OK;
}

int dgx_stage(dgx_handle h, int inttype, int stage, double dt) {
  if (!h) return fail(DGX_ERR_ARGUMENT, "null handle");
  const int ns = nstages(inttype);
  if (!ns) return fail(DGX_ERR_ARGUMENT, "unsupported inttype " + std::to_string(inttype));
  if (stage != h->next_stage || (stage > 0 && inttype != h->stage_int))
    return fail(DGX_ERR_ARGUMENT, "stages run in order 0 .. " + std::to_string(ns - 1) +
                                      " of one integrator; expected stage " + std::to_string(h->next_stage));
  int rc = run_stage(h, inttype, stage, dt, stage > 0);
  if (rc) return rc;
  h->stage_int = inttype;
  h->next_stage = stage + 1 == ns ? 0 : stage + 1;
  return DGX_OK;
}

int dgx_set_time_step(dgx_handle h, double dt, double CFLmax, double dtmax, double *cfl, double *dt_new) {
  if (!h || !cfl || !dt_new) return fail(DGX_ERR_ARGUMENT, "null argument");
  if (!(dt > 0) || !std::isfinite(dt) || !(CFLmax >= 0) || !(dtmax >= 0))
    return fail(DGX_ERR_ARGUMENT, "dt must be positive and CFLmax, dtmax >= 0");
  if (h->nsd == 3) {   // Euler3d::set_time_step (src/Euler3d.cpp:462-465) returns dt and keeps it
    *cfl = dt;
    *dt_new = dt;
    return DGX_OK;
  }
  if (!h->d_part) {
    int rc = alloc(&h->d_part, (size_t)h->ne + 1);
    if (rc) return rc;
  }
  double m = 0.0;
  int rc = dgx::cfl_max(h, h->d_part, &m);
  if (rc) return rc;
  return dgx::cfl_finish(dt, CFLmax, dtmax, m, cfl, dt_new);
}

int dgx_synchronize(dgx_handle h) {
  if (!h) return fail(DGX_ERR_ARGUMENT, "null handle");
  HIPCHK(hipStreamSynchronize(h->stream));
  int flag = 0;
  HIPCHK(hipMemcpy(&flag, h->d_flag, sizeof(int), hipMemcpyDeviceToHost));
  if (flag) {
    HIPCHK(hipMemset(h->d_flag, 0, sizeof(int)));
    return fail(DGX_ERR_NONREALIZABLE, "density or pressure below aepsM in a residual evaluation");
  }
  return DGX_OK;
}

long long dgx_stage_count(dgx_handle h) { return h ? h->stages : -1; }

int dgx_solver_iterations(dgx_handle h, int *max_iters) {
  if (!h || !max_iters) return fail(DGX_ERR_ARGUMENT, "null argument");
  *max_iters = 0;
  if (!h->d_iters) return DGX_OK;
  HIPCHK(hipStreamSynchronize(h->stream));
  HIPCHK(hipMemcpy(max_iters, h->d_iters, sizeof(int), hipMemcpyDeviceToHost));
  return DGX_OK;
}

int dgx_profile(dgx_handle h, int inttype, double dt, int nsteps, double *ms, int *count, int nkinds) {
  if (!h || !ms || nkinds < 7) return fail(DGX_ERR_ARGUMENT, "dgx_profile needs 7 slots");
  for (int k = 0; k < nkinds; ++k) { ms[k] = 0.0; if (count) count[k] = 0; }
  h->profiling = true;
  h->ev.clear();
  h->ev_kind.clear();
  int rc = dgx_advance(h, inttype, dt, nsteps);
  h->profiling = false;
  if (!rc) rc = (hipStreamSynchronize(h->stream) == hipSuccess) ? DGX_OK : fail(DGX_ERR_DEVICE, "sync");
  for (size_t i = 0; i < h->ev_kind.size(); ++i) {
    float t = 0.f;
    if (!rc && hipEventElapsedTime(&t, h->ev[2 * i], h->ev[2 * i + 1]) == hipSuccess) {
      ms[h->ev_kind[i]] += t;
      if (count) count[h->ev_kind[i]] += 1;
    }
  }
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  h->ev.clear();
  h->ev_kind.clear();
  return rc;
}

int dgx_norms(dgx_handle h, double *out) {
  if (!h || !out) return fail(DGX_ERR_ARGUMENT, "null argument");
  std::vector<double> Un((size_t)h->nf * h->ne * h->nq);
  int rc = dgx_state_download(h, Un.data(), DGX_NODAL, 0);
  if (rc) return rc;
  const int nsd = h->nsd, NQ = h->nq;
  std::vector<HostBasis> vb;   // variable order: each element's own basis
  if (h->varp)
    for (int P = 0; P <= h->p; ++P) vb.push_back(P ? make_basis(P) : HostBasis());
  std::vector<double> fr(NQ), fs(NQ), ft(NQ);
  for (int v = 0; v < h->nf; ++v) {
    double Linf = 0, L2 = 0, H1 = 0, area = 0;
    for (int e = 0; e < h->ne; ++e) {
      const HostBasis &eb = h->varp ? vb[h->porder[e]] : h->hb;
      const int Q = eb.Q, nq = nsd == 2 ? Q * Q : Q * Q * Q;
      const std::vector<double> &w = eb.w, &D = eb.D;
      const double *u = &Un[((size_t)v * h->ne + e) * NQ];
      const double *mt = &h->metric[(size_t)e * 9];
      std::fill(fr.begin(), fr.end(), 0.0); std::fill(fs.begin(), fs.end(), 0.0); std::fill(ft.begin(), ft.end(), 0.0);
      if (nsd == 2) {
        for (int i = 0; i < Q; ++i)
          for (int j = 0; j < Q; ++j)
            for (int k = 0; k < Q; ++k) {
              fr[i * Q + j] += D[i * Q + k] * u[k * Q + j];
              fs[i * Q + j] += D[j * Q + k] * u[i * Q + k];
            }
      } else {
        for (int i = 0; i < Q; ++i)
          for (int j = 0; j < Q; ++j)
            for (int k = 0; k < Q; ++k)
              for (int m = 0; m < Q; ++m) {
                fr[(i * Q + j) * Q + k] += D[i * Q + m] * u[(m * Q + j) * Q + k];
                fs[(i * Q + j) * Q + k] += D[j * Q + m] * u[(i * Q + m) * Q + k];
                ft[(i * Q + j) * Q + k] += D[k * Q + m] * u[(i * Q + j) * Q + m];
              }
      }
      const double *gm = h->curved ? &h->gm_host[(size_t)e * (nsd == 2 ? 5 : 10) * nq] : nullptr;
      for (int n = 0; n < nq; ++n) {
        double wj, gx, gy, gz = 0.0;
        if (gm) {   // nodal metrics: gm [d * nsd + x] = wJ dxi_d/dx_x, then wJ
          wj = gm[(size_t)nsd * nsd * nq + n];
          const double f[3] = {fr[n], fs[n], ft[n]};
          double g[3] = {0.0, 0.0, 0.0};
          for (int dd = 0; dd < nsd; ++dd)
            for (int x = 0; x < nsd; ++x) g[x] += gm[(size_t)(dd * nsd + x) * nq + n] * f[dd];
          gx = g[0] / wj; gy = g[1] / wj; gz = g[2] / wj;
        } else {
          wj = h->J[e];
          if (nsd == 2) wj *= w[n / Q] * w[n % Q];
          else wj *= w[n / (Q * Q)] * w[(n / Q) % Q] * w[n % Q];
          gx = mt[0] * fr[n] + mt[3] * fs[n] + mt[6] * ft[n];
          gy = mt[1] * fr[n] + mt[4] * fs[n] + mt[7] * ft[n];
          gz = mt[2] * fr[n] + mt[5] * fs[n] + mt[8] * ft[n];
        }
        Linf = std::max(Linf, std::fabs(u[n]));
        L2 += wj * u[n] * u[n];
        H1 += wj * (u[n] * u[n] + gx * gx + gy * gy + gz * gz);
        area += wj;
      }
    }
    out[3 * v] = Linf;
    out[3 * v + 1] = std::sqrt(std::fabs(L2 / area));
    out[3 * v + 2] = std::sqrt(std::fabs(H1 / area));
  }
  return DGX_OK;
}

}  // extern "C"
