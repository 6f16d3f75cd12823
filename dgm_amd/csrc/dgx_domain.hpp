// The domain behind a dgx_handle, shared by the host units of the library
// (dgx_api.cpp, dgx_adjoint.cpp).  Internal: not part of the C ABI.
#pragma once
#include "../../include/dgx.h"
#include "dgx_kernels.hpp"

#include <hip/hip_runtime.h>

#include <vector>

namespace dgx {

// LSRK14 coefficients A, B and stage times c (src/TimeInt.cpp:1870-1917):
// numeric constants that define the integrator.
inline constexpr double kLsrkA[14] = {0.0, -0.7188012108672410, -0.7785331173421570, -0.0053282796654044,
    -0.8552979934029281, -3.9564138245774565, -1.5780575380587385, -2.0837094552574054, -0.7483334182761610,
    -0.7032861106563359, 0.0013917096117681, -0.0932075369637460, -0.9514200470875948, -7.1151571693922548};
inline constexpr double kLsrkB[14] = {0.0367762454319673, 0.3136296607553959, 0.1531848691869027,
    0.0030097086818182, 0.3326293790646110, 0.2440251405350864, 0.3718879239592277, 0.6204126221582444,
    0.1524043173028741, 0.0760894927419266, 0.0077604214040978, 0.0024647284755382, 0.0780348340049386,
    5.5059777270269628};
inline constexpr double kLsrkC[14] = {0.0, 0.0367762454319673, 0.1249685262725025, 0.2446177702277698,
    0.2476149531070420, 0.2969311120382472, 0.3978149645802642, 0.5270854589440328, 0.6981269994175695,
    0.8190890835352128, 0.8527059887098624, 0.8604711817462826, 0.8627060376969976, 0.8734213127600976};

// Basis tables of one order (host copies).
struct HostBasis {
  int p, L, Q, QF;
  std::vector<double> z, w, D, BT, BW, QD, Bg, Bend, PW, Ig, PI, KE, BL, GL, NP1, NPD;
};

}  // namespace dgx

struct dgx_domain {
  int nsd = 0, ne = 0, p = 0, L = 0, Q = 0, QF = 0, nf = 0, nm = 0, nq = 0, ns = 0, qs = 0;
  int eqn = 0, flux = 0, prec = 0;
  bool visc = false;
  bool mixed = false;   // DGX_MIXED: FP32 MFMA contractions (dgx_mixed3d.hip)
  double gamma = 1.4, vis = 0.0, Pr = 0.72;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  dgx::HostBasis hb;
  std::vector<double> coords, metric, J;
  std::vector<double> gm_host;   // curved: host copy of the nodal metrics d_gm
  // device
  double *d_tab = nullptr, *d_metric = nullptr, *d_fnorm = nullptr, *d_fjr = nullptr;
  int *d_nbr = nullptr, *d_nside = nullptr, *d_flag = nullptr;
  uint8_t *d_permid = nullptr, *d_master = nullptr;
  uint16_t *d_perm = nullptr;
  double *U = nullptr, *Ut = nullptr, *R0 = nullptr, *R1 = nullptr, *T = nullptr;
  double *G = nullptr, *V = nullptr, *TV = nullptr, *nodal = nullptr, *T2 = nullptr;
  float *Vf = nullptr, *TVf = nullptr;   // viscous terms of the mixed path
  bool fused = false;   // hex fused stage kernels available for this order
  bool curved = false;  // curved hexes (dgx_curved3d.hip): nodal metrics, mass-matrix projection
  double *d_gm = nullptr, *d_fgeo = nullptr, *d_jav = nullptr;
  double *d_pre = nullptr;  // curved hexes: [ne][3][L][L] preconditioner of the mass solve
  int *d_iters = nullptr;
  double time = 0.0;       // solution time: boundary and source functions are evaluated at it
  double t_step = 0.0;     // time at the start of the current step
  int next_stage = 0, stage_int = 0;   // dgx_stage: the next stage and its integrator
  dgx::FnArgs fn;          // boundary and source functions (device programs and coordinates)
  int *d_fcode = nullptr;
  double *d_fconst = nullptr, *d_fxy = nullptr, *d_xn = nullptr;
  double *Bal = nullptr;   // shallow water: the first residual (balance term)
  bool bulk = false;       // channel bulk forcing (Channel::holdBulkMassFlow) after every residual
  double *d_J = nullptr;   // [ne] element Jacobians (bulk forcing)
  double *d_cgeo = nullptr;  // 2D: CFL geometry (x_r, x_s, y_r, y_s, J) per element or, curved, per node
  double *d_chol = nullptr;  // curved quads: packed Cholesky factors of the element mass matrices
  double *d_part = nullptr;  // 2D: CFL scratch, ne per-element maxima and the domain maximum
  int bal = 0;             // 0 no balance term, 1 store at the next residual, 2 subtract
  // variable order (quads, dgx_varp2d.hip): per-element orders, elements
  // grouped by order (d_velems, group g at vgrp_off[g] with vgrp_n[g] elements
  // of order vgrp_p[g]), Gauss points and flips of every side, side tables
  bool varp = false;
  std::vector<int> porder, vgrp_p, vgrp_off, vgrp_n;
  std::vector<double> vgrp_dz;   // [group][kMaxOrder + 2] GLL node spacings (CFL)
  int *d_velems = nullptr, *d_vqs = nullptr;
  uint8_t *d_vflip = nullptr;
  double *d_vtab = nullptr;
  dgx::Tab tab;
  dgx::Geo geo;
  dgx::Phys phys;
  long long stages = 0;
  // per-kernel timing (dgx_profile): event pairs and the kernel kind of each
  bool profiling = false;
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_kind;

  int device = 0;           // the GPU the domain's buffers live on

  size_t state_len() const { return (size_t)nf * ne * nm; }
};

namespace dgx {
// CFL evaluation of a 2D domain (dgx_api.cpp): the largest idt of
// Euler::set_time_step with `part` (ne + 1 doubles of device memory) as the
// kernels' scratch, waiting for the domain's stream; and the step rule.
int cfl_max(dgx_domain *d, double *part, double *m);
int cfl_finish(double dt, double CFLmax, double dtmax, double m, double *cfl, double *dt_new);
}  // namespace dgx
