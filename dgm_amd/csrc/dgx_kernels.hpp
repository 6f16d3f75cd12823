// Launch interface between the host engine (dgx_api.cpp) and the templated
// kernels (dgx_kernels2d.hip, dgx_kernels3d.hip, dgx_stage3d.hip, dgx_mixed3d.hip).
#pragma once
#include <string>

#include "dgx_device.hpp"

namespace dgx {

enum Op {
  OP_TRACES = 0,    // modal U -> face traces T
  OP_RHS = 1,       // U, T [, V, TV] -> R = -dU/dt
  OP_GRAD = 2,      // U, T -> lifted gradients G          (Navier-Stokes)
  OP_VISC = 3,      // U, G -> viscous volume V, traces TV  (Navier-Stokes)
  OP_FORWARD = 4,   // nodal X -> modal Y
  OP_BACKWARD = 5,  // modal X -> nodal Y
  OP_STEP = 6       // 2D: U, T [, V, TV] -> one LSRK stage in place: Ut = ca Ut - dt R, U += cb Ut
};

struct LaunchArgs {
  int op = 0, p = 0, ne = 0, nfield = 0;
  bool ns = false;
  const double *U = nullptr;
  double *T = nullptr, *R = nullptr, *G = nullptr, *V = nullptr, *TV = nullptr;
  const double *X = nullptr;
  double *Y = nullptr;
  int *flag = nullptr;
  int eqn = 0;                   // eqntype (shallow water: 30)
  double *Bal = nullptr;         // shallow water: the stored first residual
  int bal = 0;                   // 0 no balance term, 1 store the residual in Bal, 2 subtract Bal
  double *Us = nullptr, *Ut = nullptr;   // OP_STEP: the state (updated in place) and the LSRK register
  double ca = 0.0, cb = 0.0, dt = 0.0;   // OP_STEP: LSRK coefficients and time step
  double *Tn = nullptr;                  // OP_STEP: if set, the face traces of the new state go here
  bool mixed = false;            // quads, inviscid: FP32 LDS tensors and contractions (traces, residual, step)
  Tab tab;
  Geo geo;
  Phys phys;
  hipStream_t stream = nullptr;
};

// One member request of a batched 2D group (serving): its modal state U,
// LSRK register Ut, face traces T, balance term, error flag, geometry and
// physics; off = its first work item in the group's flat element list.
struct Batch2 {
  double *U = nullptr, *Ut = nullptr, *T = nullptr, *T2 = nullptr, *Bal = nullptr;
  int *flag = nullptr;
  Geo geo{};
  Phys phys{};
  int ne = 0, off = 0, bal = 0;
  double dt = 0.0;
};

// Workgroups of a batched launch: enough for every CU, each walking the work list.
constexpr int kBatchGrid = 8192;

struct BatchLaunch {
  int op = 0;                    // 0: face traces, 1: residual + LSRK14 register update (+ next traces),
                                 // 2: a persistent slice of nsteps LSRK14 steps (cooperative launch)
  int p = 0, nreq = 0, nwork = 0;
  bool shallow = false;          // shallow water (3 fields) or Euler (4 fields)
  const Batch2 *b = nullptr;     // device array of nreq members
  double A = 0.0, B = 0.0;       // LSRK14 coefficients of the stage
  int odd = 0, next = 0;         // op 1: read the traces of T2 and write into T (odd) or the reverse;
                                 // next: write the next stage's traces at all
  int nsteps = 0;                // op 2: steps of the slice
  double lsrkA[14] = {}, lsrkB[14] = {};   // op 2: the LSRK14 coefficients of every stage
  int grid = 0;                  // op 2, out: workgroups launched
  bool mixed = false;            // FP32 LDS tensors and contractions
  hipStream_t stream = nullptr;
};

hipError_t launch_batch2(BatchLaunch &a);

// Fill each translation unit's constant-memory basis tables (kTabTotal doubles).
hipError_t upload_tables_2d(const double *h);
hipError_t upload_tables_3d(const double *h);
hipError_t upload_tables_s3d(const double *h);
hipError_t upload_tables_m3d(const double *h);

hipError_t launch_2d(const LaunchArgs &a);
hipError_t launch_3d(const LaunchArgs &a);
int lds_bytes2(int p, int which);
int lds_bytes3(int p, int which);

// Fused hex stage kernels (dgx_stage3d.hip), p <= 9.
struct StageLaunch {
  int op = 1;                    // 0: BR1 viscous terms (k_vterm3), 1: stage (k_stage3)
  int p = 0, ne = 0;
  bool ns = false, update = false;
  const double *T = nullptr, *V = nullptr, *TV = nullptr;
  double *U = nullptr, *Ut = nullptr, *R = nullptr, *Tn = nullptr, *Vw = nullptr, *TVw = nullptr;
  const float *Vf = nullptr, *TVf = nullptr;   // mixed precision: FP32 viscous terms
  float *Vfw = nullptr, *TVfw = nullptr;
  double ca = 0.0, cb = 0.0, dt = 0.0;
  int *flag = nullptr;
  Tab tab;
  Geo geo;
  Phys phys;
  hipStream_t stream = nullptr;
};
bool stage3_supported(int p);
hipError_t launch_stage3(const StageLaunch &s);

// Mixed-precision fused hex stage kernels (dgx_mixed3d.hip), p <= 8.
bool mixed3_supported(int p);
hipError_t launch_mixed3(const StageLaunch &s);

// Curved hexes, Euler3d, p <= 8 (dgx_curved3d.hip): residual R = -dU/dt with
// metrics at the nodes and the element mass matrix.
// Boundary and source functions of (x, y, t) as postfix programs (dgx_function):
// bc[c] .. bc[c+1] and src[c] .. src[c+1] bound component c's instructions in code.
struct FnArgs {
  const int *code = nullptr;
  const double *consts = nullptr;
  int bc[4] = {0, 0, 0, 0}, src[4] = {0, 0, 0, 0};
  int has_bc = 0, has_src = 0;
  const double *fxy = nullptr;   // [ne][4][QF][2] Gauss-point coordinates of every side
  const double *xn = nullptr;    // [2][ne][Q^2] nodal coordinates
};

struct CurvedArgs {
  const double *U = nullptr, *T = nullptr;   // modal state and its face traces
  double *R = nullptr;                       // modal residual
  const double *gm = nullptr;                // [ne][10][Q^3]: wJ dxi_d/dx_x (d-major), wJ
  const double *fgeo = nullptr;              // [ne][6][QF^2][4]: normal, J_side/J_vol
  const double *jav = nullptr;               // [ne]: mean J
  const double *pre = nullptr;               // hexes: [ne][3][L][L] inverses of the 1D mass matrices of a
                                             // separable fit of J, the mass solve's preconditioner
  const double *chol = nullptr;              // quads: lower Cholesky factors of the element mass
                                             // matrices, rows packed, [ne][L2 (L2 + 1) / 2]
  Geo geo;
  Phys phys;
  int ne = 0, p = 0;
  int *flag = nullptr, *iters = nullptr;
  bool project = false;                      // U nodal [5][ne][Q^3] -> R modal: M^-1 B^T (wJ U)
  bool shallow = false;                      // quads: shallow water (3 fields) instead of Euler
  // quads, Navier-Stokes (BR1): op 1 writes the lifted, projected gradients of
  // (u, v, e) into G [ne][6][L^2]; op 2 the weighted nodal viscous divergence
  // into V [ne][3][Q^2] and each side's normal viscous flux into TV
  // [ne][4][3][QF]; op 0 (the residual) then adds both.
  bool ns = false;
  int op = 0;
  double *G = nullptr, *V = nullptr, *TV = nullptr;
  // shallow water: functions at time t; the balance term (Bal: bal 1 stores
  // the residual and returns R - R = 0, bal 2 returns R - Bal)
  FnArgs fn;
  double t = 0.0;
  double *Bal = nullptr;
  int bal = 0;
  // hexes: one LSRK stage in place instead of R (update): Ut = ca Ut - dt R,
  // Us += cb Ut, with Us the modal state U points to
  bool update = false;
  bool mixed = false;                        // hexes: FP32 residual arithmetic before the FP64 mass solve
  double *Us = nullptr, *Ut = nullptr;
  double ca = 0.0, cb = 0.0, dt = 0.0;
  hipStream_t stream = nullptr;
};
// Quads use the same arguments with gm [ne][5][Q^2] (wJ dxi_d/dx_x, wJ),
// fgeo [ne][4][QF][4] (normal x, y, 0, J_side/J_vol) and NF = 3 or 4 fields.
hipError_t launch_curved2(const CurvedArgs &a);
hipError_t upload_tables_c2d(const double *h);
bool curved3_supported(int p);
bool curved3_ns_supported(int p);   // Navier_Stokes3d (BR1) on curved hexes
hipError_t launch_curved3(const CurvedArgs &a);
hipError_t upload_tables_c3d(const double *h);

// Linear hyperbolic system on quads, dU/dt + Ax dU/dx + Ay dU/dy = s (the
// linearized Euler equations about a uniform mean flow), fused with one
// LSRK14 stage; or, with the adjoint system (-Ax^T, -Ay^T), one reversed
// stage of its discrete adjoint (dgx_kernels2d.hip: k_lin2).
struct LinArgs {
  const double *X = nullptr;     // modal input [4][ne][L2]: U (forward) or Ubar_t (adjoint)
  const double *T = nullptr;     // face traces of X
  double *Y = nullptr;           // forward: U += cb Ut;  adjoint: Ubar += -dt J Fadj(X / J)
  double *Yt = nullptr;          // forward: Ut = ca Ut - dt R;  adjoint: Ubar_t = ca X (may alias X)
  const double *theta = nullptr; // forward: source field, one value per element (energy equation)
  double *gtheta = nullptr;      // adjoint: dJ/dtheta accumulator per element
  double Ax[16], Ay[16];         // row-major Jacobians (adjoint: -Ax^T, -Ay^T)
  double u0 = 0, v0 = 0, c0 = 0; // Lax-Friedrichs speed |u0 nx + v0 ny| + c0
  double ca = 0, cb = 0, dt = 0, g = 0;   // LSRK coefficients, source amplitude g(t_stage)
  int ne = 0, p = 0;
  bool adjoint = false;
  bool mixed = false;            // FP32 LDS tensors and contractions, FP64 state and registers
  Tab tab;
  Geo geo;
  hipStream_t stream = nullptr;
};
hipError_t launch_lin2(const LinArgs &a);

// Backward kernels of the 2D Euler residual R = -dU/dt on periodic affine
// quads with the Lax-Friedrichs flux (dgx_adjoint2d.hip).  op 0: adjoint
// traces Lt [ne][4][4][QF] of g = lam (first formed in place as
// lam = lam + B lamU when lamU is set); op 1: out = scale (dR/dU(U))^T g, added
// to out when accumulate, with T the face traces of U and Lt those of g, and
// with `stage` lam = A g afterwards.  A reversed LSRK stage is op 0 and op 1
// with lamU = out = the state's adjoint, B = B_s, A = A_s, scale = -dt.
struct AdjArgs {
  int op = 0, p = 0, ne = 0;
  const double *U = nullptr, *T = nullptr, *lamU = nullptr;
  double *lam = nullptr, *Lt = nullptr, *out = nullptr;
  double scale = 1.0, gamma = 1.4, A = 0.0, B = 0.0;
  int flux = 0;                              // face flux: 0 Lax-Friedrichs, 1 Roe
  bool accumulate = false, stage = false;
  Geo geo;
  hipStream_t stream = nullptr;
};
hipError_t launch_adj2(const AdjArgs &a);
// Pressure (gamma-1)(E - |m|^2 / (2 rho)) of the state at receiver points,
// y_r = p(sum_m w[r][m] U[.][e_r][m]); and the adjoint of the objective
// 1/2 dt sum_r (y_r - d_r)^2 into lam, with J += the objective.  The adjoint
// takes coef [nrec][5] as scratch and the receivers grouped by element:
// group g holds element grp_e[g] and receivers grp_r[grp_ptr[g] ..
// grp_ptr[g+1]) in increasing order.
struct ObsArgs {
  const double *U = nullptr;
  const int *rec_e = nullptr;
  const double *rec_w = nullptr;
  int nrec = 0, ne = 0, L2 = 0;
  double gamma = 1.4;
  double *y = nullptr;                              // forward: y[nrec]
  double *lam = nullptr, *J = nullptr, *coef = nullptr;
  const double *d = nullptr;                        // adjoint: data d[nrec]
  double dt = 0.0;
  const int *grp_e = nullptr, *grp_ptr = nullptr, *grp_r = nullptr;
  int ngrp = 0;
  hipStream_t stream = nullptr;
};
hipError_t launch_pobs2(const ObsArgs &a);
hipError_t launch_pobs2_adj(const ObsArgs &a);
hipError_t upload_tables_a2d(const double *h);

// Receiver observations y_r = sum_f c_f sum_m w[r][m] U[f][e_r][m] into y[r]
// (observe), or the adjoint of the objective 1/2 dt sum_r (y_r - d_r)^2:
// Ubar[f][e_r][m] += dt (y_r - d_r) c_f w[r][m] and *J += 1/2 dt sum_r (y_r - d_r)^2.
hipError_t launch_observe(const double *U, const int *rec_e, const double *rec_w, const double *cobs, int nrec,
                          int ne, int L2, double *y, hipStream_t s);
hipError_t launch_observe_adjoint(const double *U, double *Ubar, const int *rec_e, const double *rec_w,
                                  const double *cobs, int nrec, int ne, int L2, const double *d, double dt,
                                  double *J, hipStream_t s);

// CFL metric of a 2D domain (dgx_cfl.hip): part[e] = max over element e's GLL
// nodes of idt (Euler::set_time_step), out[0] = max over elements.  cgeo holds
// (x_r, x_s, y_r, y_s, J) per element [ne][5] (affine) or per node [ne][5][Q^2]
// (curved); dz[i] the GLL node spacings.
struct CflArgs {
  const double *U = nullptr;     // modal state [nf][ne][L^2]
  const double *cgeo = nullptr;
  double *part = nullptr;        // [ne] scratch
  double *out = nullptr;         // [1]
  double dz[kMaxOrder + 2] = {};
  double gamma = 1.4;
  int ne = 0, p = 0;
  bool shallow = false, curved = false;
  // variable order: the group of elements of order p (elems[nel]) in slots of
  // NM modes, and no reduction (launch_cfl_max1 follows the last group)
  const int *elems = nullptr;
  int nel = 0, NM = 0;
  hipStream_t stream = nullptr;
};
hipError_t launch_cfl2(const CflArgs &a);
hipError_t launch_cfl_max1(const double *part, int n, double *out, hipStream_t s);
hipError_t upload_tables_cfl(const double *h);

// Variable-order quads (dgx_varp2d.hip), 2D Euler on affine quads, elements
// launched in groups of one order.  Side tables for element order P and n side
// points (P + 1 <= n <= kMaxOrder + 1): Bgv [n][L] (modal -> the side's Gauss
// points) and PWv [L][n] (the lift of those points into the element's modes),
// each in a block of 121 doubles at vp_block(P, n, 0 / 1).
__host__ __device__ constexpr int vp_block(int P, int n, int which) { return ((P * 12 + n) * 2 + which) * 121; }
constexpr int kVpTab = vp_block(kMaxOrder + 1, 0, 0);
struct VarpArgs {
  int op = 0, p = 0;             // OP_TRACES, OP_RHS, OP_FORWARD, OP_BACKWARD; the group's order
  const int *elems = nullptr;    // the group's elements
  int nel = 0, ne = 0;
  int NM = 0, QM = 0, QFM = 0;   // padded modes and nodes per element, points per side slot
  const double *U = nullptr;     // modal [4][ne][NM]
  double *T = nullptr;           // traces [ne][4][4][QFM]
  double *R = nullptr;           // modal residual [4][ne][NM]
  const double *X = nullptr;     // transforms: input
  double *Y = nullptr;           // transforms: output
  const int *qs = nullptr;       // [ne][4] Gauss points of each side
  const uint8_t *flip = nullptr; // [ne][4] 1: the neighbour's points run reversed
  const double *vtab = nullptr;  // side tables (kVpTab doubles)
  int *flag = nullptr;
  Geo geo;
  Phys phys;
  hipStream_t stream = nullptr;
};
hipError_t launch_varp2(const VarpArgs &a);
hipError_t upload_tables_v2d(const double *h);

// Error message of the calling thread's last failure (dgx_api.cpp).
int set_error(int code, const std::string &msg);

// Flat-array stage updates (dgx_blas.hip)
hipError_t blas_axpy(double a, const double *x, const double *y, double *z, size_t n, hipStream_t s);
hipError_t blas_axpby(double a, const double *x, double b, const double *y, double *z, size_t n,
                      hipStream_t s);
hipError_t blas_scale(double a, double *x, size_t n, hipStream_t s);
// Channel bulk forcing on a modal residual [nf][ne][nm] of affine elements with Jacobians J[ne].
hipError_t bulk_hold(double *R, const double *J, int ne, int nm, int nf, hipStream_t s);

}  // namespace dgx
