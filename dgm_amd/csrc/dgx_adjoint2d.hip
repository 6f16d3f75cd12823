// Backward kernels of the 2D Euler residual (quads, Lax-Friedrichs, affine,
// periodic): the vector-Jacobian product out = (dR/dU)^T lam of the residual
// R = -dU/dt that k_traces2 + k_rhs2 compute (Euler::convection,
// convective_flux, lift and project, src/Euler.cpp:296-757,
// src/Quad.cpp:1971-2038), and the pressure observation at receiver points
// with its adjoint.  They drive the discrete adjoint of LSRK14 time stepping
// (LSRK14::adjoint_advance, src/TimeInt.cpp:2064-2092) for gradients with
// respect to the initial state (dgx_euler_adjoint.cpp).
//
// The forward residual of element e, in modal coordinates, is
//   R = (QD (x) BW) N1 + (BW (x) QD) N2 + sum_s Bend_s (x) (PW g_s),
//   N1 = F(u).(rx, ry), N2 = F(u).(sx, sy) at the GLL nodes, u = (BT (x) BT) U,
//   g_s = J_side/J_vol (F^(u-, u+, n) - F(u-).n) at the Gauss points of side s,
//   u- = the element's trace (Bend, Bg), u+ = the neighbour's.
// Its transpose, applied to lam:
//   volume  N1bar = (QD (x) BW)^T lam, N2bar = (BW (x) QD)^T lam, then
//           ubar = A(u; rx, ry)^T N1bar + A(u; sx, sy)^T N2bar at every node
//           (A(u; k) = d(F(u).k)/du) and Ubar += (BT (x) BT)^T ubar;
//   faces   gbar_s = PW^T (the Bend collapse of lam): the adjoint traces, which
//           k_adjtraces2 evaluates for every element first, since each side's
//           trace enters both its own and its neighbour's face flux;
//           u-bar = jr (dF^/du- - A(u-; n))^T gbar_own + jr' (dF^'/du+')^T gbar_nb
//           with the neighbour's flux F^' = F^(u+, u-, n'), then the transpose
//           of the trace evaluation into the modes.
// The Lax-Friedrichs speed max(|u.n| + c) is differentiated on the side that
// attains it.
#include <vector>

#include "dgx_kernels.hpp"

namespace dgx {

// A(U; k)^T w for the 2D Euler flux F(U).k with an unnormalized vector k:
// F.k = (rho vn, m vn + p kx, n vn + p ky, (E + p) vn), vn = (m kx + n ky) / rho.
__device__ __forceinline__ void euler_jac_t(const double U[4], double kx, double ky, double gm1, const double w[4],
                                            double out[4]) {
  const double ir = 1.0 / U[0], u = U[1] * ir, v = U[2] * ir;
  const double q2 = u * u + v * v;
  const double p = gm1 * (U[3] - 0.5 * U[0] * q2);
  const double vn = u * kx + v * ky, H = (U[3] + p) * ir;
  const double dp[4] = {0.5 * gm1 * q2, -gm1 * u, -gm1 * v, gm1};
  // rows of A: r0 = (0, kx, ky, 0)
  //            r1 = (-u vn, u kx + vn, u ky, 0) + kx dp
  //            r2 = (-v vn, v kx, v ky + vn, 0) + ky dp
  //            r3 = (vn (0.5 gm1 q2 - H), H kx - gm1 u vn, H ky - gm1 v vn, (gm1 + 1) vn)
  out[0] = w[1] * (-u * vn + kx * dp[0]) + w[2] * (-v * vn + ky * dp[0]) + w[3] * (vn * (0.5 * gm1 * q2 - H));
  out[1] = w[0] * kx + w[1] * (u * kx + vn + kx * dp[1]) + w[2] * (v * kx + ky * dp[1]) + w[3] * (H * kx - gm1 * u * vn);
  out[2] = w[0] * ky + w[1] * (u * ky + kx * dp[2]) + w[2] * (v * ky + vn + ky * dp[2]) + w[3] * (H * ky - gm1 * v * vn);
  out[3] = w[1] * kx * dp[3] + w[2] * ky * dp[3] + w[3] * ((gm1 + 1.0) * vn);
}

// gradient of the wave speed |u.n| + c with respect to U
__device__ __forceinline__ void speed_grad(const double U[4], double nx, double ny, double gamma, double g[4]) {
  const double gm1 = gamma - 1.0, ir = 1.0 / U[0], u = U[1] * ir, v = U[2] * ir;
  const double p = gm1 * (U[3] - 0.5 * (U[1] * u + U[2] * v));
  const double un = u * nx + v * ny, c = sqrt(gamma * p * ir), s = un >= 0.0 ? 1.0 : -1.0;
  // d(un)/dU = (-un/rho, nx/rho, ny/rho, 0); c = sqrt(gamma p / rho):
  // dc/dU = gamma / (2 c rho) (dp/dU - p/rho e_rho)
  const double dp[4] = {0.5 * gm1 * (u * u + v * v), -gm1 * u, -gm1 * v, gm1};
  const double f = gamma / (2.0 * c) * ir;
  g[0] = s * (-un * ir) + f * (dp[0] - p * ir);
  g[1] = s * (nx * ir) + f * dp[1];
  g[2] = s * (ny * ir) + f * dp[2];
  g[3] = f * dp[3];
}

// Transposed Jacobians of the Lax-Friedrichs flux F^(UL, UR, n) =
// 1/2 (F(UL).n + F(UR).n + lam (UL - UR)), lam = max(lam(UL), lam(UR)),
// applied to w: aL = (dF^/dUL)^T w, aR = (dF^/dUR)^T w.
__device__ __forceinline__ void lf_jac_t(const double UL[4], const double UR[4], double nx, double ny, double gamma,
                                         const double w[4], double aL[4], double aR[4]) {
  const double gm1 = gamma - 1.0;
  double FL[4], FR[4], lL, lR;
  bool bad;
  euler_normal_flux2(UL, nx, ny, gm1, FL, lL, bad);
  euler_normal_flux2(UR, nx, ny, gm1, FR, lR, bad);
  const double lam = fmax(lL, lR);
  double jl[4], jr[4];
  euler_jac_t(UL, nx, ny, gm1, w, jl);
  euler_jac_t(UR, nx, ny, gm1, w, jr);
  // d lam / dU on the attaining side, times (UL - UR) . w
  double dw = 0.0;
#pragma unroll
  for (int v = 0; v < 4; ++v) dw += (UL[v] - UR[v]) * w[v];
  double gl[4] = {0, 0, 0, 0}, gr[4] = {0, 0, 0, 0};
  if (lL >= lR) speed_grad(UL, nx, ny, gamma, gl);
  else speed_grad(UR, nx, ny, gamma, gr);
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    aL[v] = 0.5 * (jl[v] + lam * w[v] + dw * gl[v]);
    aR[v] = 0.5 * (jr[v] - lam * w[v] + dw * gr[v]);
  }
}

// Forward-mode dual numbers: value and one directional derivative.  The
// Roe flux's transposed Jacobian is assembled from its 8 directional
// derivatives (one per input of UL and UR), evaluated by the same formula as
// the forward flux; min(x, 0) takes the side of its argument's value.
struct Dual {
  double v, d;
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator/(double a, Dual b) { const double q = a / b.v; return {q, -q * b.d / b.v}; }
__device__ __forceinline__ Dual dsqrt(Dual a) { const double r = sqrt(a.v); return {r, 0.5 * a.d / r}; }
__device__ __forceinline__ Dual dmin0(Dual a) { return a.v < 0.0 ? a : Dual{0.0, 0.0}; }

// The Roe face flux F^(UL, UR, n) = F(UL).n + G(UL, UR) of roe_flux2 (the same
// rotated-frame Roe average, wave strengths and upwinded waves), in duals.
__device__ __forceinline__ void roe_flux_dual(const Dual UL[4], const Dual UR[4], double nx, double ny,
                                              double gamma, Dual Fh[4]) {
  const double gm1 = gamma - 1.0;
  const Dual rl = UL[0], rul = UL[1] * nx + UL[2] * ny, rvl = UL[2] * nx - UL[1] * ny, rEl = UL[3];
  const Dual rr = UR[0], rur = UR[1] * nx + UR[2] * ny, rvr = UR[2] * nx - UR[1] * ny, rEr = UR[3];
  const Dual ul = rul / rl, vl = rvl / rl, ur = rur / rr, vr = rvr / rr;
  const Dual pl = gm1 * (rEl - 0.5 * (rul * ul + rvl * vl));
  const Dual pr = gm1 * (rEr - 0.5 * (rur * ur + rvr * vr));
  const Dual Hl = (rEl + pl) / rl, Hr = (rEr + pr) / rr;
  const Dual R = dsqrt(rr / rl), iRp1 = 1.0 / (1.0 + R);
  const Dual u = (ul + R * ur) * iRp1, v = (vl + R * vr) * iRp1, H = (Hl + R * Hr) * iRp1;
  const Dual a = dsqrt(gm1 * (H - 0.5 * (u * u + v * v))), ainv = 1.0 / a;
  const Dual c1inv = a * a * (1.0 / gm1);
  Dual al[4], lm[4], G[4];
  al[2] = (rvr - rvl) - v * (rr - rl);
  al[1] = gm1 * ainv * ainv * ((rr - rl) * (H - u * u) + u * (rur - rul) -
                               ((rEr - rEl) - ((rvr - rvl) - v * (rr - rl)) * v));
  al[0] = 0.5 * ainv * ((rr - rl) * (u + a) - (rur - rul) - a * al[1]);
  al[3] = (rr - rl) - (al[0] + al[1]);
  lm[0] = dmin0(u - a); lm[1] = dmin0(u); lm[2] = lm[1]; lm[3] = dmin0(u + a);
  G[0] = lm[0] * al[0] + lm[1] * al[1] + lm[3] * al[3];
  G[1] = lm[0] * (u - a) * al[0] + lm[1] * u * al[1] + lm[3] * (u + a) * al[3];
  G[2] = lm[0] * v * al[0] + lm[1] * v * al[1] + lm[2] * al[2] + lm[3] * v * al[3];
  G[3] = lm[0] * (H - u * a) * al[0] + lm[1] * (H - c1inv) * al[1] + lm[2] * v * al[2] +
         lm[3] * (H + u * a) * al[3];
  // F(UL).n
  const Dual ir = 1.0 / UL[0], uu = UL[1] * ir, vv = UL[2] * ir;
  const Dual p = gm1 * (UL[3] - 0.5 * (UL[1] * uu + UL[2] * vv));
  const Dual un = uu * nx + vv * ny;
  Fh[0] = UL[0] * un + G[0];
  Fh[1] = UL[1] * un + p * nx + (G[1] * nx - G[2] * ny);
  Fh[2] = UL[2] * un + p * ny + (G[1] * ny + G[2] * nx);
  Fh[3] = (UL[3] + p) * un + G[3];
}

// Transposed Jacobians of the Roe flux applied to w: aL = (dF^/dUL)^T w,
// aR = (dF^/dUR)^T w, one directional derivative per input.
__device__ __forceinline__ void roe_jac_t(const double UL[4], const double UR[4], double nx, double ny,
                                          double gamma, const double w[4], double aL[4], double aR[4]) {
#pragma unroll 1
  for (int j = 0; j < 8; ++j) {
    Dual L[4], Rr[4], F[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      L[v] = {UL[v], j == v ? 1.0 : 0.0};
      Rr[v] = {UR[v], j == 4 + v ? 1.0 : 0.0};
    }
    roe_flux_dual(L, Rr, nx, ny, gamma, F);
    double t = 0.0;
#pragma unroll
    for (int v = 0; v < 4; ++v) t = fma(w[v], F[v].d, t);
    if (j < 4) aL[j] = t;
    else aR[j - 4] = t;
  }
}

// Transposed basis tables of every order (this file's constant memory,
// derived from the shared tables at upload): BTt [L][Q] = BT^T, BWt [Q][L] =
// BW^T, QDt [Q][L] = QD^T, PWt [QF][L] = PW^T and Bgt [L][QF] = Bg^T, so that
// every transposed contraction runs through the register-blocked `contract`.
constexpr int adj_size(int p) { return 3 * (p + 1) * (p + 2) + 2 * (p + 1) * (p + 1); }
constexpr int adj_offset(int p) {
  int s = 0;
  for (int k = 1; k < p; ++k) s += adj_size(k);
  return s;
}
constexpr int kAdjTotal = adj_offset(kMaxOrder + 1);
static __constant__ double c_adj[kAdjTotal];

template <int P> struct AT {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1, b = adj_offset(P);
  __device__ static __forceinline__ const double *BTt() { return c_adj + b; }
  __device__ static __forceinline__ const double *BWt() { return c_adj + b + L * Q; }
  __device__ static __forceinline__ const double *QDt() { return c_adj + b + 2 * L * Q; }
  __device__ static __forceinline__ const double *PWt() { return c_adj + b + 3 * L * Q; }
  __device__ static __forceinline__ const double *Bgt() { return c_adj + b + 3 * L * Q + QF * L; }
};

// Adjoint traces of g: Lt[e][s][v][g] = sum_l PW[l][g] (the Bend collapse of g
// on side s)[v][l], the transpose of the lift's face integral: the same
// pattern as the face traces, with PW^T in place of Bg.  With a.lamU set, g is
// first formed in place from the stage recurrence, g = lam + B lamU (lam's
// own element only).
template <int P>
__global__ void __launch_bounds__(64) k_adjtraces2(AdjArgs a) {
  constexpr int L = P + 1, QF = P + 1, L2 = L * L, NT = 64, NF = 4;
  __shared__ double A[NF * L2], Kb[NF * 2 * L], Ka[NF * 2 * L], O[NF * 2 * QF];
  const int e = blockIdx.x, tid = threadIdx.x, ne = a.ne;
  for (int i = tid; i < NF * L2; i += NT) {
    const int v = i / L2;
    const size_t o = ((size_t)v * ne + e) * L2 + (i - v * L2);
    double g = a.lam[o];
    if (a.lamU) {
      g = fma(a.B, a.lamU[o], g);   // as the axpy lam = B lamU + lam
      a.lam[o] = g;
    }
    A[i] = g;
  }
  __syncthreads();
  double *Le = a.Lt + (size_t)e * 4 * NF * QF;
  contract<NF, L, L, 2, 2, NT>(CT<P>::Bend(), A, Kb, tid);   // [NF][L][2]: edges 0 / 2
  contract<NF, L, L, 1, 2, NT>(CT<P>::Bend(), A, Ka, tid);   // [NF][2][L]: edges 3 / 1
  __syncthreads();
  contract<NF, L, 2, 1, QF, NT>(AT<P>::PWt(), Kb, O, tid);   // [NF][QF][2]
  __syncthreads();
  for (int i = tid; i < NF * 2 * QF; i += NT) {
    const int v = i / (2 * QF), r = i - v * 2 * QF, g = r >> 1, end = r & 1;
    Le[((end ? 2 : 0) * NF + v) * QF + g] = O[i];
  }
  __syncthreads();
  contract<NF, 2, L, 2, QF, NT>(AT<P>::PWt(), Ka, O, tid);   // [NF][2][QF]
  __syncthreads();
  for (int i = tid; i < NF * 2 * QF; i += NT) {
    const int v = i / (2 * QF), r = i - v * 2 * QF, end = r / QF, g = r - end * QF;
    Le[((end ? 1 : 3) * NF + v) * QF + g] = O[i];
  }
}

// out = scale (dR/dU)^T g for element e (see the file comment); T: traces of
// U, Lt: adjoint traces of g (k_adjtraces2).  With a.stage, g's own element
// is then replaced by A g (the stage recurrence's lam = A_s g).
template <int P>
__global__ void __launch_bounds__(64) k_vjp2(AdjArgs a) {
  constexpr int L = P + 1, Q = P + 2, QF = P + 1, L2 = L * L, Q2 = Q * Q, NT = 64, NF = 4;
  constexpr int NPT = (Q2 + NT - 1) / NT;
  __shared__ double M[NF * L2], X[NF * Q * L], X2[NF * Q * L], N1[NF * Q2], N2[NF * Q2], gT[NF * 4 * QF],
      Kb[NF * 4 * L];
  const int e = blockIdx.x, tid = threadIdx.x, ne = a.ne;
  const Geo &geo = a.geo;
  const double gamma = a.gamma, gm1 = gamma - 1.0;
  // nodal state u = (BT (x) BT) U
  for (int i = tid; i < NF * L2; i += NT) {
    const int v = i / L2;
    M[i] = a.U[((size_t)v * ne + e) * L2 + (i - v * L2)];
  }
  __syncthreads();
  contract<NF, L, L, 1, Q, NT>(CT<P>::BT(), M, X, tid);   // [NF][Q][L]
  __syncthreads();
  contract<NF, Q, L, 2, Q, NT>(CT<P>::BT(), X, N1, tid);  // [NF][Q][Q]
  __syncthreads();
  double un[NPT][NF];
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int k = tid + r * NT;
#pragma unroll
    for (int v = 0; v < NF; ++v) un[r][v] = k < Q2 ? N1[v * Q2 + k] : 1.0;
  }
  // g; with a.stage the element's lam becomes A g
  for (int i = tid; i < NF * L2; i += NT) {
    const int v = i / L2;
    const size_t o = ((size_t)v * ne + e) * L2 + (i - v * L2);
    const double g = a.lam[o];
    M[i] = g;
    if (a.stage) a.lam[o] = a.A * g;
  }
  __syncthreads();
  // volume transpose: N1bar = (QD^T (x) BW^T) g, N2bar = (BW^T (x) QD^T) g
  contract<NF, L, L, 2, Q, NT>(AT<P>::BWt(), M, X, tid);    // [NF][L][Q]
  contract<NF, L, L, 2, Q, NT>(AT<P>::QDt(), M, X2, tid);   // [NF][L][Q]
  __syncthreads();
  contract<NF, L, Q, 1, Q, NT>(AT<P>::QDt(), X, N1, tid);   // [NF][Q][Q]
  contract<NF, L, Q, 1, Q, NT>(AT<P>::BWt(), X2, N2, tid);  // [NF][Q][Q]
  __syncthreads();
  // pointwise: ubar = A(u; rx, ry)^T N1bar + A(u; sx, sy)^T N2bar, into N1
  const double *mt = geo.metric + 9 * (size_t)e;
  const double rx = mt[0], ry = mt[1], sx = mt[3], sy = mt[4];
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int k = tid + r * NT;
    if (k >= Q2) continue;
    double w1[4], w2[4], o1[4], o2[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) { w1[v] = N1[v * Q2 + k]; w2[v] = N2[v * Q2 + k]; }
    euler_jac_t(un[r], rx, ry, gm1, w1, o1);
    euler_jac_t(un[r], sx, sy, gm1, w2, o2);
#pragma unroll
    for (int v = 0; v < 4; ++v) N1[v * Q2 + k] = o1[v] + o2[v];
  }
  // faces: adjoint of the own traces at the Gauss points
  for (int i = tid; i < 4 * QF; i += NT) {
    const int s = i / QF, g = i - s * QF;
    const int es = e * 4 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int gn = geo.perm[geo.permid[es] * QF + g];
    const double nx = geo.fnorm[3 * es], ny = geo.fnorm[3 * es + 1];
    const double mx = geo.fnorm[3 * (nb * 4 + ns)], my = geo.fnorm[3 * (nb * 4 + ns) + 1];
    const double jr = geo.fjr[es], jrn = geo.fjr[nb * 4 + ns];
    const double *tl = a.T + ((size_t)es) * NF * QF, *tr = a.T + ((size_t)nb * 4 + ns) * NF * QF;
    const double *ll = a.Lt + ((size_t)es) * NF * QF, *lr = a.Lt + ((size_t)nb * 4 + ns) * NF * QF;
    double UL[4], UR[4], wo[4], wn[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      UL[v] = tl[v * QF + g]; UR[v] = tr[v * QF + gn];
      wo[v] = jr * ll[v * QF + g]; wn[v] = jrn * lr[v * QF + gn];
    }
    double aL[4], aR[4], jf[4], bL[4], bR[4];
    if (a.flux == 1) {
      roe_jac_t(UL, UR, nx, ny, gamma, wo, aL, aR);  // own face flux F^(UL, UR, n)
      roe_jac_t(UR, UL, mx, my, gamma, wn, bL, bR);  // the neighbour's F^(UR, UL, n_nb)
    } else {
      lf_jac_t(UL, UR, nx, ny, gamma, wo, aL, aR);
      lf_jac_t(UR, UL, mx, my, gamma, wn, bL, bR);
    }
    euler_jac_t(UL, nx, ny, gm1, wo, jf);             // minus F(UL).n
#pragma unroll
    for (int v = 0; v < 4; ++v) gT[(v * 4 + s) * QF + g] = aL[v] - jf[v] + bR[v];
  }
  __syncthreads();
  // transpose of the trace evaluation, Kb = Bg^T gT; modal (BT^T (x) BT^T) ubar
  contract<NF, 4, QF, 2, L, NT>(AT<P>::Bgt(), gT, Kb, tid);     // [NF][4][L]
  contract<NF, Q, Q, 2, L, NT>(AT<P>::BTt(), N1, X, tid);   // [NF][Q][L]
  __syncthreads();
  contract<NF, Q, L, 1, L, NT>(AT<P>::BTt(), X, M, tid);    // [NF][L][L]
  __syncthreads();
  const double cw[4] = {1.0, 1.0, 1.0, 1.0};
  for (int t = tid; t < NF * L2; t += NT) {
    const int v = t / L2, k = t - v * L2, l = k / L, m = k - l * L;
    const double acc = M[t] + face_add2<P>(CT<P>::Bend(), Kb + v * 4 * L, l, m, cw);
    const size_t o = ((size_t)v * ne + e) * L2 + k;
    a.out[o] = a.accumulate ? fma(a.scale, acc, a.out[o]) : a.scale * acc;
  }
}

// The state at receiver r, u[f] = sum_m w[r][m] U[f][e_r][m], by one
// wavefront: lane-strided partial sums, then a butterfly over the 64 lanes
// (every lane ends with the same, deterministic value).
__device__ __forceinline__ void receiver_state(const double *__restrict__ U, int e, const double *__restrict__ w,
                                               int ne, int L2, double u[4]) {
  const int lane = threadIdx.x;
  for (int f = 0; f < 4; ++f) {
    const double *Uf = U + ((size_t)f * ne + e) * L2;
    double s = 0.0;
    for (int m = lane; m < L2; m += 64) s = fma(w[m], Uf[m], s);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    u[f] = s;
  }
}

// pressure (gamma - 1)(E - 1/2 (m u + n v)), u = m / rho, v = n / rho; the
// forward observation and its adjoint evaluate it alike, so that data the
// forward solve produced give a zero misfit
__device__ __forceinline__ double receiver_pressure(const double u[4], double gm1, double &vx, double &vy) {
#pragma clang fp contract(off)
  const double iu = 1.0 / u[0];
  vx = u[1] * iu;
  vy = u[2] * iu;
  return gm1 * (u[3] - 0.5 * (u[1] * vx + u[2] * vy));
}

// y[r] = the pressure of the state at receiver r; one wavefront per receiver
__global__ void __launch_bounds__(64) k_pobs2(ObsArgs a) {
  const int r = blockIdx.x;
  double u[4], vx, vy;
  receiver_state(a.U, a.rec_e[r], a.rec_w + (size_t)r * a.L2, a.ne, a.L2, u);
  if (threadIdx.x == 0) a.y[r] = receiver_pressure(u, a.gamma - 1.0, vx, vy);
}

// Adjoint of the objective 1/2 dt sum_r (p_r - d_r)^2 at one observation,
// first per receiver (one wavefront each): coef[r] = (dt (p_r - d_r) dp/du_f,
// f = 0..3; 1/2 dt (p_r - d_r)^2)
__global__ void __launch_bounds__(64) k_pobs2_res(ObsArgs a) {
  const int r = blockIdx.x;
  double u[4], vx, vy;
  receiver_state(a.U, a.rec_e[r], a.rec_w + (size_t)r * a.L2, a.ne, a.L2, u);
  if (threadIdx.x == 0) {
    const double gm1 = a.gamma - 1.0, p = receiver_pressure(u, gm1, vx, vy);
    const double dr = p - a.d[r], res = a.dt * dr;
    double *c = a.coef + 5 * (size_t)r;
    c[0] = res * (0.5 * gm1 * (vx * vx + vy * vy));
    c[1] = res * (-gm1 * vx);
    c[2] = res * (-gm1 * vy);
    c[3] = res * gm1;
    c[4] = 0.5 * a.dt * dr * dr;
  }
}

// then per element holding receivers (one workgroup each, receivers in their
// order, so the sums are deterministic without atomics):
// lam[f][e][m] += sum_r coef[r][f] w[r][m]; workgroup 0 adds sum_r coef[r][4]
// to J in receiver order
__global__ void __launch_bounds__(64) k_pobs2_scatter(ObsArgs a) {
  const int g = blockIdx.x, e = a.grp_e[g], r0 = a.grp_ptr[g], r1 = a.grp_ptr[g + 1];
  for (int t = threadIdx.x; t < 4 * a.L2; t += 64) {
    const int f = t / a.L2, m = t - f * a.L2;
    const size_t o = ((size_t)f * a.ne + e) * a.L2 + m;
    double v = a.lam[o];
    for (int k = r0; k < r1; ++k) {
      const int r = a.grp_r[k];
      v += a.coef[5 * (size_t)r + f] * a.rec_w[(size_t)r * a.L2 + m];
    }
    a.lam[o] = v;
  }
  if (g == 0 && threadIdx.x == 0) {
    double jacc = 0.0;
    for (int r = 0; r < a.nrec; ++r) jacc += a.coef[5 * (size_t)r + 4];
    a.J[0] += jacc;
  }
}

template <int P>
static hipError_t launch_adj2_p(const AdjArgs &a) {
  if (a.op == 0) hipLaunchKernelGGL(k_adjtraces2<P>, dim3(a.ne), dim3(64), 0, a.stream, a);
  else hipLaunchKernelGGL(k_vjp2<P>, dim3(a.ne), dim3(64), 0, a.stream, a);
  return hipGetLastError();
}

hipError_t launch_adj2(const AdjArgs &a) {
  switch (a.p) {
#define DGX_CASE(PP) case PP: return launch_adj2_p<PP>(a);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9) DGX_CASE(10)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_pobs2(const ObsArgs &a) {
  if (a.nrec < 1) return hipSuccess;
  hipLaunchKernelGGL(k_pobs2, dim3(a.nrec), dim3(64), 0, a.stream, a);
  return hipGetLastError();
}

hipError_t launch_pobs2_adj(const ObsArgs &a) {
  if (a.nrec < 1) return hipSuccess;
  hipLaunchKernelGGL(k_pobs2_res, dim3(a.nrec), dim3(64), 0, a.stream, a);
  hipLaunchKernelGGL(k_pobs2_scatter, dim3(a.ngrp), dim3(64), 0, a.stream, a);
  return hipGetLastError();
}

hipError_t upload_tables_a2d(const double *h) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), h, sizeof(double) * kTabTotal);
  if (e != hipSuccess) return e;
  std::vector<double> t(kAdjTotal);
  for (int p = 1; p <= kMaxOrder; ++p) {
    const int L = p + 1, Q = p + 2, QF = p + 1;
    const double *BT = h + tab_offset(p), *BW = BT + Q * L, *QD = BW + L * Q, *Bg = QD + L * Q;
    const double *PW = Bg + QF * L + 2 * L;
    double *o = t.data() + adj_offset(p);
    double *BTt = o, *BWt = o + L * Q, *QDt = o + 2 * L * Q, *PWt = o + 3 * L * Q, *Bgt = PWt + QF * L;
    for (int i = 0; i < Q; ++i)
      for (int l = 0; l < L; ++l) {
        BTt[l * Q + i] = BT[i * L + l];
        BWt[i * L + l] = BW[l * Q + i];
        QDt[i * L + l] = QD[l * Q + i];
      }
    for (int g = 0; g < QF; ++g)
      for (int l = 0; l < L; ++l) {
        PWt[g * L + l] = PW[l * QF + g];
        Bgt[l * QF + g] = Bg[g * L + l];
      }
  }
  return hipMemcpyToSymbol(HIP_SYMBOL(c_adj), t.data(), sizeof(double) * kAdjTotal);
}

}  // namespace dgx
