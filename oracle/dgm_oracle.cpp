// ORACLE — test infrastructure only.
//
// A plain C++ CPU restatement of DGM's explicit discontinuous-Galerkin
// residual (Domain::time_derivative) and its explicit Runge-Kutta steppers
// (TimeInt::advance) for compressible Euler and Navier-Stokes on affine quads
// and hexes.  It keeps the reference's data layout (nodal values at q = p+2
// Gauss-Lobatto-Legendre points per direction, face traces at p+1 Gauss
// points) and its operator sequence (fill_sides -> volume gradients ->
// master-side Riemann flux -> lift -> project), so that it can be checked
// against the reference's golden restart files (runs/*.ref) and then used as
// the checker for the HIP product path.
//
// Only tests/, the smoke check and the benchmark's CPU-baseline leg may load
// this library.  The product (dgm_amd/libdgx.so) never links or calls it.
//
// Reference files followed (paths relative to the reference root):
//   src/Polylib.tpp:279-500,1008-1052,1296-1434  quadrature, D, interpolation
//   src/Basis.cpp:74-137                          normalized Legendre modes
//   src/BaseQuad.cpp:28-72, src/BaseHex.cpp:41-   basis construction
//   src/Quad.cpp:263-359, 632-764, 767-946        mass, transforms, sides
//   src/Quad.cpp:1227-1452                        affine geometry, edge geometry
//   src/Quad.cpp:1971-2038, 2189-2330             lift, gradient
//   src/Hex.cpp:703-866, 1124-1494, 1786-1952, 1983-2071
//   src/Euler.cpp:296-757                         2D convection, LF/Roe flux
//   src/Euler3d.cpp:468-817, 1144-1199            3D convection, LF flux
//   src/Navier_Stokes.cpp:113-258                 2D BR1 diffusion
//   src/Navier_Stokes3d.cpp:182-391               3D BR1 diffusion
//   src/LinEuler_quasi.cpp:151-495                2D linearized Euler (uniform mean flow)
//   src/Shallow.cpp:251-525, src/ShallowBC.hpp:45-62  shallow water: convection, LF flux,
//                                                 wall, balance term (pinned by runs/shallow_mms)
//   src/BC.cpp:95-200, src/Source.cpp:11-75       function boundaries and sources (pinned
//                                                 by runs/shallow_mms)
//   src/Navier_Stokes3d.cpp:472-512, src/Channel.hpp:117-163  isothermal walls, bulk forcing
//                                                 (pinned by runs/chan: 3D NS BR1, p = 3)
//   src/Euler3d.cpp:876-909, src/BC.cpp:17-28     3D slip wall and state boundaries
//   src/EulerBC.hpp:43-62                         2D slip wall
//   src/Hex.cpp:289-318, 3992-4080                curved hexes: nodal metrics, mass matrix
//   src/Quad.cpp:263-300, 1309-1452               curved quads: nodal metrics, edges, mass matrix
//                                                 (pinned by runs/cvquad and runs/shallow_mms)
//   (no reference counterpart)                    3D BR2 face gradient: its parity
//                                                 against the reference is unpinned; the
//                                                 BR2 tests pin it by invariants
//   src/Domain.cpp:1883-1937                      average_flux
//   src/Field.cpp:87-90, 416-668, 3478-3556       links, master sides, flips
//   src/BaseQuad.hpp:53-57, src/BaseHex.hpp:291-313  sideMap
//   src/TimeInt.cpp:208-302, 738-858, 1870-2013   RK4, TVD-RK3, LSRK14
//   src/Quad.cpp:2556-2632, src/Field.cpp:4675-4697  norms
//   src/GpGpu.cu:236-262, 505-750, src/Hex.cpp:703-812  element-block transforms (pinned by
//                                                 the GLL mass identity and brute-force sums)
//
// Compile with -ffp-contract=off so that every multiply and add rounds
// separately, as in the reference's plain BLAS-1 loops.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <vector>
#include <string>
#include <algorithm>

namespace {

typedef long double ld;

// ---------------------------------------------------------------------------
// Polylib restatement (src/Polylib.tpp): Legendre recurrences evaluated in
// long double, rounded once to double.
// ---------------------------------------------------------------------------

// P_n(z) and P_n'(z)
void legendre_pd(int n, ld z, ld &P, ld &dP) {
  ld p0 = 1, p1 = z;
  if (n == 0) { P = 1; dP = 0; return; }
  for (int k = 1; k < n; ++k) {
    ld p2 = ((2 * k + 1) * z * p1 - k * p0) / (k + 1);
    p0 = p1; p1 = p2;
  }
  P = p1;
  if (fabsl(fabsl(z) - 1) < 1e-30L) {
    ld s = (z > 0) ? 1 : ((n % 2) ? 1 : -1);   // P_n'(+-1) = (+-1)^(n-1) n(n+1)/2
    dP = s * n * (n + 1) / 2.0L;
  } else {
    dP = n * (z * p1 - p0) / (z * z - 1);
  }
}

// Gauss-Lobatto-Legendre points (ascending) and weights, np points
// (src/Polylib.tpp:452-485 zwglj with alpha=beta=0)
void zwgll(int np, std::vector<double> &z, std::vector<double> &w) {
  z.assign(np, 0.0); w.assign(np, 0.0);
  if (np == 1) { z[0] = 0; w[0] = 2; return; }
  const int N = np - 1;
  std::vector<ld> zz(np);
  zz[0] = -1; zz[N] = 1;
  for (int i = 1; i < N; ++i) {
    ld x = -cosl(M_PIl * i / N);
    for (int it = 0; it < 100; ++it) {
      ld P, dP; legendre_pd(N, x, P, dP);
      ld d2P = (2 * x * dP - N * (N + 1) * P) / (1 - x * x);
      ld dx = dP / d2P;
      x -= dx;
      if (fabsl(dx) < 1e-30L) break;
    }
    zz[i] = x;
  }
  for (int i = 0; i < np; ++i) {
    ld P, dP; legendre_pd(N, zz[i], P, dP);
    z[i] = (double)zz[i];
    w[i] = (double)(2.0L / (N * (N + 1) * P * P));
  }
}

// Gauss-Legendre points (ascending) and weights (src/Polylib.tpp:389-406)
void zwgl(int np, std::vector<double> &z, std::vector<double> &w) {
  z.assign(np, 0.0); w.assign(np, 0.0);
  for (int i = 0; i < np; ++i) {
    ld x = -cosl(M_PIl * (4 * i + 3) / (4 * np + 2));
    for (int it = 0; it < 100; ++it) {
      ld P, dP; legendre_pd(np, x, P, dP);
      ld dx = P / dP;
      x -= dx;
      if (fabsl(dx) < 1e-30L) break;
    }
    ld P, dP; legendre_pd(np, x, P, dP);
    z[i] = (double)x;
    w[i] = (double)(2.0L / ((1 - x * x) * dP * dP));
  }
}

// GLL collocation derivative D[i][j] = l_j'(z_i) (src/Polylib.tpp:1008-1045)
void dgll(const std::vector<double> &z, std::vector<double> &D) {
  const int np = (int)z.size(), N = np - 1;
  D.assign(np * np, 0.0);
  std::vector<ld> PN(np);
  for (int i = 0; i < np; ++i) { ld dP; legendre_pd(N, (ld)z[i], PN[i], dP); }
  for (int i = 0; i < np; ++i)
    for (int j = 0; j < np; ++j) {
      ld v;
      if (i != j) v = PN[i] / (PN[j] * ((ld)z[i] - (ld)z[j]));
      else if (i == 0) v = -(ld)N * (N + 1) / 4;
      else if (i == N) v = (ld)N * (N + 1) / 4;
      else v = 0;
      D[i * np + j] = (double)v;
    }
}

// Lagrange interpolation matrix im[i][j] = h_j(to_i) through nodes `from`
// (src/Polylib.tpp:1368-1430 igjm/igljm)
void interp_matrix(const std::vector<double> &from, const std::vector<double> &to,
                   std::vector<double> &im) {
  const int nf = (int)from.size(), nt = (int)to.size();
  im.assign(nt * nf, 0.0);
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nf; ++j) {
      ld h = 1;
      for (int m = 0; m < nf; ++m)
        if (m != j) h *= ((ld)to[i] - from[m]) / ((ld)from[j] - from[m]);
      im[i * nf + j] = (double)h;
    }
}

// Normalized Legendre modes B[l][i] = sqrt((2l+1)/2) P_l(z_i) (src/Basis.cpp:97-102)
void legendre_modes(int L, const std::vector<double> &z, std::vector<double> &B) {
  const int q = (int)z.size();
  B.assign(L * q, 0.0);
  for (int l = 0; l < L; ++l)
    for (int i = 0; i < q; ++i) {
      ld P, dP; legendre_pd(l, (ld)z[i], P, dP);
      B[l * q + i] = (double)P * std::sqrt(0.5 * (2.0 * l + 1.0));
    }
}

// ---------------------------------------------------------------------------
// Small dense helpers with the reference's row-major matmult semantics
// (src/Utils.hpp:1487-1570): C = op(A) * op(B), ascending inner index.
// ---------------------------------------------------------------------------
// C[m][n] = sum_k A[m][k] B[k][n]
void mm(int M, int N, int K, const double *A, int lda, const double *B, int ldb,
        double *C, int ldc) {
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      double s = 0;
      for (int k = 0; k < K; ++k) s += A[i * lda + k] * B[k * ldb + j];
      C[i * ldc + j] = s;
    }
}
// C[m][n] = sum_k A[k][m] B[k][n]   (A transposed)
void mm_tn(int M, int N, int K, const double *A, int lda, const double *B, int ldb,
           double *C, int ldc) {
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      double s = 0;
      for (int k = 0; k < K; ++k) s += A[k * lda + i] * B[k * ldb + j];
      C[i * ldc + j] = s;
    }
}
// C[m][n] = sum_k A[m][k] B[n][k]   (B transposed)
void mm_nt(int M, int N, int K, const double *A, int lda, const double *B, int ldb,
           double *C, int ldc) {
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < N; ++j) {
      double s = 0;
      for (int k = 0; k < K; ++k) s += A[i * lda + k] * B[j * ldb + k];
      C[i * ldc + j] = s;
    }
}

// 2D Riemann flux at one face point (src/Euler.cpp:399-757): physical normal
// fluxes of both states (Fpl, Fmr, Cartesian) and the numerical flux Fh.
// Returns 1 for a non-realizable state.
int point_flux2d(double gamma, int flux_type, const double Ul[4], const double Ur[4], double nxl, double nyl,
                 double Fpl[4], double Fmr[4], double Fh[4]) {
  const double gm1 = gamma - 1.0, pt5 = 0.5;
  int status = 0;
  const double rl = Ul[0], rul = Ul[1] * nxl + Ul[2] * nyl, rvl = Ul[2] * nxl - Ul[1] * nyl, rEl = Ul[3];
  const double irl = 1.0 / rl;
  const double ul = irl * rul, vl = irl * rvl;
  const double pl = gm1 * (rEl - pt5 * (rul * ul + rvl * vl));
  const double Hl = irl * (rEl + pl);
  if (!(rl > 0) || !(pl > 0)) status = 1;
  Fpl[0] = rul;
  Fpl[1] = (rul * ul + pl) * nxl - (rvl * ul) * nyl;
  Fpl[2] = (rul * ul + pl) * nyl + (rvl * ul) * nxl;
  Fpl[3] = ul * (rEl + pl);
  const double rr = Ur[0], rur = Ur[1] * nxl + Ur[2] * nyl, rvr = Ur[2] * nxl - Ur[1] * nyl, rEr = Ur[3];
  const double irr = 1.0 / rr;
  const double ur = irr * rur, vr = irr * rvr;
  const double pr = gm1 * (rEr - pt5 * (rur * ur + rvr * vr));
  const double Hr = irr * (rEr + pr);
  if (!(rr > 0) || !(pr > 0)) status = 1;
  Fmr[0] = rr * ur;
  Fmr[1] = (rur * ur + pr) * nxl - (rvr * ur) * nyl;
  Fmr[2] = (rur * ur + pr) * nyl + (rvr * ur) * nxl;
  Fmr[3] = ur * (rEr + pr);
  if (flux_type == 0) {
    double lmax = 0;
    const double al = std::sqrt(gm1 * (Hl - pt5 * (ul * ul + vl * vl)));
    const double laml[4] = {ul - al, ul, ul, ul + al};
    for (int n = 0; n < 4; ++n) lmax = std::max(lmax, std::fabs(laml[n]));
    const double ar = std::sqrt(gm1 * (Hr - pt5 * (ur * ur + vr * vr)));
    const double lamr[4] = {ur - ar, ur, ur, ur + ar};
    for (int n = 0; n < 4; ++n) lmax = std::max(lmax, std::fabs(lamr[n]));
    for (int n = 0; n < 4; ++n) Fh[n] = pt5 * (Fpl[n] + Fmr[n] + lmax * (Ul[n] - Ur[n]));
  } else {
    // Roe average flux without entropy fix (src/Euler.cpp:682-741)
    const double R = std::sqrt(rr / rl);
    const double iRp1 = 1.0 / (1.0 + R);
    const double uu = (ul + R * ur) * iRp1, vv = (vl + R * vr) * iRp1, H = (Hl + R * Hr) * iRp1;
    const double a2 = gm1 * (H - pt5 * (uu * uu + vv * vv));
    const double a = std::sqrt(a2), ainv = 1.0 / a;
    const double c1 = gm1 * ainv * ainv, c1inv = 1.0 / c1;
    double alpha[4], lam[4], Gh[4];
    alpha[2] = (rvr - rvl) - vv * (rr - rl);
    alpha[1] = (gamma - 1.0) * ainv * ainv * ((rr - rl) * (H - uu * uu) + uu * (rur - rul) -
                                             ((rEr - rEl) - ((rvr - rvl) - vv * (rr - rl)) * vv));
    alpha[0] = pt5 * ainv * ((rr - rl) * (uu + a) - (rur - rul) - a * alpha[1]);
    alpha[3] = (rr - rl) - (alpha[0] + alpha[1]);
    lam[0] = std::min(uu - a, 0.0); lam[1] = std::min(uu, 0.0);
    lam[2] = std::min(uu, 0.0);     lam[3] = std::min(uu + a, 0.0);
    Gh[0] = lam[0] * alpha[0];
    Gh[1] = lam[0] * (uu - a) * alpha[0];
    Gh[2] = lam[0] * vv * alpha[0];
    Gh[3] = lam[0] * (H - uu * a) * alpha[0];
    Gh[0] += lam[1] * alpha[1];
    Gh[1] += lam[1] * uu * alpha[1];
    Gh[2] += lam[1] * vv * alpha[1];
    Gh[3] += lam[1] * (H - c1inv) * alpha[1];
    Gh[2] += lam[2] * alpha[2];
    Gh[3] += lam[2] * vv * alpha[2];
    Gh[0] += lam[3] * alpha[3];
    Gh[1] += lam[3] * (uu + a) * alpha[3];
    Gh[2] += lam[3] * vv * alpha[3];
    Gh[3] += lam[3] * (H + uu * a) * alpha[3];
    Fh[0] = Fpl[0] + Gh[0];
    Fh[1] = Fpl[1] + Gh[1] * nxl - Gh[2] * nyl;
    Fh[2] = Fpl[2] + Gh[1] * nyl + Gh[2] * nxl;
    Fh[3] = Fpl[3] + Gh[3];
  }
  return status;
}

// 3D Riemann flux at one face point in the rotated frame
// (src/Euler3d.cpp:584-817); as point_flux2d.
int point_flux3d(double gamma, int flux_type, const double Ul[5], const double Ur[5], double nx, double ny,
                 double nz, double Fpl[5], double Fmr[5], double Fh[5]) {
  const double gm1 = gamma - 1.0, pt5 = 0.5;
  int status = 0;
  double sx = ny, sy = nz, sz = nx;
  const double sdn = nx * sx + ny * sy + nz * sz;
  sx -= sdn * nx; sy -= sdn * ny; sz -= sdn * nz;
  const double snrm_inv = 1.0 / std::sqrt(sx * sx + sy * sy + sz * sz);
  sx *= snrm_inv; sy *= snrm_inv; sz *= snrm_inv;
  double tx = ny * sz - sy * nz, ty = nz * sx - sz * nx, tz = nx * sy - sx * ny;
  const double tnrm_inv = 1.0 / std::sqrt(tx * tx + ty * ty + tz * tz);
  tx *= tnrm_inv; ty *= tnrm_inv; tz *= tnrm_inv;
  const double detA = (nx * sy * tz + sx * ty * nz + tx * sz * ny) - (nz * sy * tx + sz * ty * nx + tz * sx * ny);
  const double idetA = 1.0 / detA;
  const double xn = (sy * tz - sz * ty) * idetA, yn = -(sx * tz - sz * tx) * idetA, zn = (sx * ty - sy * tx) * idetA;
  const double xs = -(ny * tz - nz * ty) * idetA, ys = (nx * tz - nz * tx) * idetA, zs = -(nx * ty - ny * tx) * idetA;
  const double xt = (ny * sz - nz * sy) * idetA, yt = -(nx * sz - nz * sx) * idetA, zt = (nx * sy - ny * sx) * idetA;
  const double rl = Ul[0];
  const double rul = Ul[1] * nx + Ul[2] * ny + Ul[3] * nz;
  const double rvl = Ul[1] * sx + Ul[2] * sy + Ul[3] * sz;
  const double rwl = Ul[1] * tx + Ul[2] * ty + Ul[3] * tz;
  const double rEl = Ul[4];
  const double irl = 1.0 / rl;
  const double ul = irl * rul, vl = irl * rvl, wl = irl * rwl;
  const double pl = gm1 * (rEl - pt5 * (rul * ul + rvl * vl + rwl * wl));
  const double Hl = irl * (rEl + pl);
  if (!(rl > 0) || !(pl > 0)) status = 1;
  Fpl[0] = rl * ul;
  Fpl[1] = (rul * ul + pl) * xn + (rvl * ul) * xs + (rwl * ul) * xt;
  Fpl[2] = (rul * ul + pl) * yn + (rvl * ul) * ys + (rwl * ul) * yt;
  Fpl[3] = (rul * ul + pl) * zn + (rvl * ul) * zs + (rwl * ul) * zt;
  Fpl[4] = ul * (rEl + pl);
  const double rr = Ur[0];
  const double rur = Ur[1] * nx + Ur[2] * ny + Ur[3] * nz;
  const double rvr = Ur[1] * sx + Ur[2] * sy + Ur[3] * sz;
  const double rwr = Ur[1] * tx + Ur[2] * ty + Ur[3] * tz;
  const double rEr = Ur[4];
  const double irr = 1.0 / rr;
  const double ur = irr * rur, vr = irr * rvr, wr = irr * rwr;
  const double pr = gm1 * (rEr - pt5 * (rur * ur + rvr * vr + rwr * wr));
  const double Hr = irr * (rEr + pr);
  if (!(rr > 0) || !(pr > 0)) status = 1;
  Fmr[0] = rr * ur;
  Fmr[1] = (rur * ur + pr) * xn + (rvr * ur) * xs + (rwr * ur) * xt;
  Fmr[2] = (rur * ur + pr) * yn + (rvr * ur) * ys + (rwr * ur) * yt;
  Fmr[3] = (rur * ur + pr) * zn + (rvr * ur) * zs + (rwr * ur) * zt;
  Fmr[4] = ur * (rEr + pr);
  if (flux_type == 0) {
    double lmax = 0;
    const double al = std::sqrt(gm1 * (Hl - pt5 * (ul * ul + vl * vl + wl * wl)));
    const double laml[5] = {ul - al, ul, ul, ul, ul + al};
    for (int n = 0; n < 5; ++n) lmax = std::max(lmax, std::fabs(laml[n]));
    const double ar = std::sqrt(gm1 * (Hr - pt5 * (ur * ur + vr * vr + wr * wr)));
    const double lamr[5] = {ur - ar, ur, ur, ur, ur + ar};
    for (int n = 0; n < 5; ++n) lmax = std::max(lmax, std::fabs(lamr[n]));
    for (int n = 0; n < 5; ++n) Fh[n] = pt5 * (Fpl[n] + Fmr[n] + lmax * (Ul[n] - Ur[n]));
  } else {
    // Roe flux without entropy fix: the reference's 2D Roe flux
    // (src/Euler.cpp:682-741) in the rotated frame (n, s, t) of
    // src/Euler3d.cpp:584-817, with a second shear wave for t.  The
    // reference has no 3D Roe flux: parity unpinned, pinned by
    // consistency, conservation and its reduction to the 2D flux.
    const double R = std::sqrt(rr / rl);
    const double iRp1 = 1.0 / (1.0 + R);
    const double uu = (ul + R * ur) * iRp1, vv = (vl + R * vr) * iRp1, ww = (wl + R * wr) * iRp1;
    const double H = (Hl + R * Hr) * iRp1;
    const double a2 = gm1 * (H - pt5 * (uu * uu + vv * vv + ww * ww));
    const double a = std::sqrt(a2), ainv = 1.0 / a;
    const double c1inv = a2 / gm1;
    double alpha[5], lam[5], Gh[5];
    alpha[2] = (rvr - rvl) - vv * (rr - rl);
    alpha[3] = (rwr - rwl) - ww * (rr - rl);
    alpha[1] = gm1 * ainv * ainv * ((rr - rl) * (H - uu * uu) + uu * (rur - rul) -
                                    ((rEr - rEl) - alpha[2] * vv - alpha[3] * ww));
    alpha[0] = pt5 * ainv * ((rr - rl) * (uu + a) - (rur - rul) - a * alpha[1]);
    alpha[4] = (rr - rl) - (alpha[0] + alpha[1]);
    lam[0] = std::min(uu - a, 0.0); lam[1] = std::min(uu, 0.0); lam[2] = lam[1];
    lam[3] = lam[1];                lam[4] = std::min(uu + a, 0.0);
    Gh[0] = lam[0] * alpha[0] + lam[1] * alpha[1] + lam[4] * alpha[4];
    Gh[1] = lam[0] * (uu - a) * alpha[0] + lam[1] * uu * alpha[1] + lam[4] * (uu + a) * alpha[4];
    Gh[2] = lam[0] * vv * alpha[0] + lam[1] * vv * alpha[1] + lam[2] * alpha[2] + lam[4] * vv * alpha[4];
    Gh[3] = lam[0] * ww * alpha[0] + lam[1] * ww * alpha[1] + lam[3] * alpha[3] + lam[4] * ww * alpha[4];
    Gh[4] = lam[0] * (H - uu * a) * alpha[0] + lam[1] * (H - c1inv) * alpha[1] +
            lam[2] * vv * alpha[2] + lam[3] * ww * alpha[3] + lam[4] * (H + uu * a) * alpha[4];
    Fh[0] = Fpl[0] + Gh[0];
    Fh[1] = Fpl[1] + Gh[1] * xn + Gh[2] * xs + Gh[3] * xt;
    Fh[2] = Fpl[2] + Gh[1] * yn + Gh[2] * ys + Gh[3] * yt;
    Fh[3] = Fpl[3] + Gh[1] * zn + Gh[2] * zs + Gh[3] * zt;
    Fh[4] = Fpl[4] + Gh[4];
  }
  return status;
}

// Shallow-water Lax-Friedrichs flux at one face point (src/Shallow.cpp:357-525)
// in (phi, uphi, vphi), phi the geopotential g(h + eta): normal fluxes of both
// states (Cartesian) and the numerical flux; the dissipation speed is the
// largest |u_n| + sqrt(phi) of the two states.  Returns 1 when phi or the
// pressure phi^2/2 of either state is below aepsM = eps^(2/3)
// (src/Constants.hpp:41-43), where the reference stops.
int point_flux_shallow(const double Ul[3], const double Ur[3], double nxl, double nyl, double Fpl[3],
                       double Fmr[3], double Fh[3]) {
  const double pt5 = 0.5, aepsM = std::pow(2.220446049250313e-16, 0.6666666666666666667);
  int status = 0;
  const double phil = Ul[0], uphil = Ul[1] * nxl + Ul[2] * nyl, vphil = Ul[2] * nxl - Ul[1] * nyl;
  const double iphil = 1.0 / phil, ul = iphil * uphil, pl = pt5 * phil * phil;
  if (!(phil >= aepsM) || !(pl >= aepsM)) status = 1;
  Fpl[0] = uphil;
  Fpl[1] = (uphil * ul + pl) * nxl - (vphil * ul) * nyl;
  Fpl[2] = (uphil * ul + pl) * nyl + (vphil * ul) * nxl;
  const double phir = Ur[0], uphir = Ur[1] * nxl + Ur[2] * nyl, vphir = Ur[2] * nxl - Ur[1] * nyl;
  const double iphir = 1.0 / phir, ur = iphir * uphir, pr = pt5 * phir * phir;
  if (!(phir >= aepsM) || !(pr >= aepsM)) status = 1;
  Fmr[0] = uphir;
  Fmr[1] = (uphir * ur + pr) * nxl - (vphir * ur) * nyl;
  Fmr[2] = (uphir * ur + pr) * nyl + (vphir * ur) * nxl;
  double lmax = 0.0;
  const double al = std::sqrt(phil), ar = std::sqrt(phir);
  const double laml[3] = {ul - al, ul, ul + al}, lamr[3] = {ur - ar, ur, ur + ar};
  for (int n = 0; n < 3; ++n) lmax = std::max(lmax, std::fabs(laml[n]));
  for (int n = 0; n < 3; ++n) lmax = std::max(lmax, std::fabs(lamr[n]));
  for (int n = 0; n < 3; ++n) Fh[n] = pt5 * (Fpl[n] + Fmr[n] + lmax * (Ul[n] - Ur[n]));
  return status;
}

// ---------------------------------------------------------------------------
// Domain
// ---------------------------------------------------------------------------

enum Eqn { EULER2D = 1, NS2D = 2, LINEULER2D = 4, EULER3D = 8, NS3D = 9, SHALLOW2D = 30 };

// LSRK14 coefficients (src/TimeInt.cpp:1870-1917)
const double kLsrkA[15] = {0.0, -0.7188012108672410, -0.7785331173421570, -0.0053282796654044,
  -0.8552979934029281, -3.9564138245774565, -1.5780575380587385, -2.0837094552574054,
  -0.7483334182761610, -0.7032861106563359, 0.0013917096117681, -0.0932075369637460,
  -0.9514200470875948, -7.1151571693922548, 0.0};
const double kLsrkB[14] = {0.0367762454319673, 0.3136296607553959, 0.1531848691869027,
  0.0030097086818182, 0.3326293790646110, 0.2440251405350864, 0.3718879239592277,
  0.6204126221582444, 0.1524043173028741, 0.0760894927419266, 0.0077604214040978,
  0.0024647284755382, 0.0780348340049386, 5.5059777270269628};
const double kLsrkC[14] = {0.0, 0.0367762454319673, 0.1249685262725025, 0.2446177702277698,
  0.2476149531070420, 0.2969311120382472, 0.3978149645802642, 0.5270854589440328, 0.6981269994175695,
  0.8190890835352128, 0.8527059887098624, 0.8604711817462826, 0.8627060376969976, 0.8734213127600976};

struct SideInfo {
  bool master = false;       // side->link set (src/SideLoop.hpp:46-50)
  int le = -1, ls = -1;      // linked element/side (both directions)
  bool flip_a = false, flip_b = false, transp = false;
  int bc = -1;               // boundary side: 0 slip wall, 1 state, 2 function of (x, y, t),
                             // 3 isothermal wall (no link)
  std::vector<double> nx, ny, nz, J;   // Gauss-point normal and J_side/J_vol
};

struct Elem {
  double J = 0;
  double rx = 0, ry = 0, rz = 0, sx = 0, sy = 0, sz = 0, tx = 0, ty = 0, tz = 0;
  std::vector<double> wJ, cx, cy, cz;
  SideInfo side[6];
  // hexes: metrics rx ry rz sx sy sz tx ty tz and J at the nodes (constant
  // when affine); curved elements also keep the Cholesky factor of their mass
  // matrix B^T diag(wJ) B (src/Hex.cpp:289-318)
  bool curved = false;
  std::vector<double> mn[9], Jn, chol;
};

struct VF {                  // vField: nodal element data + side data
  std::vector<double> u, s;
};

struct Domain {
  int nsd, ne, p, L, q, qf, qtot, nm, nsides, qs, nv, eqn;
  double gamma, vis, Pr;
  int status = 0;
  std::vector<double> z, w, zg, wg, B, D, Dt, Ig, Ii;
  std::vector<Elem> el;
  VF U, Ut, R0, R1, Uv, Ub;   // Ub: BR2 normal viscous flux on the sides

  double *eu(VF &F, int v, int e) { return &F.u[((size_t)v * ne + e) * qtot]; }
  double *es(VF &F, int v, int e, int s) {
    return &F.s[(((size_t)v * ne + e) * nsides + s) * qs];
  }
  void alloc(VF &F) {
    F.u.assign((size_t)nv * ne * qtot, 0.0);
    F.s.assign((size_t)nv * ne * nsides * qs, 0.0);
  }

  // ---- sideMap (src/BaseQuad.hpp:53-57, src/BaseHex.hpp:291-313) ----
  int sidemap(const SideInfo &s, int ql) const {
    if (nsd == 2) return s.flip_a ? (qf - 1) - ql : ql;
    const int q1 = qf, q2 = qf;
    const int il = ql / q2, jl = ql % q2;
    int i, j;
    if (s.transp) {
      j = il; i = jl;
      if (s.flip_a) j = (q1 - 1) - j;
      if (s.flip_b) i = (q2 - 1) - i;
    } else {
      i = il; j = jl;
      if (s.flip_a) i = (q1 - 1) - i;
      if (s.flip_b) j = (q2 - 1) - j;
    }
    return j + q2 * i;
  }
  // the slave side never has its flip flags set (src/Field.cpp:428)
  int sidemap_slave(int ql) const { return ql; }

  // ---- get_side (src/Quad.cpp:767-789, src/Hex.cpp:817-866) ----
  void get_side(const double *f, int sid, double *to) const {
    if (nsd == 2) {
      const int qa = q, qb = q;
      switch (sid) {
        case 0: for (int p_ = 0; p_ < qa; ++p_) to[p_] = f[0 + p_ * qb]; break;
        case 1: for (int j = 0; j < qb; ++j) to[j] = f[j + (qa - 1) * qb]; break;
        case 2: for (int p_ = 0; p_ < qa; ++p_) to[p_] = f[(qb - 1) + p_ * qb]; break;
        case 3: for (int j = 0; j < qb; ++j) to[j] = f[j]; break;
      }
      return;
    }
    const int qa = q, qb = q, qc = q;
    auto F = [&](int i, int j, int k) { return f[(i * qb + j) * qc + k]; };
    switch (sid) {
      case 0: for (int i = 0; i < qa; ++i) for (int j = 0; j < qb; ++j) to[i * qb + j] = F(i, j, 0); break;
      case 1: for (int i = 0; i < qa; ++i) for (int k = 0; k < qc; ++k) to[i * qc + k] = F(i, 0, k); break;
      case 2: for (int j = 0; j < qb; ++j) for (int k = 0; k < qc; ++k) to[j * qc + k] = F(qa - 1, j, k); break;
      case 3: for (int i = 0; i < qa; ++i) for (int k = 0; k < qc; ++k) to[i * qc + k] = F(i, qb - 1, k); break;
      case 4: for (int j = 0; j < qb; ++j) for (int k = 0; k < qc; ++k) to[j * qc + k] = F(0, j, k); break;
      case 5: for (int i = 0; i < qa; ++i) for (int j = 0; j < qb; ++j) to[i * qb + j] = F(i, j, qc - 1); break;
    }
  }

  // GLL side values -> Gauss side values (Quad::fill_sides mvmult,
  // Hex::interpolate_to_side im1 * From * im2^T)
  void to_side(const double *from, double *to) const {
    if (nsd == 2) { mm(qf, 1, q, Ig.data(), q, from, 1, to, 1); return; }
    std::vector<double> wk(qf * q);
    mm(qf, q, q, Ig.data(), q, from, q, wk.data(), q);
    mm_nt(qf, qf, q, wk.data(), q, Ig.data(), q, to, qf);
  }
  // Gauss side values -> GLL side values (G2GL)
  void from_side(const double *from, double *to) const {
    if (nsd == 2) { mm(q, 1, qf, Ii.data(), qf, from, 1, to, 1); return; }
    std::vector<double> wk(q * qf);
    mm(q, qf, qf, Ii.data(), qf, from, qf, wk.data(), qf);
    mm_nt(q, q, qf, wk.data(), qf, Ii.data(), qf, to, q);
  }

  // ---- setup ----
  void build_basis() {
    L = p + 1; q = p + 2; qf = p + 1;
    zwgll(q, z, w);
    zwgl(qf, zg, wg);
    legendre_modes(L, z, B);
    dgll(z, D);
    Dt.assign(q * q, 0.0);
    for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) Dt[j * q + i] = D[i * q + j];
    interp_matrix(z, zg, Ig);    // GL2G: qf x q
    interp_matrix(zg, z, Ii);    // G2GL: q x qf
    if (nsd == 2) { qtot = q * q; nm = L * L; nsides = 4; qs = qf; }
    else { qtot = q * q * q; nm = L * L * L; nsides = 6; qs = qf * qf; }
  }

  // vertex modes of the C0 basis (src/BaseQuad.cpp:98-116)
  static double vm(int which, double x) { return which ? 0.5 * (1.0 + x) : 0.5 * (1.0 - x); }

  void build_geometry(const double *vert) {
    const int nvert = (nsd == 2) ? 4 : 8;
    el.assign(ne, Elem());
    for (int e = 0; e < ne; ++e) {
      Elem &E = el[e];
      const double *vx = vert + (size_t)e * nsd * nvert;
      const double *vy = vx + nvert;
      const double *vz = (nsd == 3) ? vy + nvert : nullptr;
      E.cx.assign(qtot, 0.0); E.cy.assign(qtot, 0.0);
      if (nsd == 3) E.cz.assign(qtot, 0.0);
      if (nsd == 2) {
        static const int ai[4] = {0, 1, 1, 0}, bi[4] = {0, 0, 1, 1};
        for (int n = 0; n < 4; ++n)
          for (int i = 0; i < q; ++i)
            for (int j = 0; j < q; ++j) {
              const double phi = vm(ai[n], z[i]) * vm(bi[n], z[j]);
              E.cx[j + i * q] += vx[n] * phi;
              E.cy[j + i * q] += vy[n] * phi;
            }
        const double xr = (vx[1] - vx[0]) * 0.5, xs = (vx[3] - vx[0]) * 0.5;
        const double yr = (vy[1] - vy[0]) * 0.5, ys = (vy[3] - vy[0]) * 0.5;
        const double detJ = xr * ys - xs * yr;
        E.J = detJ;
        E.rx = ys / detJ; E.ry = -xs / detJ; E.sx = -yr / detJ; E.sy = xr / detJ;
        E.wJ.assign(qtot, detJ);
        for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) E.wJ[i * q + j] *= w[i];
        for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) E.wJ[i * q + j] *= w[j];
        // edge geometry (src/Quad.cpp:1367-1440)
        std::vector<double> wkx(q), wky(q), cst(q), tmp(q), dx(q), dy(q);
        for (int s = 0; s < 4; ++s) {
          SideInfo &S = E.side[s];
          S.nx.assign(qf, 0.0); S.ny.assign(qf, 0.0); S.J.assign(qf, 0.0);
          const double X = (s == 0 || s == 2) ? E.sx : E.rx;
          const double Y = (s == 0 || s == 2) ? E.sy : E.ry;
          std::fill(wkx.begin(), wkx.end(), X);
          std::fill(wky.begin(), wky.end(), Y);
          to_side(wkx.data(), S.nx.data());
          to_side(wky.data(), S.ny.data());
          if (s == 0 || s == 3)
            for (int g = 0; g < qf; ++g) { S.nx[g] *= -1.0; S.ny[g] *= -1.0; }
          std::fill(cst.begin(), cst.end(), detJ);
          std::vector<double> iJ(qf);
          to_side(cst.data(), iJ.data());
          for (int g = 0; g < qf; ++g) {
            const double len = std::sqrt(S.nx[g] * S.nx[g] + S.ny[g] * S.ny[g]);
            S.nx[g] /= len; S.ny[g] /= len;
          }
          // D along the edge applied to the edge coordinates
          get_side(E.cx.data(), s, tmp.data());
          mm(q, 1, q, D.data(), q, tmp.data(), 1, dx.data(), 1);
          get_side(E.cy.data(), s, tmp.data());
          mm(q, 1, q, D.data(), q, tmp.data(), 1, dy.data(), 1);
          for (int i = 0; i < q; ++i) dx[i] = dx[i] * dx[i];
          for (int i = 0; i < q; ++i) dx[i] = dy[i] * dy[i] + dx[i];
          for (int i = 0; i < q; ++i) tmp[i] = std::sqrt(dx[i]);
          to_side(tmp.data(), S.J.data());
          for (int g = 0; g < qf; ++g) S.J[g] = S.J[g] / iJ[g];
        }
      } else {
        // vertex n -> (a,b,c) corner: 0..3 at c=-1, 4..7 at c=+1
        static const int ai[8] = {0, 1, 1, 0, 0, 1, 1, 0};
        static const int bi[8] = {0, 0, 1, 1, 0, 0, 1, 1};
        static const int ci[8] = {0, 0, 0, 0, 1, 1, 1, 1};
        for (int n = 0; n < 8; ++n)
          for (int i = 0; i < q; ++i)
            for (int j = 0; j < q; ++j)
              for (int k = 0; k < q; ++k) {
                const double phi = vm(ai[n], z[i]) * vm(bi[n], z[j]) * vm(ci[n], z[k]);
                const int m = k + j * q + i * q * q;
                E.cx[m] += vx[n] * phi; E.cy[m] += vy[n] * phi; E.cz[m] += vz[n] * phi;
              }
        const double xr = (vx[1] - vx[0]) * 0.5, xs = (vx[3] - vx[0]) * 0.5, xt = (vx[4] - vx[0]) * 0.5;
        const double yr = (vy[1] - vy[0]) * 0.5, ys = (vy[3] - vy[0]) * 0.5, yt = (vy[4] - vy[0]) * 0.5;
        const double zr = (vz[1] - vz[0]) * 0.5, zs = (vz[3] - vz[0]) * 0.5, zt = (vz[4] - vz[0]) * 0.5;
        const double detJ = xr * ys * zt + yr * zs * xt + zr * xs * yt - xt * ys * zr - zt * xs * yr - xr * zs * yt;
        E.J = detJ;
        E.rx = (ys * zt - yt * zs) / detJ;  E.ry = -(xs * zt - xt * zs) / detJ; E.rz = (xs * yt - xt * ys) / detJ;
        E.sx = -(yr * zt - yt * zr) / detJ; E.sy = (xr * zt - xt * zr) / detJ;  E.sz = -(xr * yt - xt * yr) / detJ;
        E.tx = (yr * zs - ys * zr) / detJ;  E.ty = -(xr * zs - xs * zr) / detJ; E.tz = (xr * ys - xs * yr) / detJ;
        E.wJ.assign(qtot, detJ);
        for (int i = 0; i < q; ++i) {
          for (int j = 0; j < q; ++j) for (int k = 0; k < q; ++k) E.wJ[(i * q + j) * q + k] *= w[j];
          for (int j = 0; j < q; ++j) for (int k = 0; k < q; ++k) E.wJ[(i * q + j) * q + k] *= w[k];
        }
        for (int i = 0; i < q; ++i) for (int jk = 0; jk < q * q; ++jk) E.wJ[i * q * q + jk] *= w[i];
        const double mc[9] = {E.rx, E.ry, E.rz, E.sx, E.sy, E.sz, E.tx, E.ty, E.tz};
        for (int c = 0; c < 9; ++c) E.mn[c].assign(qtot, mc[c]);
        E.Jn.assign(qtot, detJ);
        face_geometry3(E);
      }
    }
  }

  // hex face geometry from the nodal metrics and coordinates (src/Hex.cpp:1385-1494):
  // normals from the metric row of the face's direction interpolated to the
  // Gauss points, J_side from the face tangents, divided by J interpolated alike
  void face_geometry3(Elem &E) {
    {
      {
        const int q2 = q * q;
        std::vector<double> wk(q2), X(q2), Y(q2), Zc(q2), XR(q2), XS(q2), YR(q2), YS(q2), ZR(q2), ZS(q2);
        for (int s = 0; s < 6; ++s) {
          SideInfo &S = E.side[s];
          S.nx.assign(qs, 0.0); S.ny.assign(qs, 0.0); S.nz.assign(qs, 0.0); S.J.assign(qs, 0.0);
          const int row = (s == 0 || s == 5) ? 6 : ((s == 1 || s == 3) ? 3 : 0);
          get_side(E.mn[row].data(), s, wk.data()); to_side(wk.data(), S.nx.data());
          get_side(E.mn[row + 1].data(), s, wk.data()); to_side(wk.data(), S.ny.data());
          get_side(E.mn[row + 2].data(), s, wk.data()); to_side(wk.data(), S.nz.data());
          if (s == 0 || s == 1 || s == 4)
            for (int g = 0; g < qs; ++g) { S.nx[g] *= -1.0; S.ny[g] *= -1.0; S.nz[g] *= -1.0; }
          std::vector<double> iJ(qs);
          get_side(E.Jn.data(), s, wk.data()); to_side(wk.data(), iJ.data());
          for (int g = 0; g < qs; ++g) {
            double t = S.nx[g] * S.nx[g];
            t = S.ny[g] * S.ny[g] + t;
            t = S.nz[g] * S.nz[g] + t;
            t = std::sqrt(t);
            S.nx[g] = S.nx[g] / t; S.ny[g] = S.ny[g] / t; S.nz[g] = S.nz[g] / t;
          }
          get_side(E.cx.data(), s, X.data());
          get_side(E.cy.data(), s, Y.data());
          get_side(E.cz.data(), s, Zc.data());
          mm(q, q, q, D.data(), q, X.data(), q, XR.data(), q);
          mm_nt(q, q, q, X.data(), q, D.data(), q, XS.data(), q);
          mm(q, q, q, D.data(), q, Y.data(), q, YR.data(), q);
          mm_nt(q, q, q, Y.data(), q, D.data(), q, YS.data(), q);
          mm(q, q, q, D.data(), q, Zc.data(), q, ZR.data(), q);
          mm_nt(q, q, q, Zc.data(), q, D.data(), q, ZS.data(), q);
          for (int m = 0; m < q2; ++m) {
            double cxv = YR[m] * ZS[m]; cxv = cxv - ZR[m] * YS[m];
            double cyv = ZR[m] * XS[m]; cyv = cyv - XR[m] * ZS[m];
            double czv = XR[m] * YS[m]; czv = czv - XS[m] * YR[m];
            double t = cxv * cxv; t = cyv * cyv + t; t = czv * czv + t;
            wk[m] = std::sqrt(t);
          }
          to_side(wk.data(), S.J.data());
          for (int g = 0; g < qs; ++g) S.J[g] = S.J[g] / iJ[g];
        }
      }
    }
  }

  // read_connectivity + set_flip (src/Field.cpp:3478-3556, 416-668)
  void build_links(int ncon, const int *con, const double *vert) {
    for (int r = 0; r < ncon; ++r) {
      const int e = con[4 * r], s = con[4 * r + 1], le = con[4 * r + 2], ls = con[4 * r + 3];
      if (e >= 0 && le < 0) { el[e].side[s].bc = ls; continue; }   // boundary side: (e, s, -1, bc type)
      if (e < 0 || le < 0) continue;
      SideInfo &me = el[e].side[s], &nb = el[le].side[ls];
      if (nb.master) continue;         // reverse link already set
      me.master = true; me.le = le; me.ls = ls;
      nb.master = false; nb.le = e; nb.ls = s;
    }
    static const int con2d[4][4] = {{1, 1, 0, 0}, {1, 1, 0, 0}, {0, 0, 1, 1}, {0, 0, 1, 1}};
    static const int svm[6][4] = {{0, 1, 2, 3}, {0, 1, 5, 4}, {1, 2, 6, 5},
                                  {3, 2, 6, 7}, {0, 3, 7, 4}, {4, 5, 6, 7}};
    const int nvert = (nsd == 2) ? 4 : 8;
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < nsides; ++s) {
        SideInfo &S = el[e].side[s];
        if (!S.master) continue;
        S.flip_a = S.flip_b = S.transp = false;
        if (nsd == 2) { S.flip_a = con2d[s][S.ls]; continue; }
        auto V = [&](int el_, int v, double out[3]) {
          const double *b = vert + (size_t)el_ * 3 * nvert;
          out[0] = b[v]; out[1] = b[nvert + v]; out[2] = b[2 * nvert + v];
        };
        auto vec = [&](int el_, int v1, int v0, double out[3]) {
          double a[3], b[3]; V(el_, v1, a); V(el_, v0, b);
          double n = 0;
          for (int d = 0; d < 3; ++d) { out[d] = a[d] - b[d]; n += out[d] * out[d]; }
          n = std::sqrt(n);
          for (int d = 0; d < 3; ++d) out[d] /= n;
        };
        auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        double a0[3], b0[3], a1[3], b1[3], a1p[3];
        vec(e, svm[s][1], svm[s][0], a0);
        vec(e, svm[s][3], svm[s][0], b0);
        vec(S.le, svm[S.ls][1], svm[S.ls][0], a1);
        vec(S.le, svm[S.ls][3], svm[S.ls][0], b1);
        vec(S.le, svm[S.ls][2], svm[S.ls][3], a1p);
        const double a0da1 = dot(a0, a1), a0da1p = dot(a0, a1p);
        auto near1 = [](double x) { return std::fabs(std::fabs(x) - 1.0) < 1e-8; };
        if (near1(a0da1) || near1(a0da1p)) {
          S.transp = false;
          if (a0da1 < 0) S.flip_a = true;
          if (dot(b0, b1) < 0) S.flip_b = true;
        } else {
          S.transp = true;
          if (dot(b0, a1) < 0) S.flip_a = true;
          if (dot(a0, b1) < 0) S.flip_b = true;
        }
      }
  }

  // ---- element operators ----
  // gradient in one direction (src/Quad.cpp:2309-2330, src/Hex.cpp:2030-2071)
  void grad_r(const double *f, double *fr, double *fs, double *ft) const {
    if (nsd == 2) {
      mm(q, q, q, D.data(), q, f, q, fr, q);
      mm(q, q, q, f, q, Dt.data(), q, fs, q);
      return;
    }
    const int q2 = q * q;
    mm(q, q2, q, D.data(), q, f, q2, fr, q2);
    for (int i = 0; i < q; ++i) mm(q, q, q, D.data(), q, f + i * q2, q, fs + i * q2, q);
    mm(q2, q, q, f, q, Dt.data(), q, ft, q);
  }
  void gradient_dir(int e, const double *f, double *df, char dir) const {
    const Elem &E = el[e];
    std::vector<double> fr(qtot), fs(qtot), ft(qtot);
    grad_r(f, fr.data(), fs.data(), ft.data());
    if (E.curved) {   // metrics at the nodes (src/Quad.cpp:2309-2330, src/Hex.cpp:2030-2071 with geom arrays)
      const int c = dir == 'x' ? 0 : (dir == 'y' ? 1 : 2);
      for (int i = 0; i < qtot; ++i) df[i] = E.mn[c][i] * fr[i];
      for (int i = 0; i < qtot; ++i) df[i] = E.mn[3 + c][i] * fs[i] + df[i];
      if (nsd == 3) for (int i = 0; i < qtot; ++i) df[i] = E.mn[6 + c][i] * ft[i] + df[i];
      return;
    }
    double a, b, c = 0;
    if (dir == 'x') { a = E.rx; b = E.sx; c = E.tx; }
    else if (dir == 'y') { a = E.ry; b = E.sy; c = E.ty; }
    else { a = E.rz; b = E.sz; c = E.tz; }
    for (int i = 0; i < qtot; ++i) df[i] = a * fr[i];
    for (int i = 0; i < qtot; ++i) df[i] = b * fs[i] + df[i];
    if (nsd == 3) for (int i = 0; i < qtot; ++i) df[i] = c * ft[i] + df[i];
  }
  void gradient_all(int e, const double *f, double *fx, double *fy, double *fz) const {
    const Elem &E = el[e];
    if (E.curved) {
      gradient_dir(e, f, fx, 'x'); gradient_dir(e, f, fy, 'y');
      if (nsd == 3) gradient_dir(e, f, fz, 'z');
      return;
    }
    std::vector<double> fr(qtot), fs(qtot), ft(qtot);
    grad_r(f, fr.data(), fs.data(), ft.data());
    for (int i = 0; i < qtot; ++i) fx[i] = E.rx * fr[i];
    for (int i = 0; i < qtot; ++i) fx[i] = E.sx * fs[i] + fx[i];
    if (nsd == 3) for (int i = 0; i < qtot; ++i) fx[i] = E.tx * ft[i] + fx[i];
    for (int i = 0; i < qtot; ++i) fy[i] = E.ry * fr[i];
    for (int i = 0; i < qtot; ++i) fy[i] = E.sy * fs[i] + fy[i];
    if (nsd == 3) for (int i = 0; i < qtot; ++i) fy[i] = E.ty * ft[i] + fy[i];
    if (nsd == 3) {
      for (int i = 0; i < qtot; ++i) fz[i] = E.rz * fr[i];
      for (int i = 0; i < qtot; ++i) fz[i] = E.sz * fs[i] + fz[i];
      for (int i = 0; i < qtot; ++i) fz[i] = E.tz * ft[i] + fz[i];
    }
  }

  // inner product + mass inverse (forward transform), then backward transform
  void forward(int e, double *f, double *fh) const {
    const Elem &E = el[e];
    for (int i = 0; i < qtot; ++i) f[i] = E.wJ[i] * f[i];
    if (nsd == 2) {
      std::vector<double> wk(L * q);
      mm(L, q, q, B.data(), q, f, q, wk.data(), q);
      mm_nt(L, L, q, wk.data(), q, B.data(), q, fh, L);
    } else {
      const int q2 = q * q;
      std::vector<double> H(L * q2), Ha(L * L * q);
      mm(L, q2, q, B.data(), q, f, q2, H.data(), q2);
      for (int a = 0; a < L; ++a) mm(L, q, q, B.data(), q, H.data() + a * q2, q, Ha.data() + a * L * q, q);
      mm_nt(L * L, L, q, Ha.data(), q, B.data(), q, fh, L);
    }
    if (E.curved) { chol_solve(E.chol, fh); return; }
    const double iJ = 1.0 / E.J;
    for (int i = 0; i < nm; ++i) fh[i] *= iJ;
  }
  // M = L L^T (lower factor, row-major nm x nm): solve in place
  void chol_solve(const std::vector<double> &Lf, double *b) const {
    for (int i = 0; i < nm; ++i) {
      double t = b[i];
      for (int k = 0; k < i; ++k) t -= Lf[(size_t)i * nm + k] * b[k];
      b[i] = t / Lf[(size_t)i * nm + i];
    }
    for (int i = nm - 1; i >= 0; --i) {
      double t = b[i];
      for (int k = i + 1; k < nm; ++k) t -= Lf[(size_t)k * nm + i] * b[k];
      b[i] = t / Lf[(size_t)i * nm + i];
    }
  }
  // curved hex from nodal coordinates: Hex::compute_curved_geometry
  // (src/Hex.cpp:3992-4080) at the nodes, the side geometry from those
  // metrics, and the Cholesky factor of B^T diag(wJ) B for the projection
  void set_curved(int e, const double *x, const double *y, const double *z) {
    Elem &E = el[e];
    E.cx.assign(x, x + qtot); E.cy.assign(y, y + qtot); E.cz.assign(z, z + qtot);
    std::vector<double> xr(qtot), xs(qtot), xt(qtot), yr(qtot), ys(qtot), yt(qtot), zr(qtot), zs(qtot), zt(qtot);
    grad_r(x, xr.data(), xs.data(), xt.data());
    grad_r(y, yr.data(), ys.data(), yt.data());
    grad_r(z, zr.data(), zs.data(), zt.data());
    E.Jn.assign(qtot, 0.0);
    for (int c = 0; c < 9; ++c) E.mn[c].assign(qtot, 0.0);
    for (int i = 0; i < qtot; ++i) {
      double t1 = xr[i] * ys[i]; t1 = t1 * zt[i];
      double t2 = yr[i] * zs[i]; t1 = xt[i] * t2 + t1;
      t2 = zr[i] * xs[i]; t1 = yt[i] * t2 + t1;
      t2 = xt[i] * ys[i]; t1 = t1 - zr[i] * t2;
      t2 = zt[i] * xs[i]; t1 = t1 - yr[i] * t2;
      t2 = xr[i] * zs[i]; const double J = t1 - yt[i] * t2;
      if (!(J > 0.0)) status |= 4;
      E.Jn[i] = J;
      E.mn[0][i] = (ys[i] * zt[i] - yt[i] * zs[i]) / J;
      E.mn[1][i] = -(xs[i] * zt[i] - xt[i] * zs[i]) / J;
      E.mn[2][i] = (xs[i] * yt[i] - xt[i] * ys[i]) / J;
      E.mn[3][i] = -(yr[i] * zt[i] - yt[i] * zr[i]) / J;
      E.mn[4][i] = (xr[i] * zt[i] - xt[i] * zr[i]) / J;
      E.mn[5][i] = -(xr[i] * yt[i] - xt[i] * yr[i]) / J;
      E.mn[6][i] = (yr[i] * zs[i] - ys[i] * zr[i]) / J;
      E.mn[7][i] = -(xr[i] * zs[i] - xs[i] * zr[i]) / J;
      E.mn[8][i] = (xr[i] * ys[i] - xs[i] * yr[i]) / J;
    }
    E.wJ.assign(qtot, 0.0);
    for (int i = 0; i < q; ++i)
      for (int j = 0; j < q; ++j)
        for (int k = 0; k < q; ++k) {
          const int m = (i * q + j) * q + k;
          E.wJ[m] = w[i] * w[j] * w[k] * E.Jn[m];
        }
    E.curved = true;
    face_geometry3(E);
    mass_factor(E);
  }
  // Cholesky factor of the element mass matrix M[a][b] = sum_n phi_a(n) wJ(n)
  // phi_b(n), phi from B (L x q per direction) (src/Quad.cpp:263-300,
  // src/Hex.cpp:289-318 with DGM_CHOLESKY_MASS_MATRIX)
  void mass_factor(Elem &E) {
    std::vector<double> Phi((size_t)nm * qtot);
    if (nsd == 2) {
      for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b)
          for (int i = 0; i < q; ++i)
            for (int j = 0; j < q; ++j) Phi[(size_t)(a * L + b) * qtot + i * q + j] = B[a * q + i] * B[b * q + j];
    } else {
      for (int a = 0; a < L; ++a)
        for (int b = 0; b < L; ++b)
          for (int c = 0; c < L; ++c)
            for (int i = 0; i < q; ++i)
              for (int j = 0; j < q; ++j)
                for (int k = 0; k < q; ++k)
                  Phi[(size_t)((a * L + b) * L + c) * qtot + (i * q + j) * q + k] =
                      B[a * q + i] * B[b * q + j] * B[c * q + k];
    }
    std::vector<double> M((size_t)nm * nm, 0.0);
    for (int a = 0; a < nm; ++a)
      for (int b = 0; b <= a; ++b) {
        double t = 0.0;
        for (int n = 0; n < qtot; ++n) t += Phi[(size_t)a * qtot + n] * E.wJ[n] * Phi[(size_t)b * qtot + n];
        M[(size_t)a * nm + b] = t;
      }
    E.chol.assign((size_t)nm * nm, 0.0);
    for (int i = 0; i < nm; ++i)
      for (int j = 0; j <= i; ++j) {
        double t = M[(size_t)i * nm + j];
        for (int k = 0; k < j; ++k) t -= E.chol[(size_t)i * nm + k] * E.chol[(size_t)j * nm + k];
        if (i == j) {
          if (!(t > 0.0)) status |= 4;
          E.chol[(size_t)i * nm + i] = std::sqrt(t);
        } else {
          E.chol[(size_t)i * nm + j] = t / E.chol[(size_t)j * nm + j];
        }
      }
  }
  // curved quad from nodal coordinates: Quad::compute_curved_geometry
  // (src/Quad.cpp:1309-1363) at the nodes, the edge geometry from those metrics
  // (Quad::compute_side_geometry, src/Quad.cpp:1367-1452) and the mass matrix
  void set_curved2(int e, const double *x, const double *y) {
    Elem &E = el[e];
    E.cx.assign(x, x + qtot); E.cy.assign(y, y + qtot);
    std::vector<double> xr(qtot), xs(qtot), yr(qtot), ys(qtot), dummy(qtot);
    grad_r(x, xr.data(), xs.data(), dummy.data());
    grad_r(y, yr.data(), ys.data(), dummy.data());
    E.Jn.assign(qtot, 0.0);
    for (int c = 0; c < 9; ++c) E.mn[c].assign(qtot, 0.0);
    for (int i = 0; i < qtot; ++i) {
      double J = xs[i] * yr[i];
      J = -1.0 * J;
      J = xr[i] * ys[i] + J;
      if (!(J > 0.0)) status |= 4;
      E.Jn[i] = J;
      E.mn[0][i] = ys[i] / J;            // rx
      E.mn[1][i] = -1.0 * (xs[i] / J);   // ry
      E.mn[3][i] = -1.0 * (yr[i] / J);   // sx
      E.mn[4][i] = xr[i] / J;            // sy
    }
    E.wJ = E.Jn;
    for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) E.wJ[i * q + j] *= w[i];
    for (int i = 0; i < q; ++i) for (int j = 0; j < q; ++j) E.wJ[i * q + j] *= w[j];
    E.curved = true;
    std::vector<double> wkx(q), wky(q), tmp(q), dx(q), dy(q);
    for (int s = 0; s < 4; ++s) {
      SideInfo &S = E.side[s];
      S.nx.assign(qf, 0.0); S.ny.assign(qf, 0.0); S.J.assign(qf, 0.0);
      const int row = (s == 0 || s == 2) ? 3 : 0;
      get_side(E.mn[row].data(), s, wkx.data());
      get_side(E.mn[row + 1].data(), s, wky.data());
      to_side(wkx.data(), S.nx.data());
      to_side(wky.data(), S.ny.data());
      if (s == 0 || s == 3)
        for (int g = 0; g < qf; ++g) { S.nx[g] *= -1.0; S.ny[g] *= -1.0; }
      std::vector<double> iJ(qf);
      get_side(E.Jn.data(), s, tmp.data());
      to_side(tmp.data(), iJ.data());
      for (int g = 0; g < qf; ++g) {
        const double len = std::sqrt(S.nx[g] * S.nx[g] + S.ny[g] * S.ny[g]);
        S.nx[g] /= len; S.ny[g] /= len;
      }
      get_side(E.cx.data(), s, tmp.data());
      mm(q, 1, q, D.data(), q, tmp.data(), 1, dx.data(), 1);
      get_side(E.cy.data(), s, tmp.data());
      mm(q, 1, q, D.data(), q, tmp.data(), 1, dy.data(), 1);
      for (int i = 0; i < q; ++i) dx[i] = dx[i] * dx[i];
      for (int i = 0; i < q; ++i) dx[i] = dy[i] * dy[i] + dx[i];
      for (int i = 0; i < q; ++i) tmp[i] = std::sqrt(dx[i]);
      to_side(tmp.data(), S.J.data());
      for (int g = 0; g < qf; ++g) S.J[g] = S.J[g] / iJ[g];
    }
    mass_factor(E);
  }
  void backward(const double *fh, double *f) const {
    if (nsd == 2) {
      std::vector<double> wk(q * L);
      mm_tn(q, L, L, B.data(), q, fh, L, wk.data(), L);
      mm(q, q, L, wk.data(), L, B.data(), q, f, q);
    } else {
      const int L2 = L * L;
      std::vector<double> H(q * L2), Ha(q * q * L);
      mm_tn(q, L2, L, B.data(), q, fh, L2, H.data(), L2);
      for (int i = 0; i < q; ++i) mm_tn(q, L, L, B.data(), q, H.data() + i * L2, L, Ha.data() + i * q * L, L);
      mm(q * q, q, L, Ha.data(), L, B.data(), q, f, q);
    }
  }
  void project_vec(int e, double *f) const {
    std::vector<double> fh(nm);
    forward(e, f, fh.data());
    backward(fh.data(), f);
  }

  void fill_sides(VF &F, int v, int e, const double *f) {
    std::vector<double> wk(nsd == 2 ? q : q * q);
    for (int s = 0; s < nsides; ++s) {
      get_side(f, s, wk.data());
      to_side(wk.data(), es(F, v, e, s));
    }
  }
  void fill_sides_all(VF &F) {
    for (int v = 0; v < nv; ++v)
      for (int e = 0; e < ne; ++e) fill_sides(F, v, e, eu(F, v, e));
  }
  // normal-flux side fill (src/Quad.cpp:894-934, src/Hex.cpp:950-973)
  void side_normal_flux(int e, int s, const double *fx, const double *fy, const double *fz, double *su) {
    const SideInfo &S = el[e].side[s];
    std::vector<double> wk(nsd == 2 ? q : q * q), tmp(qs);
    get_side(fx, s, wk.data()); to_side(wk.data(), su);
    if (nsd == 2) {
      get_side(fy, s, wk.data()); to_side(wk.data(), tmp.data());
      for (int g = 0; g < qs; ++g) su[g] = S.nx[g] * su[g];
      for (int g = 0; g < qs; ++g) su[g] = S.ny[g] * tmp[g] + su[g];
    } else {
      for (int g = 0; g < qs; ++g) su[g] = S.nx[g] * su[g];
      get_side(fy, s, wk.data()); to_side(wk.data(), tmp.data());
      for (int g = 0; g < qs; ++g) su[g] = S.ny[g] * tmp[g] + su[g];
      get_side(fz, s, wk.data()); to_side(wk.data(), tmp.data());
      for (int g = 0; g < qs; ++g) su[g] = S.nz[g] * tmp[g] + su[g];
    }
  }
  void fill_sides_flux(VF &F, int v, int e, const double *fx, const double *fy, const double *fz) {
    for (int s = 0; s < nsides; ++s) side_normal_flux(e, s, fx, fy, fz, es(F, v, e, s));
  }

  // lift (src/Quad.cpp:1971-2038, src/Hex.cpp:1786-1952):
  //   in += G2GL[(F_side - U_side) * J_side/J_vol (* n_dir)] / w_boundary
  void lift(VF &Us, VF &Fs, int v, int e, double *in, char dir, int only = -1) {
    const Elem &E = el[e];
    std::vector<double> f(qs), fi(nsd == 2 ? q : q * q);
    for (int s = 0; s < nsides; ++s) {
      if (only >= 0 && s != only) continue;
      const SideInfo &S = E.side[s];
      const double *fu = es(Fs, v, e, s), *uu = es(Us, v, e, s);
      for (int g = 0; g < qs; ++g) f[g] = fu[g] - uu[g];
      for (int g = 0; g < qs; ++g) f[g] = S.J[g] * f[g];
      if (dir == 'x') for (int g = 0; g < qs; ++g) f[g] = S.nx[g] * f[g];
      else if (dir == 'y') for (int g = 0; g < qs; ++g) f[g] = S.ny[g] * f[g];
      else if (dir == 'z') for (int g = 0; g < qs; ++g) f[g] = S.nz[g] * f[g];
      from_side(f.data(), fi.data());
      if (nsd == 2) {
        double W;
        switch (s) {
          case 0: W = 1.0 / w[0];
            for (int i = 0; i < q; ++i) fi[i] = W * fi[i];
            for (int i = 0; i < q; ++i) in[i * q + 0] += fi[i];
            break;
          case 1: W = 1.0 / w[q - 1];
            for (int j = 0; j < q; ++j) fi[j] = W * fi[j];
            for (int j = 0; j < q; ++j) in[(q - 1) * q + j] += fi[j];
            break;
          case 2: W = 1.0 / w[q - 1];
            for (int i = 0; i < q; ++i) fi[i] = W * fi[i];
            for (int i = 0; i < q; ++i) in[i * q + (q - 1)] += fi[i];
            break;
          case 3: W = 1.0 / w[0];
            for (int j = 0; j < q; ++j) fi[j] = W * fi[j];
            for (int j = 0; j < q; ++j) in[0 * q + j] += fi[j];
            break;
        }
      } else {
        const int q2 = q * q;
        const double Wt[6] = {1.0 / w[0], 1.0 / w[0], 1.0 / w[q - 1], 1.0 / w[q - 1], 1.0 / w[0], 1.0 / w[q - 1]};
        const double W = Wt[s];
        switch (s) {
          case 0: for (int ij = 0; ij < q2; ++ij) in[ij * q + 0] = W * fi[ij] + in[ij * q + 0]; break;
          case 1:
            for (int i = 0; i < q; ++i)
              for (int k = 0; k < q; ++k)
                in[(i * q + 0) * q + k] = W * fi[i * q + k] + in[(i * q + 0) * q + k];
            break;
          case 2: for (int jk = 0; jk < q2; ++jk) in[(q - 1) * q2 + jk] = W * fi[jk] + in[(q - 1) * q2 + jk]; break;
          case 3:
            for (int i = 0; i < q; ++i)
              for (int k = 0; k < q; ++k)
                in[(i * q + q - 1) * q + k] = W * fi[i * q + k] + in[(i * q + q - 1) * q + k];
            break;
          case 4: for (int jk = 0; jk < q2; ++jk) in[jk] = W * fi[jk] + in[jk]; break;
          case 5: for (int ij = 0; ij < q2; ++ij) in[ij * q + q - 1] = W * fi[ij] + in[ij * q + q - 1]; break;
        }
      }
    }
  }

  // ---- physics ----
  // 2D convection (src/Euler.cpp:296-390)
  void convection2d(VF &Us, VF &Uf) {
    const double gm1 = gamma - 1.0, mpt5gm1 = -0.5 * gm1;
    fill_sides_all(Us);
    std::vector<double> u(qtot), v(qtot), p_(qtot), d(qtot), d1(qtot);
    for (int e = 0; e < ne; ++e) {
      const double *r = eu(Us, 0, e), *ru = eu(Us, 1, e), *rv = eu(Us, 2, e), *rE = eu(Us, 3, e);
      double *F0 = eu(Uf, 0, e), *F1 = eu(Uf, 1, e), *F2 = eu(Uf, 2, e), *F3 = eu(Uf, 3, e);
      for (int i = 0; i < qtot; ++i) { u[i] = ru[i] / r[i]; v[i] = rv[i] / r[i]; }
      for (int i = 0; i < qtot; ++i) {
        p_[i] = u[i] * ru[i]; d[i] = v[i] * rv[i]; p_[i] = p_[i] + d[i];
        p_[i] = mpt5gm1 * p_[i]; p_[i] = gm1 * rE[i] + p_[i];
      }
      gradient_dir(e, ru, F0, 'x');
      gradient_dir(e, rv, d.data(), 'y');
      for (int i = 0; i < qtot; ++i) F0[i] = F0[i] + d[i];
      for (int i = 0; i < qtot; ++i) { d[i] = ru[i] * u[i]; d[i] = d[i] + p_[i]; }
      gradient_dir(e, d.data(), F1, 'x');
      for (int i = 0; i < qtot; ++i) d[i] = ru[i] * v[i];
      gradient_dir(e, d.data(), d1.data(), 'y');
      for (int i = 0; i < qtot; ++i) F1[i] = F1[i] + d1[i];
      for (int i = 0; i < qtot; ++i) d[i] = rv[i] * u[i];
      gradient_dir(e, d.data(), F2, 'x');
      for (int i = 0; i < qtot; ++i) { d[i] = rv[i] * v[i]; d[i] = d[i] + p_[i]; }
      gradient_dir(e, d.data(), d1.data(), 'y');
      for (int i = 0; i < qtot; ++i) F2[i] = F2[i] + d1[i];
      for (int i = 0; i < qtot; ++i) p_[i] = p_[i] + rE[i];
      for (int i = 0; i < qtot; ++i) d[i] = u[i] * p_[i];
      gradient_dir(e, d.data(), F3, 'x');
      for (int i = 0; i < qtot; ++i) d[i] = v[i] * p_[i];
      gradient_dir(e, d.data(), d1.data(), 'y');
      for (int i = 0; i < qtot; ++i) F3[i] = F3[i] + d1[i];
    }
    convective_flux2d(Us, Uf);
    for (int e = 0; e < ne; ++e)
      for (int n = 0; n < nv; ++n) lift(Us, Uf, n, e, eu(Uf, n, e), 0);
  }

  int flux_type = 0;   // 0 Lax-Friedrichs, 1 Roe

  // Euler (2D) ghost values: Euler::WallBC::apply (src/EulerBC.hpp:43-62)
  // keeps the density, reflects the momentum onto the tangent and adjusts the
  // energy to the ghost's kinetic energy; StateBC (src/BC.cpp:17-28) holds
  // bc_state
  void ghost2d(int bc, const double *Ul, double nx, double ny, double *Ub) const {
    if (bc == 1) { for (int n = 0; n < 4; ++n) Ub[n] = bc_state[n]; return; }
    Ub[0] = Ul[0];
    Ub[1] = (Ul[1] * ny - Ul[2] * nx) * ny;
    Ub[2] = (Ul[2] * nx - Ul[1] * ny) * nx;
    Ub[3] = Ul[3] - 0.5 * (Ul[1] * Ul[1] + Ul[2] * Ul[2]) / Ul[0] + 0.5 * (Ub[1] * Ub[1] + Ub[2] * Ub[2]) / Ub[0];
  }

  // 2D master-side Riemann flux (src/Euler.cpp:399-757); boundary sides pair
  // with their ghost values
  void convective_flux2d(VF &Us, VF &Uf) {
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < nsides; ++s) {
        const SideInfo &S = el[e].side[s];
        if (S.bc >= 0) {
          if (S.bc > 1) status |= 2;   // Euler has wall and state boundaries here
          for (int g = 0; g < qs; ++g) {
            double Ul[4], Ub[4], Fpl[4], Fmr[4], Fh[4];
            for (int n = 0; n < 4; ++n) Ul[n] = es(Us, n, e, s)[g];
            ghost2d(S.bc, Ul, S.nx[g], S.ny[g], Ub);
            status |= point_flux2d(gamma, flux_type, Ul, Ub, S.nx[g], S.ny[g], Fpl, Fmr, Fh);
            for (int n = 0; n < 4; ++n) { es(Us, n, e, s)[g] = Fpl[n]; es(Uf, n, e, s)[g] = Fh[n]; }
          }
          continue;
        }
        if (!S.master) continue;
        for (int g = 0; g < qs; ++g) {
          const int ql = sidemap(S, g), qr = sidemap_slave(g);
          const double nxl = S.nx[ql], nyl = S.ny[ql];
          double Ul[4], Ur[4], Fpl[4], Fmr[4], Fh[4];
          for (int n = 0; n < 4; ++n) Ul[n] = es(Us, n, e, s)[ql];
          for (int n = 0; n < 4; ++n) Ur[n] = es(Us, n, S.le, S.ls)[qr];
          status |= point_flux2d(gamma, flux_type, Ul, Ur, nxl, nyl, Fpl, Fmr, Fh);
          for (int n = 0; n < 4; ++n) es(Us, n, e, s)[ql] = Fpl[n];
          for (int n = 0; n < 4; ++n) es(Us, n, S.le, S.ls)[qr] = Fmr[n];
          for (int n = 0; n < 4; ++n) {
            es(Us, n, S.le, S.ls)[qr] *= -1.0;
            es(Uf, n, e, s)[ql] = Fh[n];
            es(Uf, n, S.le, S.ls)[qr] = -Fh[n];
          }
        }
      }
  }

  // 3D convection (src/Euler3d.cpp:468-575)
  void convection3d(VF &Us, VF &Uf) {
    const double gm1 = gamma - 1.0, mpt5gm1 = -0.5 * gm1;
    fill_sides_all(Us);
    std::vector<double> u(qtot), v(qtot), wv(qtot), p_(qtot), d(qtot), d1(qtot);
    for (int e = 0; e < ne; ++e) {
      const double *r = eu(Us, 0, e), *ru = eu(Us, 1, e), *rv = eu(Us, 2, e), *rw = eu(Us, 3, e), *rE = eu(Us, 4, e);
      double *F0 = eu(Uf, 0, e), *F1 = eu(Uf, 1, e), *F2 = eu(Uf, 2, e), *F3 = eu(Uf, 3, e), *F4 = eu(Uf, 4, e);
      for (int i = 0; i < qtot; ++i) { u[i] = ru[i] / r[i]; v[i] = rv[i] / r[i]; wv[i] = rw[i] / r[i]; }
      for (int i = 0; i < qtot; ++i) {
        p_[i] = u[i] * ru[i]; p_[i] = v[i] * rv[i] + p_[i]; p_[i] = wv[i] * rw[i] + p_[i];
        p_[i] = mpt5gm1 * p_[i]; p_[i] = gm1 * rE[i] + p_[i];
      }
      gradient_dir(e, ru, F0, 'x');
      gradient_dir(e, rv, d.data(), 'y');
      for (int i = 0; i < qtot; ++i) F0[i] = F0[i] + d[i];
      gradient_dir(e, rw, d.data(), 'z');
      for (int i = 0; i < qtot; ++i) F0[i] = F0[i] + d[i];

      for (int i = 0; i < qtot; ++i) d[i] = ru[i] * u[i] + p_[i];
      gradient_dir(e, d.data(), F1, 'x');
      for (int i = 0; i < qtot; ++i) d[i] = ru[i] * v[i];
      gradient_dir(e, d.data(), d1.data(), 'y');
      for (int i = 0; i < qtot; ++i) F1[i] = F1[i] + d1[i];
      for (int i = 0; i < qtot; ++i) d[i] = ru[i] * wv[i];
      gradient_dir(e, d.data(), d1.data(), 'z');
      for (int i = 0; i < qtot; ++i) F1[i] = F1[i] + d1[i];

      for (int i = 0; i < qtot; ++i) d[i] = rv[i] * u[i];
      gradient_dir(e, d.data(), F2, 'x');
      for (int i = 0; i < qtot; ++i) d[i] = rv[i] * v[i] + p_[i];
      gradient_dir(e, d.data(), d1.data(), 'y');
      for (int i = 0; i < qtot; ++i) F2[i] = F2[i] + d1[i];
      for (int i = 0; i < qtot; ++i) d[i] = rv[i] * wv[i];
      gradient_dir(e, d.data(), d1.data(), 'z');
      for (int i = 0; i < qtot; ++i) F2[i] = F2[i] + d1[i];

      for (int i = 0; i < qtot; ++i) d[i] = rw[i] * u[i];
      gradient_dir(e, d.data(), F3, 'x');
      for (int i = 0; i < qtot; ++i) d[i] = rw[i] * v[i];
      gradient_dir(e, d.data(), d1.data(), 'y');
      for (int i = 0; i < qtot; ++i) F3[i] = F3[i] + d1[i];
      for (int i = 0; i < qtot; ++i) d[i] = rw[i] * wv[i] + p_[i];
      gradient_dir(e, d.data(), d1.data(), 'z');
      for (int i = 0; i < qtot; ++i) F3[i] = F3[i] + d1[i];

      for (int i = 0; i < qtot; ++i) p_[i] = p_[i] + rE[i];
      for (int i = 0; i < qtot; ++i) d[i] = u[i] * p_[i];
      gradient_dir(e, d.data(), F4, 'x');
      for (int i = 0; i < qtot; ++i) d[i] = v[i] * p_[i];
      gradient_dir(e, d.data(), d1.data(), 'y');
      for (int i = 0; i < qtot; ++i) F4[i] = F4[i] + d1[i];
      for (int i = 0; i < qtot; ++i) d[i] = wv[i] * p_[i];
      gradient_dir(e, d.data(), d1.data(), 'z');
      for (int i = 0; i < qtot; ++i) F4[i] = F4[i] + d1[i];
    }
    convective_flux3d(Us, Uf);
    for (int e = 0; e < ne; ++e)
      for (int n = 0; n < nv; ++n) lift(Us, Uf, n, e, eu(Uf, n, e), 0);
  }

  double bc_state[5] = {1.0, 0.0, 0.0, 0.0, 2.5};   // state boundary values (conservative)

  // ghost side values of a boundary side (Domain::enforce_bc, src/Domain.cpp:1346-1350):
  // slip wall Euler3d::WallBC::apply (src/Euler3d.cpp:876-909): the normal
  // momentum removed, internal energy kept; state StateBC::apply (src/BC.cpp:17-28)
  // isothermal wall Navier_Stokes3d::Isothermal_WallBC::apply (src/Navier_Stokes3d.cpp:472-492):
  // density kept, no velocity, energy (Eref - 1/2) rho Tw
  double iso_Tw = 1.0, iso_Eref = 0.0;
  void ghost3d(int bc, const double *Ul, double nx, double ny, double nz, double *Ub) const {
    if (bc == 1) { for (int n = 0; n < 5; ++n) Ub[n] = bc_state[n]; return; }
    if (bc == 3) {
      Ub[0] = Ul[0]; Ub[1] = 0.0; Ub[2] = 0.0; Ub[3] = 0.0;
      Ub[4] = (iso_Eref - 0.5) * Ub[0] * iso_Tw;
      return;
    }
    double up = Ul[1], vp = Ul[2], wp = Ul[3];
    double dot = up * nx + vp * ny + wp * nz;
    up = up - dot * nx; vp = vp - dot * ny; wp = wp - dot * nz;
    Ub[0] = Ul[0]; Ub[1] = up; Ub[2] = vp; Ub[3] = wp;
    Ub[4] = Ul[4] - 0.5 * (Ul[1] * Ul[1] + Ul[2] * Ul[2] + Ul[3] * Ul[3]) / Ul[0]
          + 0.5 * (Ub[1] * Ub[1] + Ub[2] * Ub[2] + Ub[3] * Ub[3]) / Ub[0];
  }

  // 3D master-side Lax-Friedrichs (or Roe) flux in the rotated frame (src/Euler3d.cpp:584-817);
  // boundary sides pair with their ghost values.  The inviscid ghost comes from
  // Domain::BC_db, which Euler3d::setup_BCdb fills (src/Euler3d.cpp:136, 824-868):
  // Euler3d::make_BC turns every code containing "wall" -- Isothermal_Wall
  // included -- into the slip Euler3d::WallBC (:857-860), and Euler3d::convection
  // applies that database (:476, :930-933).  The isothermal ghost lives in
  // Navier_Stokes3d's own BC_db (src/Navier_Stokes3d.hpp:83, .cpp:398-465) and
  // only the viscous terms see it (primitive_ghosts, average_flux).
  void convective_flux3d(VF &Us, VF &Uf) {
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < nsides; ++s) {
        const SideInfo &S = el[e].side[s];
        if (S.bc >= 0) {
          const int inviscid_bc = (S.bc == 3) ? 0 : S.bc;
          for (int g = 0; g < qs; ++g) {
            double Ul[5], Ub[5], Fpl[5], Fmr[5], Fh[5];
            for (int n = 0; n < 5; ++n) Ul[n] = es(Us, n, e, s)[g];
            ghost3d(inviscid_bc, Ul, S.nx[g], S.ny[g], S.nz[g], Ub);
            status |= point_flux3d(gamma, flux_type, Ul, Ub, S.nx[g], S.ny[g], S.nz[g], Fpl, Fmr, Fh);
            for (int n = 0; n < 5; ++n) { es(Us, n, e, s)[g] = Fpl[n]; es(Uf, n, e, s)[g] = Fh[n]; }
          }
          continue;
        }
        if (!S.master) continue;
        for (int g = 0; g < qs; ++g) {
          const int ql = sidemap(S, g), qr = sidemap_slave(g);
          const double nx = S.nx[ql], ny = S.ny[ql], nz = S.nz[ql];
          double Ul[5], Ur[5], Fpl[5], Fmr[5], Fh[5];
          for (int n = 0; n < 5; ++n) Ul[n] = es(Us, n, e, s)[ql];
          for (int n = 0; n < 5; ++n) Ur[n] = es(Us, n, S.le, S.ls)[qr];
          status |= point_flux3d(gamma, flux_type, Ul, Ur, nx, ny, nz, Fpl, Fmr, Fh);
          for (int n = 0; n < 5; ++n) es(Us, n, e, s)[ql] = Fpl[n];
          for (int n = 0; n < 5; ++n) es(Us, n, S.le, S.ls)[qr] = Fmr[n];
          for (int n = 0; n < 5; ++n) {
            es(Us, n, S.le, S.ls)[qr] *= -1.0;
            es(Uf, n, e, s)[ql] = Fh[n];
            es(Uf, n, S.le, S.ls)[qr] = -Fh[n];
          }
        }
      }
  }

  // average_flux (src/Domain.cpp:1883-1937)
  void average_flux(VF &Us, VF &Fs, char sign) {
    const double fact = (sign == '+') ? 1.0 : -1.0;
    for (int v = 0; v < nv; ++v)
      for (int e = 0; e < ne; ++e)
        for (int s = 0; s < nsides; ++s) {
          const SideInfo &S = el[e].side[s];
          if (S.bc >= 0 && nsd == 3) {   // boundary: ghost primitives ('+') or -S ('-')
            const double *ul = es(Us, v, e, s);
            double *fl = es(Fs, v, e, s);
            for (int g = 0; g < qs; ++g) {
              const double ub = (sign == '+') ? gprim[(((size_t)v * ne + e) * nsides + s) * qs + g] : -ul[g];
              fl[g] = 0.5 * (ul[g] + fact * ub);
            }
            continue;
          }
          if (!S.master) continue;
          const double *ul = es(Us, v, e, s), *ur = es(Us, v, S.le, S.ls);
          double *fl = es(Fs, v, e, s), *fr = es(Fs, v, S.le, S.ls);
          for (int g = 0; g < qs; ++g) {
            const int ql = sidemap(S, g), qr = sidemap_slave(g);
            fl[ql] = 0.5 * (ul[ql] + fact * ur[qr]);
            fr[qr] = fact * fl[ql];
          }
        }
  }

  // conservative -> primitive side values (src/Euler.cpp:1420-1456, src/Euler3d.cpp:1162-1199)
  // Boundary sides of Navier-Stokes 3D (src/Field.cpp:4451 links them to their
  // ghost side): the ghost's primitive values, set with the interior's
  // (Domain::enforce_bc then Euler3d::primitive_sides); the viscous flux's
  // ghost is -S (Isothermal_WallBC::apply_flux, StateBC::apply_flux,
  // src/Navier_Stokes3d.cpp:494-512, src/BC.cpp:31-40), so its '-' average is S.
  std::vector<double> gprim;   // [v][e][side][qs] ghost primitives of boundary sides
  void primitive_ghosts(VF &Us) {
    gprim.assign((size_t)nv * ne * nsides * qs, 0.0);
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < nsides; ++s) {
        const SideInfo &S = el[e].side[s];
        if (S.bc < 0) continue;
        if (S.bc == 0) status |= 2;   // Euler3d::WallBC has no viscous flux (src/Euler3d.cpp:911-913)
        for (int g = 0; g < qs; ++g) {
          double Ul[5], Ub[5];
          for (int n = 0; n < 5; ++n) Ul[n] = es(Us, n, e, s)[g];
          ghost3d(S.bc, Ul, S.nx[g], S.ny[g], S.nz[g], Ub);
          double u = Ub[1] / Ub[0], v = Ub[2] / Ub[0], w = Ub[3] / Ub[0], en = Ub[4] / Ub[0];
          double wk = u * u; wk = v * v + wk; wk = w * w + wk;
          en = -0.5 * wk + en;
          const double pr[5] = {Ub[0], u, v, w, en};
          for (int n = 0; n < 5; ++n) gprim[(((size_t)n * ne + e) * nsides + s) * qs + g] = pr[n];
        }
      }
  }
  void primitive_sides(VF &Us) {
    const int d = nsd;
    if (d == 3) primitive_ghosts(Us);
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < nsides; ++s) {
        const SideInfo &S = el[e].side[s];
        if (d == 3 && S.bc >= 0) {   // the interior side of a boundary
          double *r = es(Us, 0, e, s), *uu = es(Us, 1, e, s), *vv = es(Us, 2, e, s), *ww = es(Us, 3, e, s);
          double *en = es(Us, 4, e, s);
          for (int g = 0; g < qs; ++g) {
            uu[g] = uu[g] / r[g]; vv[g] = vv[g] / r[g]; ww[g] = ww[g] / r[g]; en[g] = en[g] / r[g];
            double wk = uu[g] * uu[g];
            wk = vv[g] * vv[g] + wk;
            wk = ww[g] * ww[g] + wk;
            en[g] = -0.5 * wk + en[g];
          }
          continue;
        }
        if (!S.master) continue;
        for (int side = 0; side < 2; ++side) {
          const int ee = side ? S.le : e, ss = side ? S.ls : s;
          double *r = es(Us, 0, ee, ss), *uu = es(Us, 1, ee, ss), *vv = es(Us, 2, ee, ss);
          double *ww = (d == 3) ? es(Us, 3, ee, ss) : nullptr, *en = es(Us, d + 1, ee, ss);
          for (int g = 0; g < qs; ++g) {
            uu[g] = uu[g] / r[g]; vv[g] = vv[g] / r[g];
            if (d == 3) ww[g] = ww[g] / r[g];
            en[g] = en[g] / r[g];
            double wk = uu[g] * uu[g];
            wk = vv[g] * vv[g] + wk;
            if (d == 3) wk = ww[g] * ww[g] + wk;
            en[g] = -0.5 * wk + en[g];
          }
        }
      }
  }

  // lifted, projected gradient (src/Euler.cpp:1406-1417, src/Euler3d.cpp:1144-1160)
  void lifted_gradient(VF &Us, VF &Fv, int v, int e, const double *f, double *fx, double *fy, double *fz) {
    gradient_all(e, f, fx, fy, fz);
    lift(Us, Fv, v, e, fx, 'x'); project_vec(e, fx);
    lift(Us, Fv, v, e, fy, 'y'); project_vec(e, fy);
    if (nsd == 3) { lift(Us, Fv, v, e, fz, 'z'); project_vec(e, fz); }
  }

  // 2D BR1 diffusion (src/Navier_Stokes.cpp:113-258); periodic meshes only
  void diffusion2d(VF &Us, VF &Uf, VF &Ur) {
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < nsides; ++s)
        if (el[e].side[s].bc >= 0) status |= 2;   // no viscous boundaries restated in 2D
    fill_sides_all(Us);
    primitive_sides(Us);
    average_flux(Us, Uf, '+');
    const int n = qtot;
    std::vector<double> u(n), v(n), e_(n), wk(n), ux(n), uy(n), vx(n), vy(n), t11(n), t12(n), t22(n);
    const double mu = vis;
    for (int k = 0; k < ne; ++k) {
      const double *r = eu(Us, 0, k), *ru = eu(Us, 1, k), *rv = eu(Us, 2, k), *rE = eu(Us, 3, k);
      for (int i = 0; i < n; ++i) { u[i] = ru[i] / r[i]; v[i] = rv[i] / r[i]; e_[i] = rE[i] / r[i]; }
      for (int i = 0; i < n; ++i) { wk[i] = u[i] * u[i]; wk[i] = v[i] * v[i] + wk[i]; e_[i] = -0.5 * wk[i] + e_[i]; }
      lifted_gradient(Us, Uf, 1, k, u.data(), ux.data(), uy.data(), nullptr);
      lifted_gradient(Us, Uf, 2, k, v.data(), vx.data(), vy.data(), nullptr);
      for (int i = 0; i < n; ++i) { t12[i] = uy[i] + vx[i]; t12[i] = mu * t12[i]; }
      for (int i = 0; i < n; ++i) { t11[i] = 2.0 * ux[i] - vy[i]; t11[i] = (2.0 / 3.0) * t11[i]; t11[i] = mu * t11[i]; }
      for (int i = 0; i < n; ++i) { t22[i] = 2.0 * vy[i] - ux[i]; t22[i] = (2.0 / 3.0) * t22[i]; t22[i] = mu * t22[i]; }
      double *F1 = eu(Uf, 1, k), *F2 = eu(Uf, 2, k), *F3 = eu(Uf, 3, k);
      fill_sides_flux(Us, 1, k, t11.data(), t12.data(), nullptr);
      gradient_dir(k, t11.data(), ux.data(), 'x');
      gradient_dir(k, t12.data(), wk.data(), 'y');
      for (int i = 0; i < n; ++i) F1[i] = ux[i] + wk[i];
      fill_sides_flux(Us, 2, k, t12.data(), t22.data(), nullptr);
      gradient_dir(k, t12.data(), ux.data(), 'x');
      gradient_dir(k, t22.data(), wk.data(), 'y');
      for (int i = 0; i < n; ++i) F2[i] = ux[i] + wk[i];
      for (int i = 0; i < n; ++i) {
        ux[i] = t11[i] * u[i]; ux[i] = t12[i] * v[i] + ux[i];
        uy[i] = t12[i] * u[i]; uy[i] = t22[i] * v[i] + uy[i];
      }
      lifted_gradient(Us, Uf, 3, k, e_.data(), vx.data(), vy.data(), nullptr);
      const double gp = gamma / Pr;
      for (int i = 0; i < n; ++i) { vx[i] = mu * vx[i]; ux[i] = gp * vx[i] + ux[i]; }
      for (int i = 0; i < n; ++i) { vy[i] = mu * vy[i]; uy[i] = gp * vy[i] + uy[i]; }
      fill_sides_flux(Us, 3, k, ux.data(), uy.data(), nullptr);
      gradient_dir(k, ux.data(), wk.data(), 'x');
      gradient_dir(k, uy.data(), ux.data(), 'y');
      for (int i = 0; i < n; ++i) F3[i] = wk[i] + ux[i];
    }
    average_flux(Us, Uf, '-');
    for (int nn = 1; nn < nv; ++nn)
      for (int k = 0; k < ne; ++k) lift(Us, Uf, nn, k, eu(Uf, nn, k), 0);
    for (int nn = 1; nn < nv; ++nn)
      for (int k = 0; k < ne; ++k) {
        double *a = eu(Uf, nn, k), *b = eu(Ur, nn, k);
        for (int i = 0; i < n; ++i) { a[i] = -1.0 * a[i]; b[i] = b[i] + a[i]; }
      }
  }

  int visc_scheme = 0;   // 0 BR1, 1 BR2
  double br2_eta = 6.0;  // BR2 penalty

  // BR2 normal viscous flux on every side of hex k (no reference counterpart:
  // the reference's diffusion is BR1 only).  On side s the stresses and heat
  // flux come from grad_h q + eta r_s(q) instead of the BR1 lifted gradient
  // grad_h q + sum_s' r_s'(q), where grad_h q is the projected element gradient
  // and r_s the projected lifting of side s's jump ({q} - q-) alone, both
  // built from the same gradient / lift / project steps as BR1
  // (src/Euler3d.cpp:1144-1160).  out[(eq - 1) * nsides + s][qs], eq = 1..4.
  void br2_sides3d(VF &Us, VF &Fv, int k, const double *u, const double *v, const double *w,
                   const double *en, std::vector<double> &out) {
    const int n = qtot;
    const double *prim[4] = {u, v, w, en};
    std::vector<double> bg(12 * n), rl(12 * n), G(12 * n);
    for (int c = 0; c < 4; ++c) {
      double *gx = &bg[(3 * c) * n], *gy = &bg[(3 * c + 1) * n], *gz = &bg[(3 * c + 2) * n];
      gradient_all(k, prim[c], gx, gy, gz);
      project_vec(k, gx); project_vec(k, gy); project_vec(k, gz);
    }
    const double mu = vis, gp = gamma / Pr;
    std::vector<double> t(6 * n), fx(n), fy(n), fz(n);
    out.assign(4 * nsides * qs, 0.0);
    for (int s = 0; s < nsides; ++s) {
      std::fill(rl.begin(), rl.end(), 0.0);
      for (int c = 0; c < 4; ++c) {
        const char dirs[3] = {'x', 'y', 'z'};
        for (int x = 0; x < 3; ++x) {
          double *r = &rl[(3 * c + x) * n];
          lift(Us, Fv, c + 1, k, r, dirs[x], s);
          project_vec(k, r);
        }
      }
      for (int i = 0; i < 12 * n; ++i) G[i] = br2_eta * rl[i] + bg[i];
      const double *ux = &G[0], *uy = &G[n], *uz = &G[2 * n], *vx = &G[3 * n], *vy = &G[4 * n],
                   *vz = &G[5 * n], *wx = &G[6 * n], *wy = &G[7 * n], *wz = &G[8 * n];
      const double *ex = &G[9 * n], *ey = &G[10 * n], *ez = &G[11 * n];
      double *t11 = &t[0], *t12 = &t[n], *t13 = &t[2 * n], *t22 = &t[3 * n], *t23 = &t[4 * n], *t33 = &t[5 * n];
      for (int i = 0; i < n; ++i) {
        t12[i] = mu * (uy[i] + vx[i]);
        t13[i] = mu * (uz[i] + wx[i]);
        t23[i] = mu * (vz[i] + wy[i]);
        t11[i] = mu * ((2.0 / 3.0) * (2.0 * ux[i] - vy[i] - wz[i]));
        t22[i] = mu * ((2.0 / 3.0) * (2.0 * vy[i] - ux[i] - wz[i]));
        t33[i] = mu * ((2.0 / 3.0) * (2.0 * wz[i] - ux[i] - vy[i]));
      }
      side_normal_flux(k, s, t11, t12, t13, &out[(0 * nsides + s) * qs]);
      side_normal_flux(k, s, t12, t22, t23, &out[(1 * nsides + s) * qs]);
      side_normal_flux(k, s, t13, t23, t33, &out[(2 * nsides + s) * qs]);
      for (int i = 0; i < n; ++i) {
        fx[i] = t11[i] * u[i] + t12[i] * v[i] + t13[i] * w[i] + gp * (mu * ex[i]);
        fy[i] = t12[i] * u[i] + t22[i] * v[i] + t23[i] * w[i] + gp * (mu * ey[i]);
        fz[i] = t13[i] * u[i] + t23[i] * v[i] + t33[i] * w[i] + gp * (mu * ez[i]);
      }
      side_normal_flux(k, s, fx.data(), fy.data(), fz.data(), &out[(3 * nsides + s) * qs]);
    }
  }

  // 3D BR1 diffusion (src/Navier_Stokes3d.cpp:182-391)
  void diffusion3d(VF &Us, VF &Uf, VF &Ur) {
    fill_sides_all(Us);
    primitive_sides(Us);
    average_flux(Us, Uf, '+');
    const int n = qtot;
    std::vector<double> u(n), v(n), w_(n), e_(n), wk(n);
    std::vector<double> ux(n), uy(n), uz(n), vx(n), vy(n), vz(n), wx(n), wy(n), wz(n);
    std::vector<double> t11(n), t12(n), t13(n), t22(n), t23(n), t33(n), br2;
    const double mu = vis;
    for (int k = 0; k < ne; ++k) {
      const double *r = eu(Us, 0, k), *ru = eu(Us, 1, k), *rv = eu(Us, 2, k), *rw = eu(Us, 3, k), *rE = eu(Us, 4, k);
      for (int i = 0; i < n; ++i) { u[i] = ru[i] / r[i]; v[i] = rv[i] / r[i]; w_[i] = rw[i] / r[i]; e_[i] = rE[i] / r[i]; }
      for (int i = 0; i < n; ++i) {
        wk[i] = u[i] * u[i]; wk[i] = v[i] * v[i] + wk[i]; wk[i] = w_[i] * w_[i] + wk[i];
        e_[i] = -0.5 * wk[i] + e_[i];
      }
      if (visc_scheme == 1) br2_sides3d(Us, Uf, k, u.data(), v.data(), w_.data(), e_.data(), br2);
      lifted_gradient(Us, Uf, 1, k, u.data(), ux.data(), uy.data(), uz.data());
      lifted_gradient(Us, Uf, 2, k, v.data(), vx.data(), vy.data(), vz.data());
      lifted_gradient(Us, Uf, 3, k, w_.data(), wx.data(), wy.data(), wz.data());
      for (int i = 0; i < n; ++i) {
        t12[i] = uy[i] + vx[i]; t12[i] = mu * t12[i];
        t13[i] = uz[i] + wx[i]; t13[i] = mu * t13[i];
        t23[i] = vz[i] + wy[i]; t23[i] = mu * t23[i];
        wk[i] = 2.0 * ux[i] - vy[i]; t11[i] = 1.0 * wk[i] - wz[i]; t11[i] = (2.0 / 3.0) * t11[i]; t11[i] = mu * t11[i];
        wk[i] = 2.0 * vy[i] - ux[i]; t22[i] = 1.0 * wk[i] - wz[i]; t22[i] = (2.0 / 3.0) * t22[i]; t22[i] = mu * t22[i];
        wk[i] = 2.0 * wz[i] - ux[i]; t33[i] = 1.0 * wk[i] - vy[i]; t33[i] = (2.0 / 3.0) * t33[i]; t33[i] = mu * t33[i];
      }
      double *F1 = eu(Uf, 1, k), *F2 = eu(Uf, 2, k), *F3 = eu(Uf, 3, k), *F4 = eu(Uf, 4, k);
      fill_sides_flux(Us, 1, k, t11.data(), t12.data(), t13.data());
      gradient_dir(k, t11.data(), ux.data(), 'x');
      gradient_dir(k, t12.data(), wk.data(), 'y');
      for (int i = 0; i < n; ++i) F1[i] = ux[i] + wk[i];
      gradient_dir(k, t13.data(), wk.data(), 'z');
      for (int i = 0; i < n; ++i) F1[i] = F1[i] + wk[i];
      fill_sides_flux(Us, 2, k, t12.data(), t22.data(), t23.data());
      gradient_dir(k, t12.data(), ux.data(), 'x');
      gradient_dir(k, t22.data(), wk.data(), 'y');
      for (int i = 0; i < n; ++i) F2[i] = ux[i] + wk[i];
      gradient_dir(k, t23.data(), wk.data(), 'z');
      for (int i = 0; i < n; ++i) F2[i] = F2[i] + wk[i];
      fill_sides_flux(Us, 3, k, t13.data(), t23.data(), t33.data());
      gradient_dir(k, t13.data(), ux.data(), 'x');
      gradient_dir(k, t23.data(), wk.data(), 'y');
      for (int i = 0; i < n; ++i) F3[i] = ux[i] + wk[i];
      gradient_dir(k, t33.data(), wk.data(), 'z');
      for (int i = 0; i < n; ++i) F3[i] = F3[i] + wk[i];
      for (int i = 0; i < n; ++i) {
        ux[i] = t11[i] * u[i]; ux[i] = t12[i] * v[i] + ux[i]; ux[i] = t13[i] * w_[i] + ux[i];
        uy[i] = t12[i] * u[i]; uy[i] = t22[i] * v[i] + uy[i]; uy[i] = t23[i] * w_[i] + uy[i];
        uz[i] = t13[i] * u[i]; uz[i] = t23[i] * v[i] + uz[i]; uz[i] = t33[i] * w_[i] + uz[i];
      }
      lifted_gradient(Us, Uf, 4, k, e_.data(), vx.data(), vy.data(), vz.data());
      const double gp = gamma / Pr;
      for (int i = 0; i < n; ++i) {
        vx[i] = mu * vx[i]; ux[i] = gp * vx[i] + ux[i];
        vy[i] = mu * vy[i]; uy[i] = gp * vy[i] + uy[i];
        vz[i] = mu * vz[i]; uz[i] = gp * vz[i] + uz[i];
      }
      fill_sides_flux(Us, 4, k, ux.data(), uy.data(), uz.data());
      gradient_dir(k, ux.data(), wk.data(), 'x');
      gradient_dir(k, uy.data(), ux.data(), 'y');
      for (int i = 0; i < n; ++i) F4[i] = wk[i] + ux[i];
      gradient_dir(k, uz.data(), wk.data(), 'z');
      for (int i = 0; i < n; ++i) F4[i] = F4[i] + wk[i];
      if (visc_scheme == 1)
        for (int eq = 0; eq < 4; ++eq)
          for (int s = 0; s < nsides; ++s)
            std::memcpy(es(Ub, eq + 1, k, s), &br2[(eq * nsides + s) * qs], qs * sizeof(double));
    }
    // face flux: average of the side fluxes (BR1: of the lifted-gradient flux
    // also used in the volume; BR2: of the per-side BR2 flux); the lift below
    // corrects the volume term's own trace F(G)- . n to it
    average_flux(visc_scheme == 1 ? Ub : Us, Uf, '-');
    for (int nn = 1; nn < nv; ++nn)
      for (int k = 0; k < ne; ++k) lift(Us, Uf, nn, k, eu(Uf, nn, k), 0);
    for (int nn = 1; nn < nv; ++nn)
      for (int k = 0; k < ne; ++k) {
        double *a = eu(Uf, nn, k), *b = eu(Ur, nn, k);
        for (int i = 0; i < n; ++i) { a[i] = -1.0 * a[i]; b[i] = b[i] + a[i]; }
      }
  }

  // Linear hyperbolic system on quads, dU/dt + Ax dU/dx + Ay dU/dy = 0, with the
  // Lax-Friedrichs flux 1/2 (An UL + An UR) + 1/2 lam (UL - UR): the
  // reference's LinEuler_quasi residual (src/LinEuler_quasi.cpp:151-495:
  // Jacobian-vector fluxes, gradient, Lax-Friedrichs with the largest wave
  // speed, lift, project) for a uniform mean flow, whose Jacobians Ax, Ay and
  // speed |u0 . n| + c0 the caller sets; with -Ax^T, -Ay^T it is the adjoint system.
  double linA[2][16] = {{0}}, lin_u0 = 0.0, lin_v0 = 0.0, lin_c0 = 0.0;
  void convection_lin2d(VF &Us, VF &Uf) {
    fill_sides_all(Us);
    std::vector<double> fx(qtot), fy(qtot), d(qtot);
    for (int e = 0; e < ne; ++e) {
      for (int i = 0; i < 4; ++i) {
        for (int q = 0; q < qtot; ++q) {
          double sx = 0.0, sy = 0.0;
          for (int j = 0; j < 4; ++j) {
            sx += linA[0][4 * i + j] * eu(Us, j, e)[q];
            sy += linA[1][4 * i + j] * eu(Us, j, e)[q];
          }
          fx[q] = sx; fy[q] = sy;
        }
        double *F = eu(Uf, i, e);
        gradient_dir(e, fx.data(), F, 'x');
        gradient_dir(e, fy.data(), d.data(), 'y');
        for (int q = 0; q < qtot; ++q) F[q] = F[q] + d[q];
      }
    }
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < nsides; ++s) {
        const SideInfo &S = el[e].side[s];
        if (!S.master) continue;
        for (int g = 0; g < qs; ++g) {
          const int ql = sidemap(S, g), qr = sidemap_slave(g);
          const double nx = S.nx[ql], ny = S.ny[ql];
          const double lin_lam = std::fabs(lin_u0 * nx + lin_v0 * ny) + lin_c0;
          double UL[4], UR[4], FL[4], FR[4];
          for (int n = 0; n < 4; ++n) { UL[n] = es(Us, n, e, s)[ql]; UR[n] = es(Us, n, S.le, S.ls)[qr]; }
          for (int i = 0; i < 4; ++i) {
            double a = 0.0, b = 0.0;
            for (int j = 0; j < 4; ++j) {
              const double An = nx * linA[0][4 * i + j] + ny * linA[1][4 * i + j];
              a += An * UL[j];
              b += An * UR[j];
            }
            FL[i] = a; FR[i] = b;
          }
          for (int n = 0; n < 4; ++n) {
            const double Fh = 0.5 * (FL[n] + FR[n] + lin_lam * (UL[n] - UR[n]));
            es(Us, n, e, s)[ql] = FL[n];
            es(Us, n, S.le, S.ls)[qr] = -FR[n];
            es(Uf, n, e, s)[ql] = Fh;
            es(Uf, n, S.le, S.ls)[qr] = -Fh;
          }
        }
      }
    for (int e = 0; e < ne; ++e)
      for (int n = 0; n < 4; ++n) lift(Us, Uf, n, e, eu(Uf, n, e), 0);
  }

  // ---- shallow water (src/Shallow.cpp) in (phi, uphi, vphi) ----
  // Function boundaries and sources (FuncBC, src/BC.cpp:95-200; FuncSource,
  // src/Source.cpp:11-75) are callbacks of (t, n, x, y) -> 3 x n values; the
  // balance term is the first residual, subtracted from every residual
  // (REO_SHALLOW_ENFORCE_BALANCE, src/Shallow.cpp:251-266).
  typedef void (*field_fn)(double t, int n, const double *x, const double *y, double *out);
  field_fn bc_fn = nullptr, src_fn = nullptr;
  double t = 0.0;
  bool balance = true, have_bal = false;
  std::vector<double> Usrc;

  // Gauss-point coordinates of side s of element e (Element::get_side_coordinates)
  void side_coords(int e, int s, double *x, double *y) const {
    std::vector<double> wk(q);
    get_side(el[e].cx.data(), s, wk.data()); to_side(wk.data(), x);
    get_side(el[e].cy.data(), s, wk.data()); to_side(wk.data(), y);
  }
  // ghost values of a shallow-water boundary side: Shallow::WallBC::apply
  // (src/ShallowBC.hpp:45-62) keeps phi and the tangential momentum; state
  // boundaries hold bc_state; function boundaries the callback's values
  void ghost_shallow(int e, int s, const SideInfo &S, VF &Us, std::vector<double> &Ub) {
    Ub.assign(3 * qs, 0.0);
    if (S.bc == 0) {
      for (int g = 0; g < qs; ++g) {
        const double m1 = es(Us, 1, e, s)[g], m2 = es(Us, 2, e, s)[g], nx = S.nx[g], ny = S.ny[g];
        Ub[g] = es(Us, 0, e, s)[g];
        Ub[qs + g] = (m1 * ny - m2 * nx) * ny;
        Ub[2 * qs + g] = (m2 * nx - m1 * ny) * nx;
      }
    } else if (S.bc == 1) {
      for (int n = 0; n < 3; ++n) for (int g = 0; g < qs; ++g) Ub[n * qs + g] = bc_state[n];
    } else {
      std::vector<double> x(qs), y(qs);
      side_coords(e, s, x.data(), y.data());
      if (bc_fn) bc_fn(t, qs, x.data(), y.data(), Ub.data());
    }
  }
  // Shallow::convection (src/Shallow.cpp:276-336): flux divergence at the
  // quadrature points, master-side flux, lift
  void convection_shallow2d(VF &Us, VF &Uf) {
    fill_sides_all(Us);
    std::vector<double> u(qtot), v(qtot), p_(qtot), d(qtot), d1(qtot);
    for (int e = 0; e < ne; ++e) {
      const double *phi = eu(Us, 0, e), *uphi = eu(Us, 1, e), *vphi = eu(Us, 2, e);
      double *F0 = eu(Uf, 0, e), *F1 = eu(Uf, 1, e), *F2 = eu(Uf, 2, e);
      for (int i = 0; i < qtot; ++i) { u[i] = uphi[i] / phi[i]; v[i] = vphi[i] / phi[i]; }
      for (int i = 0; i < qtot; ++i) { p_[i] = phi[i] * phi[i]; p_[i] = 0.5 * p_[i]; }
      gradient_dir(e, uphi, F0, 'x');
      gradient_dir(e, vphi, d.data(), 'y');
      for (int i = 0; i < qtot; ++i) F0[i] = d[i] + F0[i];
      for (int i = 0; i < qtot; ++i) { d[i] = uphi[i] * u[i]; d[i] = d[i] + p_[i]; }
      gradient_dir(e, d.data(), F1, 'x');
      for (int i = 0; i < qtot; ++i) d[i] = uphi[i] * v[i];
      gradient_dir(e, d.data(), d1.data(), 'y');
      for (int i = 0; i < qtot; ++i) F1[i] = d1[i] + F1[i];
      for (int i = 0; i < qtot; ++i) d[i] = vphi[i] * u[i];
      gradient_dir(e, d.data(), F2, 'x');
      for (int i = 0; i < qtot; ++i) { d[i] = vphi[i] * v[i]; d[i] = d[i] + p_[i]; }
      gradient_dir(e, d.data(), d1.data(), 'y');
      for (int i = 0; i < qtot; ++i) F2[i] = d1[i] + F2[i];
    }
    std::vector<double> Ub;
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < nsides; ++s) {
        const SideInfo &S = el[e].side[s];
        if (S.bc >= 0) {
          ghost_shallow(e, s, S, Us, Ub);
          for (int g = 0; g < qs; ++g) {
            double Ul[3], Ur[3], Fpl[3], Fmr[3], Fh[3];
            for (int n = 0; n < 3; ++n) { Ul[n] = es(Us, n, e, s)[g]; Ur[n] = Ub[n * qs + g]; }
            status |= point_flux_shallow(Ul, Ur, S.nx[g], S.ny[g], Fpl, Fmr, Fh);
            for (int n = 0; n < 3; ++n) { es(Us, n, e, s)[g] = Fpl[n]; es(Uf, n, e, s)[g] = Fh[n]; }
          }
          continue;
        }
        if (!S.master) continue;
        for (int g = 0; g < qs; ++g) {
          const int ql = sidemap(S, g), qr = sidemap_slave(g);
          double Ul[3], Ur[3], Fpl[3], Fmr[3], Fh[3];
          for (int n = 0; n < 3; ++n) { Ul[n] = es(Us, n, e, s)[ql]; Ur[n] = es(Us, n, S.le, S.ls)[qr]; }
          status |= point_flux_shallow(Ul, Ur, S.nx[ql], S.ny[ql], Fpl, Fmr, Fh);
          for (int n = 0; n < 3; ++n) {
            es(Us, n, e, s)[ql] = Fpl[n];
            es(Us, n, S.le, S.ls)[qr] = -Fmr[n];
            es(Uf, n, e, s)[ql] = Fh[n];
            es(Uf, n, S.le, S.ls)[qr] = -Fh[n];
          }
        }
      }
    for (int e = 0; e < ne; ++e)
      for (int n = 0; n < 3; ++n) lift(Us, Uf, n, e, eu(Uf, n, e), 0);
  }
  // FuncSource::apply (src/Source.cpp:11-24): Uf -= s at the quadrature points
  void add_source_shallow(VF &Uf) {
    if (!src_fn) return;
    std::vector<double> s(3 * qtot);
    for (int e = 0; e < ne; ++e) {
      src_fn(t, qtot, el[e].cx.data(), el[e].cy.data(), s.data());
      for (int n = 0; n < 3; ++n) {
        double *f = eu(Uf, n, e);
        for (int i = 0; i < qtot; ++i) f[i] = f[i] - s[n * qtot + i];
      }
    }
  }

  void project_all(VF &F) {
    for (int v = 0; v < nv; ++v)
      for (int e = 0; e < ne; ++e) project_vec(e, eu(F, v, e));
  }

  // time_derivative: F = -dU/dt (src/Euler.cpp:282-287, src/Euler3d.cpp:455-458,
  // src/Navier_Stokes.cpp:89-94, src/Navier_Stokes3d.cpp:135-161)
  void time_derivative(VF &Us, VF &Uf) {
    if (eqn == LINEULER2D) { convection_lin2d(Us, Uf); project_all(Uf); return; }
    if (eqn == SHALLOW2D) {   // Shallow::time_derivative (src/Shallow.cpp:255-266)
      convection_shallow2d(Us, Uf);
      add_source_shallow(Uf);
      project_all(Uf);
      if (balance) {
        if (!have_bal) { Usrc = Uf.u; have_bal = true; }
        for (size_t i = 0; i < Uf.u.size(); ++i) Uf.u[i] = -1.0 * Usrc[i] + Uf.u[i];
      }
      return;
    }
    if (nsd == 2) convection2d(Us, Uf); else convection3d(Us, Uf);
    if (eqn == NS2D) diffusion2d(Us, Uv, Uf);
    if (eqn == NS3D) diffusion3d(Us, Uv, Uf);
    project_all(Uf);
    if (channel) hold_bulk(Uf);
  }

  // Euler::set_time_step (src/Euler.cpp:804-877) and Shallow::set_time_step
  // (src/Shallow.cpp:573-637) with the stability metrics of
  // Element::compute_metrics_2d_tensor_prod (src/Element.cpp:485-555): the
  // largest idt over the GLL nodes of the current state, then the step rule.
  // Hexes: Euler3d::set_time_step (src/Euler3d.cpp:462-465) returns dt.
  // out = {cfl, new dt}
  // largest idt over the GLL nodes of element e for the nodal state r, ru, rv, rE
  double cfl_element(int e, const double *r, const double *ru, const double *rv, const double *rE) {
    std::vector<double> xr(qtot), xs(qtot), yr(qtot), ys(qtot);
    {
      const Elem &E = el[e];
      // local_gradient(C->x, xr, xs), local_gradient(C->y, yr, ys)
      for (int i = 0; i < q; ++i)
        for (int j = 0; j < q; ++j) {
          double a0 = 0, a1 = 0, c0 = 0, c1 = 0;
          for (int m = 0; m < q; ++m) {
            a0 += D[i * q + m] * E.cx[m * q + j]; a1 += D[j * q + m] * E.cx[i * q + m];
            c0 += D[i * q + m] * E.cy[m * q + j]; c1 += D[j * q + m] * E.cy[i * q + m];
          }
          xr[i * q + j] = a0; xs[i * q + j] = a1; yr[i * q + j] = c0; ys[i * q + j] = c1;
        }
      double emax = 0.0;
      for (int i = 0; i < q; ++i) {
        const double dr = i == 0 ? z[i + 1] - z[i] : (i == q - 1 ? z[i] - z[i - 1] : 0.5 * (z[i + 1] - z[i - 1]));
        for (int j = 0; j < q; ++j) {
          const double ds = j == 0 ? z[j + 1] - z[j] : (j == q - 1 ? z[j] - z[j - 1] : 0.5 * (z[j + 1] - z[j - 1]));
          const int n = j + i * q;
          const double Jn = E.curved ? xr[n] * ys[n] - xs[n] * yr[n] : E.J;
          const double dxr = xr[n] * dr, dxs = xs[n] * ds, dyr = yr[n] * dr, dys = ys[n] * ds;
          const double ivol = 1.0 / (dr * ds * Jn);
          const double acoustic = std::sqrt((dxr + dyr) * (dxr + dyr) + (dxs + dys) * (dxs + dys));
          const double u = ru[n] / r[n], v = rv[n] / r[n];
          double c;
          if (eqn == SHALLOW2D) {
            c = std::sqrt(1.0 * r[n]);
          } else {   // c = sqrt(gamma (gamma-1) (rE - 1/2 (r u u + r v v)) / r)
            c = u * ru[n];
            c = v * rv[n] + c;
            c = -0.5 * c;
            c = rE[n] + c;
            c = gamma * (gamma - 1.0) * c;
            c = c / r[n];
            c = std::sqrt(c);
          }
          double idt = u * dxr;
          idt = v * dyr + idt;
          idt = std::fabs(idt);
          double wk = u * dxs;
          wk = v * dys + wk;
          wk = std::fabs(wk);
          idt = wk + idt;
          idt = c * acoustic + idt;
          idt = ivol * idt;
          if (!(idt == idt)) status |= 1;
          emax = std::max(emax, idt);
        }
      }
      return emax;
    }
  }
  void set_time_step(double dt, double CFLmax, double dtmax, double *out) {
    if (nsd == 3) { out[0] = dt; out[1] = dt; return; }
    double cfl = 0.0;
    for (int e = 0; e < ne; ++e) {
      const double *rE = eqn == SHALLOW2D ? nullptr : eu(U, 3, e);
      cfl = std::max(cfl, dt * cfl_element(e, eu(U, 0, e), eu(U, 1, e), eu(U, 2, e), rE));
    }
    double dn = dt;
    if (CFLmax != 0 && dtmax != 0) dn = std::min(dtmax, CFLmax * dt / cfl);
    else if (CFLmax != 0) dn = CFLmax * dt / cfl;
    out[0] = cfl * dn / dt;
    out[1] = dn;
  }

  // Channel::holdBulkMassFlow (src/Channel.hpp:117-163): subtract the domain
  // means of the density, x-momentum and energy residuals
  bool channel = false;
  void hold_bulk(VF &Uf) {
    const int fields[3] = {0, 1, nv - 1};
    double mean[3] = {0, 0, 0}, vol = 0.0;
    for (int e = 0; e < ne; ++e) {
      double lv = 0.0;
      for (int i = 0; i < qtot; ++i) lv += el[e].wJ[i];
      for (int k = 0; k < 3; ++k) {
        const double *f = eu(Uf, fields[k], e);
        double a = 0.0;
        for (int i = 0; i < qtot; ++i) a += el[e].wJ[i] * f[i];
        mean[k] += a;
      }
      vol += lv;
    }
    for (int k = 0; k < 3; ++k) {
      mean[k] /= vol;
      for (int e = 0; e < ne; ++e) {
        double *f = eu(Uf, fields[k], e);
        for (int i = 0; i < qtot; ++i) f[i] -= mean[k];
      }
    }
  }

  // vField BLAS-1 (src/vField.hpp:1002-1031) on element data
  static void axpy(double a, const VF &X, const VF &Y, VF &Z) {
    for (size_t i = 0; i < X.u.size(); ++i) Z.u[i] = a * X.u[i] + Y.u[i];
  }
  static void axpy2(double a, const VF &X, VF &Y) {
    for (size_t i = 0; i < X.u.size(); ++i) Y.u[i] = a * X.u[i] + Y.u[i];
  }
  static void scal(double a, VF &X) {
    for (size_t i = 0; i < X.u.size(); ++i) X.u[i] = a * X.u[i];
  }
  static void axpby(double a, const VF &X, double b, const VF &Y, VF &Z) {
    for (size_t i = 0; i < X.u.size(); ++i) Z.u[i] = a * X.u[i] + b * Y.u[i];
  }

  // RungeKutta::advance (src/TimeInt.cpp:267-302)
  void rk4(double dt) {
    static const double a[3] = {0.5, 0.5, 1.0};
    static const double b[4] = {0.16666666666666667, 0.33333333333333333, 0.33333333333333333, 0.16666666666666667};
    time_derivative(U, R0);
    axpy(-a[0] * dt, R0, U, Ut);
    scal(-b[0], R0);
    t += 0.5 * dt;
    time_derivative(Ut, R1);
    axpy(-a[1] * dt, R1, U, Ut);
    axpy2(-b[1], R1, R0);
    time_derivative(Ut, R1);
    axpy(-a[2] * dt, R1, U, Ut);
    axpy2(-b[2], R1, R0);
    t += 0.5 * dt;
    time_derivative(Ut, R1);
    axpy2(-b[3], R1, R0);
    axpy2(dt, R0, U);
  }
  // TVD::advance (src/TimeInt.cpp:824-856); stage times c = {0, 1, 1/2} (src/TimeInt.cpp:740)
  void tvd(double dt) {
    static const double a[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
    static const double b[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
    const double tn = t;
    time_derivative(U, R0);
    axpy(-dt, R0, U, Ut);
    t = tn + dt;
    time_derivative(Ut, R0);
    axpy(-dt, R0, Ut, R0);
    axpby(a[1], U, b[1], R0, Ut);
    t = tn + 0.5 * dt;
    time_derivative(Ut, R0);
    axpy(-dt, R0, Ut, R0);
    axpby(b[2], R0, a[2], U, U);
    t = tn + dt;
  }
  // LSRK14::advance (src/TimeInt.cpp:1989-2013); coefficients from
  // src/TimeInt.cpp:1870-1917
  void lsrk14(double dt) {
    const double tn = t;
    for (int s = 0; s < 14; ++s) {
      t = tn + kLsrkC[s] * dt;
      time_derivative(U, R0);
      axpby(-dt, R0, kLsrkA[s], Ut, Ut);
      axpy2(kLsrkB[s], Ut, U);
    }
    t = tn + dt;
  }
};

// Variable-order quads (runs/vortex): 2D Euler with an order per element, as
// the reference builds it from root.ord (Field::Field, src/Field.cpp:2290-2320).
// Every element keeps its own basis (q_e = p_e + 2 GLL points); the two sides
// of a link share max(L_left, L_right) Gauss points (Side::match,
// src/Side.cpp:106-117, called by Field::adjust_side_order,
// src/Field.cpp:3208-3230).  fill_sides interpolates the element's GLL edge
// values to the side's Gauss points (Quad::fill_sides, src/Quad.cpp:845-869:
// GL2G of size qs x q_e); the lift interpolates back with G2GL (q_e x qs) and
// scales by the element's end weight (Quad::local_lift, src/Quad.cpp:1971-2038:
// the reference's default build has no DGM_SIDE_FLUX_TO_VOLUME, so var_p_lift
// is never taken).  Affine quads; the per-element work reuses one uniform-order
// Domain per order present (same mesh, same links).
struct VarDomain2 {
  int ne = 0, nv = 4, pmax = 0, Lmax = 0, qmax = 0;
  double gamma = 1.4;
  int status = 0;
  std::vector<int> pe;                           // order of each element
  std::vector<Domain *> dom;                     // [p] uniform-order domain (nullptr if absent)
  std::vector<size_t> off;                       // nodal offset of element e within a field
  size_t ntot = 0;                               // nodes per field
  std::vector<int> qs;                           // [e*4 + s] Gauss points of the side
  std::vector<std::vector<double>> gl2g, g2gl;   // [e*4 + s] GL2G (qs x q_e), G2GL (q_e x qs)
  VF U, Ut, R0, R1, Sf;                          // flat nodal [v][ntot]; sides in .s of Us/Uf
  double t = 0.0;

  ~VarDomain2() { for (Domain *d : dom) delete d; }
  Domain &D(int e) const { return *dom[pe[e]]; }
  int q_of(int e) const { return pe[e] + 2; }
  double *nod(VF &F, int v, int e) { return &F.u[(size_t)v * ntot + off[e]]; }
  double *sid(VF &F, int v, int e, int s) { return &F.s[(((size_t)v * ne + e) * 4 + s) * Lmax]; }
  void alloc(VF &F) {
    F.u.assign((size_t)nv * ntot, 0.0);
    F.s.assign((size_t)nv * ne * 4 * Lmax, 0.0);
  }

  void build(int ne_, const int *orders, const double *vert, int ncon, const int *con, double g, int flux) {
    ne = ne_; gamma = g;
    pe.assign(orders, orders + ne);
    pmax = *std::max_element(pe.begin(), pe.end());
    Lmax = pmax + 1; qmax = pmax + 2;
    dom.assign(pmax + 1, nullptr);
    for (int e = 0; e < ne; ++e) {
      const int p = pe[e];
      if (dom[p]) continue;
      Domain *d = new Domain();
      d->nsd = 2; d->ne = ne; d->p = p; d->eqn = 1; d->gamma = g; d->vis = 0.0; d->Pr = 0.72;
      d->flux_type = flux; d->nv = 4;
      d->build_basis();
      d->build_geometry(vert);
      d->build_links(ncon, con, vert);
      dom[p] = d;
    }
    off.assign(ne, 0);
    ntot = 0;
    for (int e = 0; e < ne; ++e) { off[e] = ntot; ntot += (size_t)q_of(e) * q_of(e); }
    // side point counts and interpolation matrices
    qs.assign(4 * (size_t)ne, 0);
    gl2g.assign(4 * (size_t)ne, {});
    g2gl.assign(4 * (size_t)ne, {});
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < 4; ++s) {
        const SideInfo &S = D(e).el[e].side[s];
        int n = pe[e] + 1;
        if (S.bc < 0 && S.le >= 0) n = std::max(n, pe[S.le] + 1);
        qs[4 * e + s] = n;
        std::vector<double> zg, wg;
        zwgl(n, zg, wg);
        interp_matrix(D(e).z, zg, gl2g[4 * e + s]);   // n x q_e
        interp_matrix(zg, D(e).z, g2gl[4 * e + s]);   // q_e x n
      }
    alloc(U); alloc(Ut); alloc(R0); alloc(R1); alloc(Sf);
  }

  void fill_sides(VF &F) {
    std::vector<double> wk(qmax);
    for (int v = 0; v < nv; ++v)
      for (int e = 0; e < ne; ++e) {
        const int q = q_of(e);
        for (int s = 0; s < 4; ++s) {
          D(e).get_side(nod(F, v, e), s, wk.data());
          mm(qs[4 * e + s], 1, q, gl2g[4 * e + s].data(), q, wk.data(), 1, sid(F, v, e, s), 1);
        }
      }
  }

  // Euler::convection volume term of one element (src/Euler.cpp:296-390)
  void volume(VF &Us, VF &Uf, int e) {
    const Domain &Dm = D(e);
    const int qt = q_of(e) * q_of(e);
    const double gm1 = gamma - 1.0, mpt5gm1 = -0.5 * gm1;
    std::vector<double> u(qt), v(qt), p_(qt), d(qt), d1(qt);
    const double *r = nod(Us, 0, e), *ru = nod(Us, 1, e), *rv = nod(Us, 2, e), *rE = nod(Us, 3, e);
    double *F0 = nod(Uf, 0, e), *F1 = nod(Uf, 1, e), *F2 = nod(Uf, 2, e), *F3 = nod(Uf, 3, e);
    for (int i = 0; i < qt; ++i) { u[i] = ru[i] / r[i]; v[i] = rv[i] / r[i]; }
    for (int i = 0; i < qt; ++i) {
      p_[i] = u[i] * ru[i]; d[i] = v[i] * rv[i]; p_[i] = p_[i] + d[i];
      p_[i] = mpt5gm1 * p_[i]; p_[i] = gm1 * rE[i] + p_[i];
    }
    Dm.gradient_dir(e, ru, F0, 'x');
    Dm.gradient_dir(e, rv, d.data(), 'y');
    for (int i = 0; i < qt; ++i) F0[i] = F0[i] + d[i];
    for (int i = 0; i < qt; ++i) { d[i] = ru[i] * u[i]; d[i] = d[i] + p_[i]; }
    Dm.gradient_dir(e, d.data(), F1, 'x');
    for (int i = 0; i < qt; ++i) d[i] = ru[i] * v[i];
    Dm.gradient_dir(e, d.data(), d1.data(), 'y');
    for (int i = 0; i < qt; ++i) F1[i] = F1[i] + d1[i];
    for (int i = 0; i < qt; ++i) d[i] = rv[i] * u[i];
    Dm.gradient_dir(e, d.data(), F2, 'x');
    for (int i = 0; i < qt; ++i) { d[i] = rv[i] * v[i]; d[i] = d[i] + p_[i]; }
    Dm.gradient_dir(e, d.data(), d1.data(), 'y');
    for (int i = 0; i < qt; ++i) F2[i] = F2[i] + d1[i];
    for (int i = 0; i < qt; ++i) p_[i] = p_[i] + rE[i];
    for (int i = 0; i < qt; ++i) d[i] = u[i] * p_[i];
    Dm.gradient_dir(e, d.data(), F3, 'x');
    for (int i = 0; i < qt; ++i) d[i] = v[i] * p_[i];
    Dm.gradient_dir(e, d.data(), d1.data(), 'y');
    for (int i = 0; i < qt; ++i) F3[i] = F3[i] + d1[i];
  }

  // master-side Riemann flux on the shared Gauss points (src/Euler.cpp:399-757);
  // a flipped side reverses the order of its qs points (sideMap,
  // src/BaseQuad.hpp:53-57)
  void faces(VF &Us, VF &Uf) {
    for (int e = 0; e < ne; ++e)
      for (int s = 0; s < 4; ++s) {
        const SideInfo &S = D(e).el[e].side[s];
        const int n = qs[4 * e + s];
        if (S.bc >= 0) {
          if (S.bc > 1) status |= 2;
          for (int g = 0; g < n; ++g) {
            double Ul[4], Ub[4], Fpl[4], Fmr[4], Fh[4];
            for (int c = 0; c < 4; ++c) Ul[c] = sid(Us, c, e, s)[g];
            D(e).ghost2d(S.bc, Ul, S.nx[0], S.ny[0], Ub);
            status |= point_flux2d(gamma, D(e).flux_type, Ul, Ub, S.nx[0], S.ny[0], Fpl, Fmr, Fh);
            for (int c = 0; c < 4; ++c) { sid(Us, c, e, s)[g] = Fpl[c]; sid(Uf, c, e, s)[g] = Fh[c]; }
          }
          continue;
        }
        if (!S.master) continue;
        for (int g = 0; g < n; ++g) {
          const int ql = S.flip_a ? (n - 1) - g : g, qr = g;
          const double nxl = S.nx[0], nyl = S.ny[0];   // affine: constant along the side
          double Ul[4], Ur[4], Fpl[4], Fmr[4], Fh[4];
          for (int c = 0; c < 4; ++c) Ul[c] = sid(Us, c, e, s)[ql];
          for (int c = 0; c < 4; ++c) Ur[c] = sid(Us, c, S.le, S.ls)[qr];
          status |= point_flux2d(gamma, D(e).flux_type, Ul, Ur, nxl, nyl, Fpl, Fmr, Fh);
          for (int c = 0; c < 4; ++c) sid(Us, c, e, s)[ql] = Fpl[c];
          for (int c = 0; c < 4; ++c) sid(Us, c, S.le, S.ls)[qr] = Fmr[c];
          for (int c = 0; c < 4; ++c) {
            sid(Us, c, S.le, S.ls)[qr] *= -1.0;
            sid(Uf, c, e, s)[ql] = Fh[c];
            sid(Uf, c, S.le, S.ls)[qr] = -Fh[c];
          }
        }
      }
  }

  // Quad::lift / local_lift with G2GL (q_e x qs) (src/Quad.cpp:1971-2038)
  void lift(VF &Us, VF &Uf, int v, int e) {
    const Domain &Dm = D(e);
    const int q = q_of(e);
    double *in = nod(Uf, v, e);
    std::vector<double> f(Lmax), fi(q);
    for (int s = 0; s < 4; ++s) {
      const SideInfo &S = Dm.el[e].side[s];
      const int n = qs[4 * e + s];
      const double *fu = sid(Uf, v, e, s), *uu = sid(Us, v, e, s);
      for (int g = 0; g < n; ++g) f[g] = fu[g] - uu[g];
      for (int g = 0; g < n; ++g) f[g] = S.J[0] * f[g];
      mm(q, 1, n, g2gl[4 * e + s].data(), n, f.data(), 1, fi.data(), 1);
      double W;
      switch (s) {
        case 0: W = 1.0 / Dm.w[0];
          for (int i = 0; i < q; ++i) fi[i] = W * fi[i];
          for (int i = 0; i < q; ++i) in[i * q + 0] += fi[i];
          break;
        case 1: W = 1.0 / Dm.w[q - 1];
          for (int j = 0; j < q; ++j) fi[j] = W * fi[j];
          for (int j = 0; j < q; ++j) in[(q - 1) * q + j] += fi[j];
          break;
        case 2: W = 1.0 / Dm.w[q - 1];
          for (int i = 0; i < q; ++i) fi[i] = W * fi[i];
          for (int i = 0; i < q; ++i) in[i * q + (q - 1)] += fi[i];
          break;
        case 3: W = 1.0 / Dm.w[0];
          for (int j = 0; j < q; ++j) fi[j] = W * fi[j];
          for (int j = 0; j < q; ++j) in[0 * q + j] += fi[j];
          break;
      }
    }
  }

  void project_all(VF &F) {
    for (int v = 0; v < nv; ++v)
      for (int e = 0; e < ne; ++e) D(e).project_vec(e, nod(F, v, e));
  }

  // Euler::time_derivative: F = -dU/dt (src/Euler.cpp:282-287)
  void time_derivative(VF &Us, VF &Uf) {
    fill_sides(Us);
    for (int e = 0; e < ne; ++e) volume(Us, Uf, e);
    faces(Us, Uf);
    for (int e = 0; e < ne; ++e)
      for (int v = 0; v < nv; ++v) lift(Us, Uf, v, e);
    project_all(Uf);
  }

  // Euler::set_time_step (src/Euler.cpp:804-877) with each element's own nodes
  void set_time_step(double dt, double CFLmax, double dtmax, double *out) {
    double cfl = 0.0;
    for (int e = 0; e < ne; ++e) {
      Domain &Dm = *dom[pe[e]];
      cfl = std::max(cfl, dt * Dm.cfl_element(e, nod(U, 0, e), nod(U, 1, e), nod(U, 2, e), nod(U, 3, e)));
      status |= Dm.status;
    }
    double dn = dt;
    if (CFLmax != 0 && dtmax != 0) dn = std::min(dtmax, CFLmax * dt / cfl);
    else if (CFLmax != 0) dn = CFLmax * dt / cfl;
    out[0] = cfl * dn / dt;
    out[1] = dn;
  }

  // the steppers of Domain (RungeKutta, TVD, LSRK14) on the flat nodal arrays
  void rk4(double dt) {
    static const double a[3] = {0.5, 0.5, 1.0};
    static const double b[4] = {0.16666666666666667, 0.33333333333333333, 0.33333333333333333, 0.16666666666666667};
    time_derivative(U, R0);
    Domain::axpy(-a[0] * dt, R0, U, Ut);
    Domain::scal(-b[0], R0);
    t += 0.5 * dt;
    time_derivative(Ut, R1);
    Domain::axpy(-a[1] * dt, R1, U, Ut);
    Domain::axpy2(-b[1], R1, R0);
    time_derivative(Ut, R1);
    Domain::axpy(-a[2] * dt, R1, U, Ut);
    Domain::axpy2(-b[2], R1, R0);
    t += 0.5 * dt;
    time_derivative(Ut, R1);
    Domain::axpy2(-b[3], R1, R0);
    Domain::axpy2(dt, R0, U);
  }
  void tvd(double dt) {
    static const double a[3] = {0.0, 3.0 / 4.0, 1.0 / 3.0};
    static const double b[3] = {1.0, 1.0 / 4.0, 2.0 / 3.0};
    const double tn = t;
    time_derivative(U, R0);
    Domain::axpy(-dt, R0, U, Ut);
    t = tn + dt;
    time_derivative(Ut, R0);
    Domain::axpy(-dt, R0, Ut, R0);
    Domain::axpby(a[1], U, b[1], R0, Ut);
    t = tn + 0.5 * dt;
    time_derivative(Ut, R0);
    Domain::axpy(-dt, R0, Ut, R0);
    Domain::axpby(b[2], R0, a[2], U, U);
    t = tn + dt;
  }
  void lsrk14(double dt) {
    const double tn = t;
    for (int s = 0; s < 14; ++s) {
      t = tn + kLsrkC[s] * dt;
      time_derivative(U, R0);
      Domain::axpby(-dt, R0, kLsrkA[s], Ut, Ut);
      Domain::axpy2(kLsrkB[s], Ut, U);
    }
    t = tn + dt;
  }
};

}  // namespace

// ---------------------------------------------------------------------------
// C interface (ctypes)
// ---------------------------------------------------------------------------
extern "C" {

// vert: [e][dim][nvert] as laid out in a .msh file; con: ncon rows of
// (eid, sid, leid, lsid) in .con file order.  eqn: 1 Euler, 2 NS (2D), 30 Shallow,
// 8 Euler3d, 9 NS3d (src/Reo.hpp:33-61).
void *orc_create(int nsd, int ne, int p, const double *vert, int ncon, const int *con,
                 int eqn, double gamma, double vis, double Pr, int flux) {
  Domain *d = new Domain();
  d->nsd = nsd; d->ne = ne; d->p = p; d->eqn = eqn;
  d->gamma = gamma; d->vis = vis; d->Pr = Pr; d->flux_type = flux;
  d->nv = eqn == SHALLOW2D ? 3 : nsd + 2;
  d->build_basis();
  d->build_geometry(vert);
  d->build_links(ncon, con, vert);
  d->alloc(d->U); d->alloc(d->Ut); d->alloc(d->R0); d->alloc(d->R1); d->alloc(d->Uv); d->alloc(d->Ub);
  return d;
}
void orc_destroy(void *h) { delete (Domain *)h; }
// Numerical flux at one face point: nsd 2 (U, F: 4 values) or 3 (5 values);
// flux 0 Lax-Friedrichs, 1 Roe.  Returns 1 for a non-realizable state.
int orc_riemann(int nsd, int flux, double gamma, const double *Ul, const double *Ur, const double *n,
                double *Fh) {
  double Fpl[5], Fmr[5];
  if (nsd == 2) return point_flux2d(gamma, flux, Ul, Ur, n[0], n[1], Fpl, Fmr, Fh);
  return point_flux3d(gamma, flux, Ul, Ur, n[0], n[1], n[2], Fpl, Fmr, Fh);
}
// curved elements: nodal coordinates [nsd][ne][qtot] at the GLL points replace
// the affine map of every element
void orc_set_nodes(void *h, const double *xyz) {
  Domain *d = (Domain *)h;
  const size_t n = (size_t)d->qtot, ne = (size_t)d->ne;
  for (int e = 0; e < d->ne; ++e) {
    if (d->nsd == 2) d->set_curved2(e, xyz + (0 * ne + e) * n, xyz + (1 * ne + e) * n);
    else d->set_curved(e, xyz + (0 * ne + e) * n, xyz + (1 * ne + e) * n, xyz + (2 * ne + e) * n);
  }
}
// conservative state of the state boundary condition (3D Euler)
void orc_set_bc_state(void *h, const double *U) {
  Domain *d = (Domain *)h;
  std::memcpy(d->bc_state, U, 5 * sizeof(double));
}
// viscous scheme of the Navier-Stokes residual: 0 BR1 (the reference's), 1 BR2
// with penalty eta (3D only)
void orc_set_viscous(void *h, int scheme, double eta) {
  Domain *d = (Domain *)h;
  d->visc_scheme = scheme;
  d->br2_eta = eta;
}
// matrices Ax, Ay (row-major 4x4) of the linear system (eqn 4) and its
// Lax-Friedrichs speed |u0 nx + v0 ny| + c0
void orc_set_linear(void *h, const double *Ax, const double *Ay, double u0, double v0, double c0) {
  Domain *d = (Domain *)h;
  std::memcpy(d->linA[0], Ax, 16 * sizeof(double));
  std::memcpy(d->linA[1], Ay, 16 * sizeof(double));
  d->lin_u0 = u0; d->lin_v0 = v0; d->lin_c0 = c0;
}
// Shallow water: boundary and source callbacks (either may be null), the
// domain time, and the first-residual balance (on by default, as compiled in
// the reference); turning it on again discards the stored residual.
void orc_set_functions(void *h, void (*bc)(double, int, const double *, const double *, double *),
                       void (*src)(double, int, const double *, const double *, double *)) {
  Domain *d = (Domain *)h;
  d->bc_fn = bc; d->src_fn = src;
}
// Navier-Stokes 3D: isothermal walls (boundary type 3) with wall temperature
// Tw and reference energy Eref, and the channel's bulk forcing (on/off).
void orc_set_isothermal(void *h, double Tw, double Eref) {
  Domain *d = (Domain *)h;
  d->iso_Tw = Tw; d->iso_Eref = Eref;
}
void orc_set_channel(void *h, int on) { ((Domain *)h)->channel = on != 0; }
void orc_set_time(void *h, double t) { ((Domain *)h)->t = t; }
double orc_get_time(void *h) { return ((Domain *)h)->t; }
void orc_set_balance(void *h, int on) {
  Domain *d = (Domain *)h;
  d->balance = on != 0; d->have_bal = false; d->Usrc.clear();
}
int orc_qtot(void *h) { return ((Domain *)h)->qtot; }
int orc_nmodes(void *h) { return ((Domain *)h)->nm; }
int orc_status(void *h) { return ((Domain *)h)->status; }

// coordinates of the nodes, [dim][e][qtot]
void orc_coords(void *h, double *x) {
  Domain *d = (Domain *)h;
  const size_t n = (size_t)d->qtot;
  for (int e = 0; e < d->ne; ++e) {
    std::memcpy(x + (0 * (size_t)d->ne + e) * n, d->el[e].cx.data(), n * sizeof(double));
    std::memcpy(x + (1 * (size_t)d->ne + e) * n, d->el[e].cy.data(), n * sizeof(double));
    if (d->nsd == 3) std::memcpy(x + (2 * (size_t)d->ne + e) * n, d->el[e].cz.data(), n * sizeof(double));
  }
}
// element geometry and side tables, for checks on the product's setup
void orc_geometry(void *h, double *J, double *metrics /* ne x 9 */) {
  Domain *d = (Domain *)h;
  for (int e = 0; e < d->ne; ++e) {
    const Elem &E = d->el[e];
    J[e] = E.J;
    const double m[9] = {E.rx, E.ry, E.rz, E.sx, E.sy, E.sz, E.tx, E.ty, E.tz};
    std::memcpy(metrics + 9 * (size_t)e, m, sizeof(m));
  }
}
void orc_basis(void *h, double *z, double *w, double *zg, double *wg, double *D, double *B) {
  Domain *d = (Domain *)h;
  std::memcpy(z, d->z.data(), d->q * sizeof(double));
  std::memcpy(w, d->w.data(), d->q * sizeof(double));
  std::memcpy(zg, d->zg.data(), d->qf * sizeof(double));
  std::memcpy(wg, d->wg.data(), d->qf * sizeof(double));
  std::memcpy(D, d->D.data(), d->q * d->q * sizeof(double));
  std::memcpy(B, d->B.data(), d->L * d->q * sizeof(double));
}
void orc_set_state(void *h, const double *U) {
  Domain *d = (Domain *)h;
  std::memcpy(d->U.u.data(), U, d->U.u.size() * sizeof(double));
}
void orc_get_state(void *h, double *U) {
  Domain *d = (Domain *)h;
  std::memcpy(U, d->U.u.data(), d->U.u.size() * sizeof(double));
}
// forward transform to modal then back (vField::write side effect,
// src/vField.cpp:1102-1181)
void orc_project_state(void *h) { Domain *d = (Domain *)h; d->project_all(d->U); }
// modal coefficients [v][e][nModes] of the current state
void orc_modal(void *h, double *Uh) {
  Domain *d = (Domain *)h;
  std::vector<double> f(d->qtot);
  for (int v = 0; v < d->nv; ++v)
    for (int e = 0; e < d->ne; ++e) {
      std::memcpy(f.data(), d->eu(d->U, v, e), d->qtot * sizeof(double));
      d->forward(e, f.data(), Uh + ((size_t)v * d->ne + e) * d->nm);
    }
}
// state from modal coefficients (vField::read + backward_transform)
void orc_set_modal(void *h, const double *Uh) {
  Domain *d = (Domain *)h;
  for (int v = 0; v < d->nv; ++v)
    for (int e = 0; e < d->ne; ++e)
      d->backward(Uh + ((size_t)v * d->ne + e) * d->nm, d->eu(d->U, v, e));
}
// one residual evaluation F = -dU/dt of the given nodal state
int orc_rhs(void *h, const double *U, double *F) {
  Domain *d = (Domain *)h;
  d->status = 0;
  std::memcpy(d->Ut.u.data(), U, d->Ut.u.size() * sizeof(double));
  d->time_derivative(d->Ut, d->R0);
  std::memcpy(F, d->R0.u.data(), d->R0.u.size() * sizeof(double));
  return d->status;
}
// one residual evaluation on a subset of elements only is not meaningful
// (faces couple them); the baseline leg times whole residuals instead.
int orc_advance(void *h, int inttype, double dt, int nsteps) {
  Domain *d = (Domain *)h;
  d->status = 0;
  for (int s = 0; s < nsteps && !d->status; ++s) {
    if (inttype == 4) d->rk4(dt);
    else if (inttype == 3) d->tvd(dt);
    else if (inttype == 14) d->lsrk14(dt);
    else return -1;
  }
  return d->status;
}
// Linf, L2, H1 per field of the current state (src/Field.cpp:4675-4697)
int orc_set_time_step(void *h, double dt, double CFLmax, double dtmax, double *out) {
  Domain *d = (Domain *)h;
  d->status = 0;
  d->set_time_step(dt, CFLmax, dtmax, out);
  return d->status;
}
void orc_norms(void *h, double *out) {
  Domain *d = (Domain *)h;
  const int n = d->qtot;
  std::vector<double> fx(n), fy(n), fz(n);
  for (int v = 0; v < d->nv; ++v) {
    double Linf = 0, L2 = 0, H1 = 0, area = 0;
    for (int e = 0; e < d->ne; ++e) {
      const double *u = d->eu(d->U, v, e);
      const std::vector<double> &b = d->el[e].wJ;
      double l2 = 0, ar = 0, h1 = 0, li = 0;
      for (int i = 0; i < n; ++i) { l2 += b[i] * (u[i] * u[i]); ar += b[i]; li = std::max(li, std::fabs(u[i])); }
      d->gradient_all(e, u, fx.data(), fy.data(), fz.data());
      for (int i = 0; i < n; ++i) {
        double g = fx[i] * fx[i];
        g = fy[i] * fy[i] + g;
        if (d->nsd == 3) g = fz[i] * fz[i] + g;
        double f = u[i] * u[i];
        f = f + g;
        h1 += b[i] * f;
      }
      Linf = std::max(Linf, li); L2 += l2; H1 += h1; area += ar;
    }
    out[3 * v + 0] = Linf;
    out[3 * v + 1] = std::sqrt(std::fabs(L2 / area));
    out[3 * v + 2] = std::sqrt(std::fabs(H1 / area));
  }
}

// Element-block transforms (GpGpu::hex_kernels, src/GpGpu.cu:236-262 and
// 505-750, used by HexBlock::forward_transform / backward_transform,
// src/HexBlock.cpp:54-130), element by element as Hex::inner_product
// (src/Hex.cpp:703-747: f *= wJ, H = Ba F, Ha = Bb H per a, fh = Ha Bc^T)
// and Hex::backward_transform (src/Hex.cpp:796-812: H = Ba^T fh,
// Ha = Bb^T H per a, f = Ha Bc).  Slots of qa*qb*qc doubles per element, the
// nm^3 modes first; in place.  wj_stride 0: one wJ for all elements.
void oracle_hexblock_inner_product(long long ne, int nm, int qa, int qb, int qc, const double *wJ,
                                   int wj_stride, const double *Ba, const double *Bb, const double *Bc,
                                   double *u) {
  const int qt = qa * qb * qc, bc = qb * qc;
  std::vector<double> f(qt), H((size_t)nm * bc), Ha((size_t)nm * nm * qc);
  for (long long e = 0; e < ne; ++e) {
    double *ue = u + e * qt;
    const double *w = wJ + e * wj_stride;
    for (int i = 0; i < qt; ++i) f[i] = ue[i] * w[i];
    for (int i = 0; i < nm; ++i)
      for (int r = 0; r < bc; ++r) {
        double s = 0.0;
        for (int a = 0; a < qa; ++a) s = s + Ba[i * qa + a] * f[a * bc + r];
        H[i * bc + r] = s;
      }
    for (int i = 0; i < nm; ++i)
      for (int j = 0; j < nm; ++j)
        for (int c = 0; c < qc; ++c) {
          double s = 0.0;
          for (int b = 0; b < qb; ++b) s = s + Bb[j * qb + b] * H[i * bc + b * qc + c];
          Ha[(i * nm + j) * qc + c] = s;
        }
    for (int ij = 0; ij < nm * nm; ++ij)
      for (int k = 0; k < nm; ++k) {
        double s = 0.0;
        for (int c = 0; c < qc; ++c) s = s + Ha[ij * qc + c] * Bc[k * qc + c];
        ue[ij * nm + k] = s;
      }
  }
}

void oracle_hexblock_backward_transform(long long ne, int nm, int qa, int qb, int qc, const double *Ba,
                                        const double *Bb, const double *Bc, double *u) {
  const int qt = qa * qb * qc, n2 = nm * nm, n3 = n2 * nm;
  std::vector<double> fh(n3), H((size_t)qa * n2), Ha((size_t)qa * qb * nm);
  for (long long e = 0; e < ne; ++e) {
    double *ue = u + e * qt;
    for (int i = 0; i < n3; ++i) fh[i] = ue[i];
    for (int a = 0; a < qa; ++a)
      for (int r = 0; r < n2; ++r) {
        double s = 0.0;
        for (int i = 0; i < nm; ++i) s = s + Ba[i * qa + a] * fh[i * n2 + r];
        H[a * n2 + r] = s;
      }
    for (int a = 0; a < qa; ++a)
      for (int b = 0; b < qb; ++b)
        for (int k = 0; k < nm; ++k) {
          double s = 0.0;
          for (int j = 0; j < nm; ++j) s = s + Bb[j * qb + b] * H[a * n2 + j * nm + k];
          Ha[(a * qb + b) * nm + k] = s;
        }
    for (int ab = 0; ab < qa * qb; ++ab)
      for (int c = 0; c < qc; ++c) {
        double s = 0.0;
        for (int k = 0; k < nm; ++k) s = s + Ha[ab * nm + k] * Bc[k * qc + c];
        ue[ab * qc + c] = s;
      }
  }
}

// Variable-order quads (VarDomain2).  Arrays are padded per element: nodal
// [v][e][qmax^2] with element e's q_e^2 nodes first, modal [v][e][Lmax^2]
// with its L_e^2 modes first (the rest zero on output, ignored on input).
void *orc_varp_create(int ne, const int *orders, const double *vert, int ncon, const int *con, double gamma,
                      int flux) {
  VarDomain2 *d = new VarDomain2();
  d->build(ne, orders, vert, ncon, con, gamma, flux);
  return d;
}
void orc_varp_destroy(void *h) { delete (VarDomain2 *)h; }
int orc_varp_qmax(void *h) { return ((VarDomain2 *)h)->qmax; }
static void varp_pad_out(VarDomain2 *d, const VF &F, double *out, int nv) {
  const size_t Q2 = (size_t)d->qmax * d->qmax;
  std::fill(out, out + (size_t)nv * d->ne * Q2, 0.0);
  for (int v = 0; v < nv; ++v)
    for (int e = 0; e < d->ne; ++e) {
      const size_t n = (size_t)d->q_of(e) * d->q_of(e);
      std::memcpy(out + ((size_t)v * d->ne + e) * Q2, &F.u[(size_t)v * d->ntot + d->off[e]], n * sizeof(double));
    }
}
static void varp_pad_in(VarDomain2 *d, const double *in, VF &F, int nv) {
  const size_t Q2 = (size_t)d->qmax * d->qmax;
  for (int v = 0; v < nv; ++v)
    for (int e = 0; e < d->ne; ++e) {
      const size_t n = (size_t)d->q_of(e) * d->q_of(e);
      std::memcpy(&F.u[(size_t)v * d->ntot + d->off[e]], in + ((size_t)v * d->ne + e) * Q2, n * sizeof(double));
    }
}
void orc_varp_coords(void *h, double *x) {
  VarDomain2 *d = (VarDomain2 *)h;
  VF c;
  c.u.assign(2 * d->ntot, 0.0);
  for (int e = 0; e < d->ne; ++e) {
    const Elem &E = d->D(e).el[e];
    std::memcpy(&c.u[d->off[e]], E.cx.data(), E.cx.size() * sizeof(double));
    std::memcpy(&c.u[d->ntot + d->off[e]], E.cy.data(), E.cy.size() * sizeof(double));
  }
  varp_pad_out(d, c, x, 2);
}
int orc_varp_set_time_step(void *h, double dt, double CFLmax, double dtmax, double *out) {
  VarDomain2 *d = (VarDomain2 *)h;
  d->status = 0;
  for (Domain *dm : d->dom)
    if (dm) dm->status = 0;
  d->set_time_step(dt, CFLmax, dtmax, out);
  return d->status;
}
void orc_varp_set_bc_state(void *h, const double *U) {
  VarDomain2 *d = (VarDomain2 *)h;
  for (Domain *dm : d->dom)
    if (dm) std::memcpy(dm->bc_state, U, 5 * sizeof(double));
}
void orc_varp_set_state(void *h, const double *U) { VarDomain2 *d = (VarDomain2 *)h; varp_pad_in(d, U, d->U, 4); }
void orc_varp_get_state(void *h, double *U) { VarDomain2 *d = (VarDomain2 *)h; varp_pad_out(d, d->U, U, 4); }
void orc_varp_project_state(void *h) { VarDomain2 *d = (VarDomain2 *)h; d->project_all(d->U); }
void orc_varp_modal(void *h, double *Uh) {
  VarDomain2 *d = (VarDomain2 *)h;
  const size_t L2 = (size_t)d->Lmax * d->Lmax;
  std::fill(Uh, Uh + 4 * d->ne * L2, 0.0);
  std::vector<double> f((size_t)d->qmax * d->qmax);
  for (int v = 0; v < 4; ++v)
    for (int e = 0; e < d->ne; ++e) {
      const int q = d->q_of(e);
      std::memcpy(f.data(), d->nod(d->U, v, e), (size_t)q * q * sizeof(double));
      d->D(e).forward(e, f.data(), Uh + ((size_t)v * d->ne + e) * L2);
    }
}
void orc_varp_set_modal(void *h, const double *Uh) {
  VarDomain2 *d = (VarDomain2 *)h;
  const size_t L2 = (size_t)d->Lmax * d->Lmax;
  for (int v = 0; v < 4; ++v)
    for (int e = 0; e < d->ne; ++e) d->D(e).backward(Uh + ((size_t)v * d->ne + e) * L2, d->nod(d->U, v, e));
}
int orc_varp_rhs(void *h, const double *U, double *F) {
  VarDomain2 *d = (VarDomain2 *)h;
  d->status = 0;
  varp_pad_in(d, U, d->Ut, 4);
  d->time_derivative(d->Ut, d->R0);
  varp_pad_out(d, d->R0, F, 4);
  return d->status;
}
int orc_varp_advance(void *h, int inttype, double dt, int nsteps) {
  VarDomain2 *d = (VarDomain2 *)h;
  d->status = 0;
  for (int s = 0; s < nsteps && !d->status; ++s) {
    if (inttype == 4) d->rk4(dt);
    else if (inttype == 3) d->tvd(dt);
    else if (inttype == 14) d->lsrk14(dt);
    else return -1;
  }
  return d->status;
}

// Linf, L2, H1 per field of the current state, as orc_norms with each
// element's own basis
void orc_varp_norms(void *h, double *out) {
  VarDomain2 *d = (VarDomain2 *)h;
  for (int v = 0; v < 4; ++v) {
    double Linf = 0, L2 = 0, H1 = 0, area = 0;
    for (int e = 0; e < d->ne; ++e) {
      const Domain &Dm = d->D(e);
      const int n = d->q_of(e) * d->q_of(e);
      std::vector<double> fx(n), fy(n);
      const double *u = d->nod(d->U, v, e);
      const std::vector<double> &b = Dm.el[e].wJ;
      double l2 = 0, ar = 0, h1 = 0, li = 0;
      for (int i = 0; i < n; ++i) { l2 += b[i] * (u[i] * u[i]); ar += b[i]; li = std::max(li, std::fabs(u[i])); }
      Dm.gradient_all(e, u, fx.data(), fy.data(), nullptr);
      for (int i = 0; i < n; ++i) {
        double g = fx[i] * fx[i];
        g = fy[i] * fy[i] + g;
        double f = u[i] * u[i];
        f = f + g;
        h1 += b[i] * f;
      }
      Linf = std::max(Linf, li); L2 += l2; H1 += h1; area += ar;
    }
    out[3 * v + 0] = Linf;
    out[3 * v + 1] = std::sqrt(std::fabs(L2 / area));
    out[3 * v + 2] = std::sqrt(std::fabs(H1 / area));
  }
}

}  // extern "C"
