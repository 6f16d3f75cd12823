// Device-side building blocks of the DG residual: LDS tensor contractions
// (sum factorization), pointwise Euler/Navier-Stokes fluxes and the
// Lax-Friedrichs / Roe numerical fluxes.  Everything here is FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dgx {

// Basis tables of one polynomial order (device memory, row-major).
//   BT  [Q][L]  B_l(xi_i)                 modal -> GLL nodes
//   BW  [L][Q]  B_l(xi_i) w_i             GLL inner product
//   QD  [L][Q]  sum_i B_l(xi_i) w_i D_ik  inner product of a collocation derivative
//   Bg  [QF][L] B_l(eta_a)                modal -> Gauss face points
//   Bend[2][L]  B_l(-1), B_l(+1)
//   PW  [L][QF] wg_a B_l(eta_a)           Gauss face inner product
//   Ig  [QF][Q] GLL -> Gauss interpolation
struct Tab {
  const double *BT, *BW, *QD, *Bg, *Bend, *PW, *Ig;
};

// The same tables for every supported order p = 1..10 in constant memory, so
// that uniform coefficient reads become scalar loads.  Layout per order:
// BT, BW, QD, Bg, Bend, PW, Ig (sizes above), then PI = PW Ig [L][Q] (GLL
// face layer -> modal face coefficients), KE [2][Q] (KE[s][i] =
// sum_l Bend[s][l] BT[i][l], the end traces sum_l phi_l(-+1) phi_l(xi_i)),
// BL [L][Q], GL [Q][QF], DN [Q][Q], NP1 [Q][Q] and NPD [Q][Q] (see CT), orders
// packed in ascending p.
constexpr int kMaxOrder = 10;
constexpr int tab_size(int p) {
  return 3 * (p + 2) * (p + 1) + 2 * (p + 1) * (p + 1) + 2 * (p + 1) + (p + 1) * (p + 2) +
         (p + 1) * (p + 2) + 2 * (p + 2) + (p + 1) * (p + 2) + (p + 2) * (p + 1) + 3 * (p + 2) * (p + 2);
}
constexpr int tab_offset(int p) {
  int s = 0;
  for (int k = 1; k < p; ++k) s += tab_size(k);
  return s;
}
constexpr int kTabTotal = tab_offset(kMaxOrder + 1);

// Each translation unit holds its own copy, filled by its upload_tables_* ().
static __constant__ double c_tab[kTabTotal];

template <int P> struct CT {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1, b = tab_offset(P);
  __device__ static __forceinline__ const double *BT() { return c_tab + b; }
  __device__ static __forceinline__ const double *BW() { return c_tab + b + Q * L; }
  __device__ static __forceinline__ const double *QD() { return c_tab + b + 2 * Q * L; }
  __device__ static __forceinline__ const double *Bg() { return c_tab + b + 3 * Q * L; }
  __device__ static __forceinline__ const double *Bend() { return c_tab + b + 3 * Q * L + QF * L; }
  __device__ static __forceinline__ const double *PW() { return c_tab + b + 3 * Q * L + QF * L + 2 * L; }
  __device__ static __forceinline__ const double *Ig() { return c_tab + b + 3 * Q * L + 2 * QF * L + 2 * L; }
  __device__ static __forceinline__ const double *PI() { return Ig() + QF * Q; }
  __device__ static __forceinline__ const double *KE() { return PI() + L * Q; }
  // curved elements: BL [L][Q] = phi_l(xi_i) (unweighted projection), GL [Q][QF]
  // Gauss -> GLL interpolation (= BT PW), DN [Q][Q] GLL collocation derivative
  __device__ static __forceinline__ const double *BL() { return KE() + 2 * Q; }
  __device__ static __forceinline__ const double *GL() { return BL() + L * Q; }
  __device__ static __forceinline__ const double *DN() { return GL() + Q * QF; }
  // nodal-to-nodal [Q][Q]: NP1 = BT BW (the L2 projection onto the order-P
  // modes, evaluated back at the GLL points) and NPD = BT QD (the projected
  // derivative, at the GLL points)
  __device__ static __forceinline__ const double *NP1() { return DN() + Q * Q; }
  __device__ static __forceinline__ const double *NPD() { return NP1() + Q * Q; }
};

// Per-element affine geometry and face links (device memory).
struct Geo {
  const double *metric;          // [ne][9]: rx ry rz sx sy sz tx ty tz
  const double *fnorm;           // [ne][nsides][3] outward unit normal
  const double *fjr;             // [ne][nsides] J_side / J_vol
  const int *nbr;                // [ne][nsides] neighbour element
  const int *nside;              // [ne][nsides] neighbour side
  const uint8_t *permid;         // [ne][nsides] row of `perm`
  const uint16_t *perm;          // [nperm][QF^(d-1)] own point -> neighbour point
  const uint8_t *master;         // [ne][nsides] 1: master or boundary side, 0: slave
};

struct Phys {
  double gamma, vis, Pr;
  int flux;                      // 0 Lax-Friedrichs, 1 Roe
  int visc;                      // viscous face gradient: 0 BR1, 1 BR2
  double eta;                    // BR2 penalty
  double bcU[5];                 // conservative state of state boundaries
  double wallT, wallE;           // isothermal walls: temperature Tw and reference energy Eref
};

// Shallow water in (phi, uphi, vphi) (src/Shallow.cpp:357-525): Cartesian
// normal flux of one state and its largest wave speed |u_n| + sqrt(phi);
// returns true when phi or phi^2/2 is below aepsM = eps^(2/3), where the
// reference stops.
__device__ __forceinline__ bool shallow_normal_flux2(const double U[3], double nx, double ny, double F[3],
                                                     double &lam) {
  constexpr double aepsM = 3.666852862501036e-11;
  const double phi = U[0], mn = U[1] * nx + U[2] * ny, mt = U[2] * nx - U[1] * ny;
  const double un = (1.0 / phi) * mn, p = 0.5 * phi * phi;
  F[0] = mn;
  F[1] = (mn * un + p) * nx - (mt * un) * ny;
  F[2] = (mn * un + p) * ny + (mt * un) * nx;
  lam = fabs(un) + sqrt(phi);
  return !(phi >= aepsM) || !(p >= aepsM);
}

// Shallow-water ghost values: Shallow::WallBC::apply (src/ShallowBC.hpp:45-62)
// keeps phi and the tangential momentum; state boundaries hold bcU[0..2].
__device__ __forceinline__ void ghost_shallow2(int bc, const double UL[3], double nx, double ny,
                                               const double bcU[5], double UB[3]) {
#pragma clang fp contract(off)
  if (bc == 1) {
    UB[0] = bcU[0]; UB[1] = bcU[1]; UB[2] = bcU[2];
    return;
  }
  UB[0] = UL[0];
  UB[1] = (UL[1] * ny - UL[2] * nx) * ny;
  UB[2] = (UL[2] * nx - UL[1] * ny) * nx;
}

// Ghost values of an Euler (2D) boundary side: bc 0 slip wall
// (Euler::WallBC::apply, src/EulerBC.hpp:43-62: the density kept, the momentum
// reflected onto the tangent, the energy moved to the ghost's kinetic energy),
// bc 1 state (StateBC::apply, src/BC.cpp:17-28).
__device__ __forceinline__ void ghost2(int bc, const double UL[4], double nx, double ny, const double bcU[5],
                                       double UB[4]) {
#pragma clang fp contract(off)
  if (bc == 1) {
    UB[0] = bcU[0]; UB[1] = bcU[1]; UB[2] = bcU[2]; UB[3] = bcU[3];
    return;
  }
  UB[0] = UL[0];
  UB[1] = (UL[1] * ny - UL[2] * nx) * ny;
  UB[2] = (UL[2] * nx - UL[1] * ny) * nx;
  UB[3] = UL[3] - 0.5 * (UL[1] * UL[1] + UL[2] * UL[2]) / UL[0] + 0.5 * (UB[1] * UB[1] + UB[2] * UB[2]) / UB[0];
}

// One component of a boundary or source function at (x, y, t): the postfix
// program code[b] .. code[e] (dgx_function).  The host checks every program at
// domain creation (opcodes, constant indices, a stack depth within 16 that
// ends at one value), so the stack index stays in bounds.
__device__ __noinline__ double eval_program(const int *__restrict__ code, int b, int e,
                                            const double *__restrict__ consts, double x, double y, double t) {
#pragma clang fp contract(off)
  double st[16];
  int sp = 0;
  for (int i = b; i < e; ++i) {
    const int ins = code[i], op = ins & 255;
    if (op <= 4) {
      st[sp++] = op == 0 ? consts[ins >> 8] : (op == 1 ? x : (op == 2 ? y : (op == 3 ? 0.0 : t)));
    } else if (op <= 9 || op == 25) {
      const double bb = st[--sp], a = st[sp - 1];
      double r;
      switch (op) {
        case 5: r = a + bb; break;
        case 6: r = a - bb; break;
        case 7: r = a * bb; break;
        case 8: r = a / bb; break;
        case 9: r = pow(a, bb); break;
        default: r = atan2(a, bb); break;
      }
      st[sp - 1] = r;
    } else {
      const double a = st[sp - 1];
      double r;
      switch (op) {
        case 10: r = -a; break;
        case 11: r = sin(a); break;
        case 12: r = cos(a); break;
        case 13: r = tan(a); break;
        case 14: r = asin(a); break;
        case 15: r = acos(a); break;
        case 16: r = atan(a); break;
        case 17: r = sinh(a); break;
        case 18: r = cosh(a); break;
        case 19: r = tanh(a); break;
        case 20: r = exp(a); break;
        case 21: r = log(a); break;
        case 22: r = log10(a); break;
        case 23: r = sqrt(a); break;
        default: r = fabs(a); break;
      }
      st[sp - 1] = r;
    }
  }
  return st[0];
}

// Ghost values of a boundary side, as Domain::enforce_bc fills them
// (src/Domain.cpp:1346-1350): bc 0 slip wall (Euler3d::WallBC::apply,
// src/Euler3d.cpp:876-909: the normal momentum removed, the internal energy
// kept), bc 1 state (StateBC::apply, src/BC.cpp:17-28).  Isothermal walls
// (bc 3) take the slip wall here: Domain::BC_db holds the Euler3d::WallBC that
// Euler3d::make_BC makes of every "wall" code (src/Euler3d.cpp:857-860), and
// the isothermal ghost enters only the viscous terms (ghost_prim3).  A side
// whose neighbour index nb is negative is a boundary of type -1 - nb.
__device__ __forceinline__ void ghost3(int bc, const double UL[5], double nx, double ny, double nz,
                                       const double bcU[5], double UB[5]) {
  // no fused multiply-adds: the ghost energy is a difference of O(1/Ma^2)
  // energies, and the wall's Riemann wave strengths are differences of it
#pragma clang fp contract(off)
  if (bc == 1) {
#pragma unroll
    for (int v = 0; v < 5; ++v) UB[v] = bcU[v];
    return;
  }
  const double dot = UL[1] * nx + UL[2] * ny + UL[3] * nz;
  UB[0] = UL[0];
  UB[1] = UL[1] - dot * nx;
  UB[2] = UL[2] - dot * ny;
  UB[3] = UL[3] - dot * nz;
  UB[4] = UL[4] - 0.5 * (UL[1] * UL[1] + UL[2] * UL[2] + UL[3] * UL[3]) / UL[0] +
          0.5 * (UB[1] * UB[1] + UB[2] * UB[2] + UB[3] * UB[3]) / UB[0];
}

// Primitive values (u, v, w, e) of the ghost of a Navier-Stokes boundary side
// at one face point, as Navier_Stokes3d::enforce_bc and Euler3d::primitive_sides
// give them (src/Navier_Stokes3d.cpp:457-492, src/Euler3d.cpp:1162-1199):
// bc 3 isothermal wall (rho-, 0, 0, 0, (Eref - 1/2) rho- Tw), bc 1 state bcU.
__device__ __forceinline__ void ghost_prim3(int bc, double rho, const Phys &ph, double pq[4]) {
#pragma clang fp contract(off)
  if (bc == 3) {
    const double E = (ph.wallE - 0.5) * rho * ph.wallT;
    pq[0] = 0.0; pq[1] = 0.0; pq[2] = 0.0; pq[3] = E / rho;
    return;
  }
  const double r = ph.bcU[0], u = ph.bcU[1] / r, v = ph.bcU[2] / r, w = ph.bcU[3] / r;
  double wk = u * u; wk = v * v + wk; wk = w * w + wk;
  pq[0] = u; pq[1] = v; pq[2] = w; pq[3] = -0.5 * wk + ph.bcU[4] / r;
}

// Contract dimension DIMC of the LDS tensor `in` (shape N0 x N1 x N2, last
// index fastest) with the NO x NI matrix M; output has NO in place of NI.
// One thread owns one pencil at a time and computes OC outputs per pass from
// its NI inputs (register blocking).  When every thread owns at most one
// pencil, its inputs are read once and all outputs computed from registers;
// otherwise passes over output chunks are outermost so that only OC x NI
// coefficients are live in scalar registers at a time.  Both orders give
// bitwise the same sums.  ACC adds into `out`.  T is the arithmetic type of
// the matrix, the tensors and the sums (double, or float on the mixed paths).
template <int N0, int N1, int N2, int DIMC, int NO, int NT, bool ACC = false, typename T = double>
__device__ __forceinline__ void contract(const T *__restrict__ M, const T *in, T *out, int tid) {
  constexpr int NI = DIMC == 0 ? N0 : (DIMC == 1 ? N1 : N2);
  constexpr int NP = (N0 * N1 * N2) / NI;
  constexpr int O1 = DIMC == 1 ? NO : N1;
  constexpr int O2 = DIMC == 2 ? NO : N2;
  constexpr int SI = DIMC == 0 ? N1 * N2 : (DIMC == 1 ? N2 : 1);
  constexpr int SO = DIMC == 0 ? O1 * O2 : (DIMC == 1 ? O2 : 1);
  constexpr int OC = NI * 3 <= 40 ? 3 : 2;    // coefficients live per pass <= ~40 doubles
  auto pencil = [&](int pc, int o0, int o1) {
    int i0, i1, i2;
    if (DIMC == 0) { i0 = 0; i1 = pc / N2; i2 = pc - i1 * N2; }
    else if (DIMC == 1) { i0 = pc / N2; i1 = 0; i2 = pc - i0 * N2; }
    else { i0 = pc / N1; i1 = pc - i0 * N1; i2 = 0; }
    const int bi = (i0 * N1 + i1) * N2 + i2;
    const int bo = (i0 * O1 + i1) * O2 + i2;
    T x[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) x[i] = in[bi + i * SI];
#pragma unroll
    for (int o = o0; o < o1; ++o) {
      T s = 0;
#pragma unroll
      for (int i = 0; i < NI; ++i) s = fma(M[o * NI + i], x[i], s);
      if (ACC) out[bo + o * SO] += s;
      else out[bo + o * SO] = s;
    }
  };
  if constexpr (NP <= NT) {
    if (tid < NP) pencil(tid, 0, NO);
  } else {
#pragma unroll
    for (int o0 = 0; o0 < NO; o0 += OC)
      for (int pc = tid; pc < NP; pc += NT) pencil(pc, o0, o0 + OC < NO ? o0 + OC : NO);
  }
}

// Edge contributions at modal (l,m): edges 0,2 normal to b, edges 1,3 normal to a
// (arithmetic in T: double, or float on the mixed paths).
template <int P, typename T>
__device__ __forceinline__ T face_add2(const T *__restrict__ Bend, const T *H, int l, int m, const double c[4]) {
  constexpr int L = P + 1;
  const T *B0 = Bend, *B1 = Bend + L;
  T s = 0;
  s = fma((T)c[0] * B0[m], H[0 * L + l], s);
  s = fma((T)c[2] * B1[m], H[2 * L + l], s);
  s = fma((T)c[1] * B1[l], H[1 * L + m], s);
  s = fma((T)c[3] * B0[l], H[3 * L + m], s);
  return s;
}

// Contribution of the six face blocks H[s][.][.] at modal index (l,m,n),
// weighted per face by c[s].  Face normal dims: 4,2 -> a; 1,3 -> b; 0,5 -> c.
template <int P>
__device__ __forceinline__ double face_add3(const double *__restrict__ Bend, const double *H,
                                            int l, int m, int n, const double c[6]) {
  constexpr int L = P + 1, L2 = L * L;
  const double *B0 = Bend, *B1 = Bend + L;
  double s = 0.0;
  s = fma(c[4] * B0[l], H[4 * L2 + m * L + n], s);
  s = fma(c[2] * B1[l], H[2 * L2 + m * L + n], s);
  s = fma(c[1] * B0[m], H[1 * L2 + l * L + n], s);
  s = fma(c[3] * B1[m], H[3 * L2 + l * L + n], s);
  s = fma(c[0] * B0[n], H[0 * L2 + l * L + m], s);
  s = fma(c[5] * B1[n], H[5 * L2 + l * L + m], s);
  return s;
}

// Element of workgroup i.  Workgroups are dispatched round-robin over the 8
// XCDs, each with its own L2; this bijection of [0, ne) gives XCD x the
// contiguous block of elements x*q + min(x, r) + [0, q + (x < r)) (ne = 8q + r),
// so that face neighbours (which read each other's traces) mostly share an L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_element(int i, int ne) {
  const int q = ne / kXcds, r = ne - q * kXcds;
  const int x = i % kXcds, j = i / kXcds;
  return x * q + (x < r ? x : r) + j;
}

// Copy n doubles global -> LDS (or LDS -> global) with the workgroup.
template <int NT>
__device__ __forceinline__ void copy_block(const double *__restrict__ src, double *dst, int n,
                                           int tid) {
  for (int i = tid; i < n; i += NT) dst[i] = src[i];
}

// aepsM = eps^0.66 (src/Constants.hpp:43): realizability threshold.
constexpr double kAepsM = 4.6012e-11;

// Physical normal flux F(U).n of compressible Euler (3D), pressure and sound speed.
__device__ __forceinline__ void euler_normal_flux3(const double U[5], double nx, double ny,
                                                   double nz, double gm1, double F[5],
                                                   double &lam, bool &bad) {
  const double ir = 1.0 / U[0];
  const double u = U[1] * ir, v = U[2] * ir, w = U[3] * ir;
  const double p = gm1 * (U[4] - 0.5 * (U[1] * u + U[2] * v + U[3] * w));
  const double un = u * nx + v * ny + w * nz;
  F[0] = U[0] * un;
  F[1] = U[1] * un + p * nx;
  F[2] = U[2] * un + p * ny;
  F[3] = U[3] * un + p * nz;
  F[4] = un * (U[4] + p);
  const double c = sqrt((gm1 + 1.0) * p * ir);
  lam = fabs(un) + c;
  bad = !(U[0] > kAepsM) || !(p > kAepsM);
}

__device__ __forceinline__ void euler_normal_flux2(const double U[4], double nx, double ny,
                                                   double gm1, double F[4], double &lam,
                                                   bool &bad) {
  const double ir = 1.0 / U[0];
  const double u = U[1] * ir, v = U[2] * ir;
  const double p = gm1 * (U[3] - 0.5 * (U[1] * u + U[2] * v));
  const double un = u * nx + v * ny;
  F[0] = U[0] * un;
  F[1] = U[1] * un + p * nx;
  F[2] = U[2] * un + p * ny;
  F[3] = un * (U[3] + p);
  const double c = sqrt((gm1 + 1.0) * p * ir);
  lam = fabs(un) + c;
  bad = !(U[0] > kAepsM) || !(p > kAepsM);
}

// 3D Roe flux without entropy fix: the reference's 2D Roe flux
// (src/Euler.cpp:682-741) extended by the second shear wave, written without a
// tangent basis (the two shear waves share the eigenvalue u.n, so only the
// tangential jump vector enters).  FL = F(UL).n; Fh = FL + sum_k min(lambda_k,0)
// alpha_k r_k.  The reference has no 3D Roe flux.
__device__ __forceinline__ void roe_flux3(const double UL[5], const double UR[5], double nx, double ny,
                                          double nz, double gamma, const double FL[5], double Fh[5]) {
  const double gm1 = gamma - 1.0;
  const double rl = UL[0], rr = UR[0];
  const double irl = 1.0 / rl, irr = 1.0 / rr;
  const double ulx = UL[1] * irl, uly = UL[2] * irl, ulz = UL[3] * irl;
  const double urx = UR[1] * irr, ury = UR[2] * irr, urz = UR[3] * irr;
  const double pl = gm1 * (UL[4] - 0.5 * (UL[1] * ulx + UL[2] * uly + UL[3] * ulz));
  const double pr = gm1 * (UR[4] - 0.5 * (UR[1] * urx + UR[2] * ury + UR[3] * urz));
  const double Hl = (UL[4] + pl) * irl, Hr = (UR[4] + pr) * irr;
  const double R = sqrt(rr / rl), iRp1 = 1.0 / (1.0 + R);
  const double ux = (ulx + R * urx) * iRp1, uy = (uly + R * ury) * iRp1, uz = (ulz + R * urz) * iRp1;
  const double H = (Hl + R * Hr) * iRp1;
  const double un = ux * nx + uy * ny + uz * nz;
  const double tx = ux - un * nx, ty = uy - un * ny, tz = uz - un * nz;   // tangential velocity
  const double a2 = gm1 * (H - 0.5 * (ux * ux + uy * uy + uz * uz));
  const double a = sqrt(a2), ainv = 1.0 / a, c1inv = a2 / gm1;
  const double dr = rr - rl, dE = UR[4] - UL[4];
  const double dmx = UR[1] - UL[1], dmy = UR[2] - UL[2], dmz = UR[3] - UL[3];
  const double dmn = dmx * nx + dmy * ny + dmz * nz;
  // shear strengths (tangential momentum jump minus tangential velocity x density jump)
  const double sx = (dmx - dmn * nx) - tx * dr, sy = (dmy - dmn * ny) - ty * dr, sz = (dmz - dmn * nz) - tz * dr;
  const double st = sx * tx + sy * ty + sz * tz;
  const double a1 = gm1 * ainv * ainv * (dr * (H - un * un) + un * dmn - (dE - st));
  const double a0 = 0.5 * ainv * (dr * (un + a) - dmn - a * a1);
  const double a4 = dr - (a0 + a1);
  const double l0 = fmin(un - a, 0.0), l1 = fmin(un, 0.0), l4 = fmin(un + a, 0.0);
  const double g0 = l0 * a0 + l1 * a1 + l4 * a4;
  const double gn = l0 * (un - a) * a0 + l1 * un * a1 + l4 * (un + a) * a4;
  Fh[0] = FL[0] + g0;
  Fh[1] = FL[1] + gn * nx + g0 * tx + l1 * sx;
  Fh[2] = FL[2] + gn * ny + g0 * ty + l1 * sy;
  Fh[3] = FL[3] + gn * nz + g0 * tz + l1 * sz;
  Fh[4] = FL[4] + l0 * (H - un * a) * a0 + l1 * (H - c1inv) * a1 + l1 * st + l4 * (H + un * a) * a4;
}

// Numerical flux of Euler3d at one face point: FL = F(UL).n and Fh by
// Lax-Friedrichs (flux 0, src/Euler3d.cpp:772-795) or Roe (flux 1).
__device__ __forceinline__ void riemann3(const double UL[5], const double UR[5], double nx, double ny,
                                         double nz, double gamma, int flux, double FL[5], double Fh[5],
                                         bool &bad) {
  const double gm1 = gamma - 1.0;
  double FR[5], lL, lR;
  bool bl, br;
  euler_normal_flux3(UL, nx, ny, nz, gm1, FL, lL, bl);
  euler_normal_flux3(UR, nx, ny, nz, gm1, FR, lR, br);
  bad |= bl | br;
  if (flux == 1) {
    roe_flux3(UL, UR, nx, ny, nz, gamma, FL, Fh);
  } else {
    const double lmax = fmax(lL, lR);
#pragma unroll
    for (int v = 0; v < 5; ++v) Fh[v] = 0.5 * (FL[v] + FR[v] + lmax * (UL[v] - UR[v]));
  }
}

// 2D Roe flux without entropy fix in the face frame (src/Euler.cpp:682-741),
// returns the full numerical flux in Cartesian components.
__device__ __forceinline__ void roe_flux2(const double UL[4], const double UR[4], double nx,
                                          double ny, double gamma, const double FL[4],
                                          double Fh[4]) {
  const double gm1 = gamma - 1.0;
  const double rl = UL[0], rul = UL[1] * nx + UL[2] * ny, rvl = UL[2] * nx - UL[1] * ny, rEl = UL[3];
  const double rr = UR[0], rur = UR[1] * nx + UR[2] * ny, rvr = UR[2] * nx - UR[1] * ny, rEr = UR[3];
  const double ul = rul / rl, vl = rvl / rl, ur = rur / rr, vr = rvr / rr;
  const double pl = gm1 * (rEl - 0.5 * (rul * ul + rvl * vl));
  const double pr = gm1 * (rEr - 0.5 * (rur * ur + rvr * vr));
  const double Hl = (rEl + pl) / rl, Hr = (rEr + pr) / rr;
  const double R = sqrt(rr / rl), iRp1 = 1.0 / (1.0 + R);
  const double u = (ul + R * ur) * iRp1, v = (vl + R * vr) * iRp1, H = (Hl + R * Hr) * iRp1;
  const double a = sqrt(gm1 * (H - 0.5 * (u * u + v * v))), ainv = 1.0 / a;
  const double c1inv = a * a / gm1;
  double al[4], lm[4], G[4];
  al[2] = (rvr - rvl) - v * (rr - rl);
  al[1] = gm1 * ainv * ainv * ((rr - rl) * (H - u * u) + u * (rur - rul) -
                               ((rEr - rEl) - ((rvr - rvl) - v * (rr - rl)) * v));
  al[0] = 0.5 * ainv * ((rr - rl) * (u + a) - (rur - rul) - a * al[1]);
  al[3] = (rr - rl) - (al[0] + al[1]);
  lm[0] = fmin(u - a, 0.0); lm[1] = fmin(u, 0.0); lm[2] = lm[1]; lm[3] = fmin(u + a, 0.0);
  G[0] = lm[0] * al[0] + lm[1] * al[1] + lm[3] * al[3];
  G[1] = lm[0] * (u - a) * al[0] + lm[1] * u * al[1] + lm[3] * (u + a) * al[3];
  G[2] = lm[0] * v * al[0] + lm[1] * v * al[1] + lm[2] * al[2] + lm[3] * v * al[3];
  G[3] = lm[0] * (H - u * a) * al[0] + lm[1] * (H - c1inv) * al[1] + lm[2] * v * al[2] +
         lm[3] * (H + u * a) * al[3];
  Fh[0] = FL[0] + G[0];
  Fh[1] = FL[1] + G[1] * nx - G[2] * ny;
  Fh[2] = FL[2] + G[1] * ny + G[2] * nx;
  Fh[3] = FL[3] + G[3];
}

}  // namespace dgx
