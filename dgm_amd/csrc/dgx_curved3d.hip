// Curved hexes (Euler3d, p <= 8): the residual with metrics at the nodes and
// the projection through the element mass matrix.
//
//   volume term   Euler3d::convection with Hex::gradient on curved geometry:
//                 dF_x/dx = sum_d (dxi_d/dx)(node) D_d F_x at every node
//                 (src/Euler3d.cpp:468-575, src/Hex.cpp:2030-2071, 3992-4080)
//   faces         the Riemann flux with the master side's normal at each Gauss
//                 point (Euler3d::convective_flux evaluates a pair once, on its
//                 master side), lifted at the face nodes:
//                 G2GL((Fhat - F-.n) J_side/J_vol) / w_end (src/Hex.cpp:1786-1952)
//   project       b = B^T (wJ f) and M x = b with M = B^T diag(wJ) B, the curved
//                 element's mass matrix (src/Hex.cpp:289-318), solved by
//                 preconditioned conjugate gradients inside the workgroup.  The
//                 preconditioner is the exact inverse for a separable fit
//                 J ~ ja(xi) jb(eta) jc(zeta): the tensor product of three
//                 L x L inverses per element, built on the host.  On smoothly
//                 curved elements the fit leaves M's spectrum within ~1e-5 of
//                 the preconditioner's, so two or three iterations reach rounding.
#include <type_traits>
#include <vector>

#include "dgx_kernels.hpp"

namespace dgx {

// Float copies of the basis tables for the mixed residual (FP32 contractions
// and derivatives; FP64 metrics, fluxes, state and the mass solve), filled
// beside c_tab at upload.  TB3<R, P> gives the tables of arithmetic type R.
static __constant__ float c_tabf3[kTabTotal];
template <typename R, int P> struct TB3;
template <int P> struct TB3<double, P> : CT<P> {};
template <int P> struct TB3<float, P> {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1, b = tab_offset(P);
  __device__ static __forceinline__ const float *BT() { return c_tabf3 + b; }
  __device__ static __forceinline__ const float *BL() { return c_tabf3 + (CT<P>::BL() - c_tab); }
  __device__ static __forceinline__ const float *GL() { return c_tabf3 + (CT<P>::GL() - c_tab); }
  __device__ static __forceinline__ const float *DN() { return c_tabf3 + (CT<P>::DN() - c_tab); }
};

template <int P> struct C3 {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1;
  static constexpr int L2 = L * L, L3 = L2 * L, Q2 = Q * Q, Q3 = Q2 * Q, QF2 = QF * QF;
  // 256 threads up to p = 6 (three workgroups fit a CU's LDS at p = 5); one
  // 512-thread workgroup per CU at p = 7, 8, where the LDS holds only one
  static constexpr int NT = P >= 7 ? 512 : 256, NW = NT / 64;
  static constexpr int NPT = (Q3 + NT - 1) / NT;          // nodes per thread
  static constexpr int NML = (5 * L3 + NT - 1) / NT;      // modal entries per thread (5 fields)
  // LDS (doubles): Fw [5][Q3] (nodal residual, then the CG direction); scratch
  // S1, S2 [SZ] (the largest intermediate: a transform of 5 fields or the 30
  // face blocks at the GLL points; S1 also holds the 3 flux components of the
  // volume term); G [GZ] (the face terms gF [5][6][QF2], then b, M p and the
  // preconditioned residual [5][L3]); the element's preconditioner [3][L][L];
  // reductions [NW][16] and the mass solve's per-field totals [32] (the same
  // in every thread, so read from LDS instead of held in registers).  The
  // nodal state sits in S2 through the volume term.
  static constexpr int SZ = (5 * Q3 > 30 * Q2) ? 5 * Q3 : 30 * Q2;
  static constexpr int GZ = (30 * QF2 > 5 * L3) ? 30 * QF2 : 5 * L3;
  static constexpr int lds = (5 * Q3 + 2 * SZ + GZ + 3 * L2 + 16 * NW + 32) * 8;
};

// [NF][N]^3 blocks (compact) through matrix M [NO][N] along all three dimensions
template <int NF, int N, int NO, int NT, typename T>
__device__ __forceinline__ void chain3(const T *M, const T *in, T *s1, T *s2, T *out, int tid) {
  contract<NF, N, N * N, 1, NO, NT>(M, in, s1, tid);             // [NF][NO][N N]
  __syncthreads();
  contract<NF * NO, N, N, 1, NO, NT>(M, s1, s2, tid);            // [NF NO][NO][N]
  __syncthreads();
  contract<NF * NO * NO, N, 1, 1, NO, NT>(M, s2, out, tid);      // [NF NO NO][NO]
  __syncthreads();
}

// N per-thread partial sums, reduced over the workgroup into dst[0 .. N) in
// LDS: each wave's shuffle sum, then the waves' sums in ascending order
template <int P, int N>
__device__ __forceinline__ void wg_sums_lds(const double (&part)[N], double *red, double *dst, int tid) {
  using S = C3<P>;
  static_assert(N <= 16, "reduction slots");
#pragma unroll
  for (int f = 0; f < N; ++f) {
    double t = part[f];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if ((tid & 63) == 0) red[(tid >> 6) * 16 + f] = t;
  }
  __syncthreads();
  if (tid < N) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < S::NW; ++k) t += red[k * 16 + tid];
    dst[tid] = t;
  }
  __syncthreads();
}

// AP = M Pv with M = B^T diag(wJ) B, on [NF][L3] (S1, S2 scratch)
template <int P, int NT, int NF>
__device__ __forceinline__ void mass_apply(const double *Pv, double *AP, double *S1, double *S2, const double *ge,
                                           int tid) {
  constexpr int L = P + 1, Q = P + 2, Q3 = Q * Q * Q;
  contract<NF, L, L * L, 1, Q, NT>(CT<P>::BT(), Pv, S1, tid);        // [NF][Q][L L]
  __syncthreads();
  contract<NF * Q, L, L, 1, Q, NT>(CT<P>::BT(), S1, S2, tid);        // [NF Q][Q][L]
  __syncthreads();
  contract<NF * Q * Q, L, 1, 1, Q, NT>(CT<P>::BT(), S2, S1, tid);    // nodal [NF][Q3]
  __syncthreads();
  for (int i = tid; i < NF * Q3; i += NT) S1[i] *= ge[9 * Q3 + (i % Q3)];
  __syncthreads();
  contract<NF, Q, Q * Q, 1, L, NT>(CT<P>::BL(), S1, S2, tid);        // [NF][L][Q Q]
  __syncthreads();
  contract<NF * L, Q, Q, 1, L, NT>(CT<P>::BL(), S2, S1, tid);        // [NF L][L][Q]
  __syncthreads();
  contract<NF * L * L, Q, 1, 1, L, NT>(CT<P>::BL(), S1, AP, tid);    // [NF][L3]
  __syncthreads();
}

// out = (Pa (x) Pb (x) Pc) in on [NF][L][L][L] (in and out may alias), s1, s2 scratch
template <int L, int NT, int NF>
__device__ __forceinline__ void precondition(const double *Pre, const double *in, double *s1, double *s2,
                                             double *out, int tid) {
  contract<NF, L, L * L, 1, L, NT>(Pre, in, s1, tid);                  // along a
  __syncthreads();
  contract<NF * L, L, L, 1, L, NT>(Pre + L * L, s1, s2, tid);          // along b
  __syncthreads();
  contract<NF * L * L, L, 1, 1, L, NT>(Pre + 2 * L * L, s2, out, tid); // along c
  __syncthreads();
}

// Preconditioned conjugate gradients on M x = b for NF fields at once, M =
// B^T diag(wJ) B, with the element's separable preconditioner P: x0 = P b,
// r0 = b - M x0, z = P r, p = z, each field with its own step lengths and
// stopping test |r|^2 <= 1e-30 |b|^2.  bvec holds this thread's entries of b
// (index tid + k NT of [NF][L3]), which Bv also holds in LDS; AP = M p, then
// r and z = P r, reuse Bv, and Pv holds the direction p.  Per-field totals in
// LDS, Tot: |b|^2 [0, NF), r.r [NF, 2 NF), r.z [2 NF, 3 NF), p.Ap [3 NF, 4 NF),
// the previous r.z [4 NF, 5 NF).  The solution is left in x (registers);
// returns the iteration count.
template <int P, int NF, int ML>
__device__ __forceinline__ int pcg_solve(const double (&bvec)[ML], double *Bv, double *Pv, double *S1, double *S2,
                                         const double *Pre, const double *ge, double *red, double *Tot,
                                         double (&x)[ML], int tid) {
  using S = C3<P>;
  constexpr int L = S::L, L3 = S::L3, NT = S::NT, N = NF * L3;
  static_assert(ML * NT >= N && 5 * NF <= 32, "entries per thread, totals");
  double *AP = Bv;
  double res[ML];
  {
    double part[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) part[f] = 0.0;
#pragma unroll
    for (int k = 0; k < ML; ++k) {
      const int g = tid + k * NT;
      if (g < N) part[g / L3] = fma(bvec[k], bvec[k], part[g / L3]);
    }
    wg_sums_lds<P, NF>(part, red, Tot, tid);
  }
  precondition<L, NT, NF>(Pre, Bv, S1, S2, Pv, tid);   // x0 = P b, into Pv
#pragma unroll
  for (int k = 0; k < ML; ++k) {
    const int g = tid + k * NT;
    x[k] = g < N ? Pv[g] : 0.0;
  }
  // AP = M x0, r0 = b - AP
  mass_apply<P, NT, NF>(Pv, AP, S1, S2, ge, tid);
#pragma unroll
  for (int k = 0; k < ML; ++k) {
    const int g = tid + k * NT;
    res[k] = 0.0;
    if (g < N) {
      res[k] = bvec[k] - AP[g];
      AP[g] = res[k];
    }
  }
  __syncthreads();
  // z = P r from AP holding r (each thread's own entries); r.r and r.z
  auto residual_step = [&]() {
    precondition<L, NT, NF>(Pre, AP, S1, S2, AP, tid);
    double part2[2 * NF];
#pragma unroll
    for (int f = 0; f < 2 * NF; ++f) part2[f] = 0.0;
#pragma unroll
    for (int k = 0; k < ML; ++k) {
      const int g = tid + k * NT;
      if (g < N) {
        const int f = g / L3;
        part2[f] = fma(res[k], res[k], part2[f]);
        part2[NF + f] = fma(res[k], AP[g], part2[NF + f]);
      }
    }
    wg_sums_lds<P, 2 * NF>(part2, red, Tot + NF, tid);
  };
  residual_step();
  if (tid < NF) Tot[4 * NF + tid] = Tot[2 * NF + tid];
#pragma unroll
  for (int k = 0; k < ML; ++k) {
    const int g = tid + k * NT;
    if (g < N) Pv[g] = AP[g];
  }
  __syncthreads();
  constexpr double tol2 = 1e-30;   // (1e-15 relative)^2
  int it = 0;
  for (; it < 60; ++it) {
    bool conv = true;
#pragma unroll
    for (int f = 0; f < NF; ++f) conv = conv && (Tot[NF + f] <= tol2 * Tot[f]);
    if (conv) break;
    mass_apply<P, NT, NF>(Pv, AP, S1, S2, ge, tid);
    {
      double part[NF];
#pragma unroll
      for (int f = 0; f < NF; ++f) part[f] = 0.0;
#pragma unroll
      for (int k = 0; k < ML; ++k) {
        const int g = tid + k * NT;
        if (g < N) part[g / L3] = fma(Pv[g], AP[g], part[g / L3]);
      }
      wg_sums_lds<P, NF>(part, red, Tot + 3 * NF, tid);
    }
#pragma unroll
    for (int k = 0; k < ML; ++k) {
      const int g = tid + k * NT;
      if (g < N) {
        const int f = g / L3;
        const double rz = Tot[4 * NF + f], pap = Tot[3 * NF + f];
        const double alpha = (rz > 0.0 && pap > 0.0) ? rz / pap : 0.0;
        x[k] = fma(alpha, Pv[g], x[k]);
        res[k] = fma(-alpha, AP[g], res[k]);
        AP[g] = res[k];
      }
    }
    __syncthreads();
    residual_step();
#pragma unroll
    for (int k = 0; k < ML; ++k) {
      const int g = tid + k * NT;
      if (g < N) {
        const int f = g / L3;
        const double rz = Tot[4 * NF + f];
        const double beta = rz > 0.0 ? Tot[2 * NF + f] / rz : 0.0;
        Pv[g] = fma(beta, Pv[g], AP[g]);
      }
    }
    __syncthreads();
    if (tid < NF) Tot[4 * NF + tid] = Tot[2 * NF + tid];
    __syncthreads();
  }
  return it;
}

// PROJ: project the nodal fields a.U [5][ne][Q^3] instead (the forward
// transform of a curved element: x = M^-1 B^T (wJ u)).
// Three workgroups fit a CU's LDS at p = 5, and the kernel asks the compiler
// for three waves per SIMD (at most 168 VGPRs; it needs about 150 at p = 5,
// with no spills).  At p = 7, 8 one 512-thread workgroup per CU gives two
// waves per SIMD.
// R: arithmetic of the residual before the mass solve (its element tensors,
// contractions and derivatives): double, or float on the mixed path, whose
// LDS views alias the same memory as the FP64 buffers of the mass solve.
// NS: Navier-Stokes (BR1), adding the viscous terms k_grad3c and k_visc3c
// wrote: the weighted nodal divergence V is subtracted before the projection,
// and each interior side's correction gains J_side/J_vol (1/2)(F_v.n- + F_v+.n+)
// (the BR1 face flux is the average of the two sides' own normal fluxes; a
// boundary side's ghost flux is -F_v, so its face flux is its own).
template <int P, bool PROJ, typename R = double, bool NS = false>
__global__ void __launch_bounds__(C3<P>::NT) __attribute__((amdgpu_waves_per_eu(P >= 7 ? 2 : 3)))
k_rhs3c(CurvedArgs a) {
  using S = C3<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L3 = S::L3, Q2 = S::Q2, Q3 = S::Q3, QF2 = S::QF2;
  constexpr int NT = S::NT, NPT = S::NPT, NML = S::NML;
  extern __shared__ double sm[];
  double *Fw = sm, *S1 = Fw + 5 * Q3, *S2 = S1 + S::SZ, *G = S2 + S::SZ, *Pre = G + S::GZ, *red = Pre + 3 * S::L2;
  double *Bv = G;
  // the residual's views (type R) of the same buffers
  R *Fr = reinterpret_cast<R *>(Fw), *S1r = reinterpret_cast<R *>(S1), *S2r = reinterpret_cast<R *>(S2);
  R *gF = reinterpret_cast<R *>(G), *Fc = S1r, *Bvr = reinterpret_cast<R *>(G);
  static_assert(30 * QF * Q <= S::SZ && 5 * L3 <= S::SZ && 3 * Q3 <= S::SZ && 5 * Q * L * L <= S::SZ, "scratch");
  static_assert(S::lds <= 160 * 1024, "LDS");
  const int ne = a.ne, e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;
  const Geo &geo = a.geo;
  const double *ge = a.gm + (size_t)e * 10 * Q3;   // wJ dxi_d/dx_x (9 rows), wJ
  const double gm1 = a.phys.gamma - 1.0;
  for (int i = tid; i < 3 * S::L2; i += NT) Pre[i] = a.pre[(size_t)e * 3 * S::L2 + i];

  if (PROJ) {
    for (int i = tid; i < 5 * Q3; i += NT) {
      const int v = i / Q3, n = i - v * Q3;
      Fr[i] = (R)(ge[9 * Q3 + n] * a.U[((size_t)v * ne + e) * Q3 + n]);
    }
    __syncthreads();
  } else {
  // the modal state's loads are issued first; they land in LDS after the face
  // loop, whose dependent loads (links, orientation, neighbour traces) they overlap
  double um[NML];
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int i = tid + k * NT, v = i / L3;
    um[k] = i < 5 * L3 ? a.U[((size_t)v * ne + e) * L3 + (i - v * L3)] : 0.0;
  }
  // face corrections g = (J_side/J_vol)(Fhat - F(U-).n) at the Gauss points,
  // before the nodal state and the metric terms are live (the Riemann
  // solver's temporaries and those would not fit in registers beside each other)
  bool bad = false;
  for (int i = tid; i < 6 * QF2; i += NT) {
    const int s = i / QF2, q = i - s * QF2;
    const int es = e * 6 + s;
    const double *fg = a.fgeo + ((size_t)es * QF2 + q) * 4;
    double nx = fg[0], ny = fg[1], nz = fg[2];
    const double jr = fg[3];
    const int nb = geo.nbr[es], ns = geo.nside[es];
    double UL[5], UR[5], FL[5], Fh[5];
    const double *tl = a.T + ((size_t)e * 6 + s) * 5 * QF2;
#pragma unroll
    for (int v = 0; v < 5; ++v) UL[v] = tl[v * QF2 + q];
    if (nb >= 0) {
      const int qn = geo.perm[geo.permid[es] * QF2 + q];
      if (!geo.master[es]) {   // the pair's flux uses the master side's normal (both sides' normals
                               // come from their own interpolated metrics and differ slightly)
        const double *fm = a.fgeo + (((size_t)nb * 6 + ns) * QF2 + qn) * 4;
        nx = -fm[0]; ny = -fm[1]; nz = -fm[2];
      }
      const double *tr = a.T + ((size_t)nb * 6 + ns) * 5 * QF2;
#pragma unroll
      for (int v = 0; v < 5; ++v) UR[v] = tr[v * QF2 + qn];
    } else {
      ghost3(-1 - nb, UL, nx, ny, nz, a.phys.bcU, UR);
    }
    riemann3(UL, UR, nx, ny, nz, a.phys.gamma, a.phys.flux, FL, Fh, bad);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      double g = jr * (Fh[v] - FL[v]);
      if (NS && v > 0 && nb >= 0) {
        const double *vl = a.TV + ((size_t)e * 6 + s) * 4 * QF2, *vr = a.TV + ((size_t)nb * 6 + ns) * 4 * QF2;
        g += 0.5 * jr * (vl[(v - 1) * QF2 + q] + vr[(v - 1) * QF2 + geo.perm[geo.permid[es] * QF2 + q]]);
      }
      gF[(v * 6 + s) * QF2 + q] = (R)g;
    }
  }
  if (bad) atomicOr(a.flag, 1);


  // modal state -> nodal (5 fields), held in S2 through the volume term
  // (S2 is free until the face lift; registers are not)
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int i = tid + k * NT;
    if (i < 5 * L3) S1r[i] = (R)um[k];
  }
  __syncthreads();
  chain3<5, L, Q, NT>(TB3<R, P>::BT(), S1r, Fr, S1r, S2r, tid);
  const R *Un = S2r;

  // volume: wJ * sum_x sum_d (dxi_d/dx_x) D_d F_x at every node, field by
  // field; each thread's nodes keep their 9 metric terms in registers
  double mg[NPT][9];
#pragma unroll
  for (int rr = 0; rr < NPT; ++rr) {
    const int n = tid + rr * NT;
#pragma unroll
    for (int c = 0; c < 9; ++c) mg[rr][c] = n < Q3 ? ge[c * Q3 + n] : 0.0;
  }
  for (int v = 0; v < 5; ++v) {
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
      if (n >= Q3) continue;
      const double r = Un[n], mx = Un[Q3 + n], my = Un[2 * Q3 + n], mz = Un[3 * Q3 + n], E = Un[4 * Q3 + n];
      const double ir = 1.0 / r, u = mx * ir, vv = my * ir, w = mz * ir;
      const double p = gm1 * (E - 0.5 * (mx * u + my * vv + mz * w));
      double fx, fy, fz;
      if (v == 0) { fx = mx; fy = my; fz = mz; }
      else if (v == 1) { fx = mx * u + p; fy = mx * vv; fz = mx * w; }
      else if (v == 2) { fx = my * u; fy = my * vv + p; fz = my * w; }
      else if (v == 3) { fx = mz * u; fy = mz * vv; fz = mz * w + p; }
      else { const double h = E + p; fx = h * u; fy = h * vv; fz = h * w; }
      Fc[n] = (R)fx; Fc[Q3 + n] = (R)fy; Fc[2 * Q3 + n] = (R)fz;
    }
    __syncthreads();
    const R *Dn = TB3<R, P>::DN();
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
      if (n < Q3) {
        const int i = n / Q2, j = (n / Q) % Q, k = n % Q;
        double acc = 0.0;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          const R *F = Fc + x * Q3;
          R dr = 0, ds = 0, dt = 0;
#pragma unroll
          for (int m = 0; m < Q; ++m) {
            dr = fma(Dn[i * Q + m], F[(m * Q + j) * Q + k], dr);
            ds = fma(Dn[j * Q + m], F[(i * Q + m) * Q + k], ds);
            dt = fma(Dn[k * Q + m], F[(i * Q + j) * Q + m], dt);
          }
          acc = fma(mg[rr][x], (double)dr, acc);
          acc = fma(mg[rr][3 + x], (double)ds, acc);
          acc = fma(mg[rr][6 + x], (double)dt, acc);
        }
        if (NS && v > 0) acc -= a.V[((size_t)e * 4 + v - 1) * Q3 + n];
        Fr[v * Q3 + n] = (R)acc;
      }
    }
  }
  __syncthreads();
  // faces: Gauss -> GLL face nodes for 30 blocks, then wJ / w_end at the face nodes
  contract<30, QF, QF, 1, Q, NT>(TB3<R, P>::GL(), gF, S1r, tid);   // [30][Q][QF]
  __syncthreads();
  contract<30, Q, QF, 2, Q, NT>(TB3<R, P>::GL(), S1r, S2r, tid);    // [30][Q][Q]
  __syncthreads();
  {
    const double wend = CT<P>::BW()[0] / CT<P>::BT()[0];   // GLL end weight: phi_0 w_0 / phi_0
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
      if (n < Q3) {
        const int i = n / Q2, j = (n / Q) % Q, k = n % Q;
        const double c = ge[9 * Q3 + n] / wend;
        // the faces the node lies on, in a fixed order, each with a constant
        // face (no indexed arrays, which would live in scratch)
        auto add = [&](int f, int a1, int a2) {
#pragma unroll
          for (int v = 0; v < 5; ++v)
            Fr[v * Q3 + n] = fma((R)c, S2r[((v * 6 + f) * Q + a1) * Q + a2], Fr[v * Q3 + n]);
        };
        if (k == 0) add(0, i, j);
        if (k == Q - 1) add(5, i, j);
        if (j == 0) add(1, i, k);
        if (j == Q - 1) add(3, i, k);
        if (i == 0) add(4, j, k);
        if (i == Q - 1) add(2, j, k);
      }
    }
  }
  __syncthreads();
  }   // !PROJ
  // b = B^T (wJ f): into Bv (compact [5][L3], the face-term region, free now)
  chain3<5, Q, L, NT>(TB3<R, P>::BL(), Fr, S1r, S2r, Bvr, tid);
  // the mass solve: Pv (compact [5][L3]) reuses Fw; totals after the reductions.
  // The update's RK register is loaded first: the solve's barriers cover it.
  double x[NML], utin[NML];
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int g = tid + k * NT, v = g / L3;
    utin[k] = (!PROJ && a.update && g < 5 * L3) ? a.Ut[((size_t)v * ne + e) * L3 + (g - v * L3)] : 0.0;
  }
  int it;
  {
    double bvec[NML];
#pragma unroll
    for (int k = 0; k < NML; ++k) {
      const int g = tid + k * NT;
      bvec[k] = g < 5 * L3 ? (double)Bvr[g] : 0.0;
    }
    if (!std::is_same<R, double>::value) {   // b to the FP64 view of the same buffer
      __syncthreads();
#pragma unroll
      for (int k = 0; k < NML; ++k) {
        const int g = tid + k * NT;
        if (g < 5 * L3) Bv[g] = bvec[k];
      }
    }
    it = pcg_solve<P, 5>(bvec, Bv, Fw, S1, S2, Pre, ge, red, red + 16 * S::NW, x, tid);
  }
  // the residual, or (update) the LSRK stage in place: this element's modes
  // were read at the start and no other workgroup reads them
  const double mdt = -a.dt;
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int g = tid + k * NT;
    if (g < 5 * L3) {
      const int v = g / L3;
      const size_t o = ((size_t)v * ne + e) * L3 + (g - v * L3);
      if (!PROJ && a.update) {
        const double ut = fma(mdt, x[k], a.ca * utin[k]);   // as the axpby Ut = ca Ut - dt R
        a.Ut[o] = ut;
        a.Us[o] = fma(a.cb, ut, a.Us[o]);                   // as the axpy U += cb Ut
      } else {
        a.R[o] = x[k];
      }
    }
  }
  if (tid == 0 && a.iters) atomicMax(a.iters, it);
}


// Navier-Stokes (BR1) on curved hexes, in the order of Navier_Stokes3d::diffusion
// (src/Navier_Stokes3d.cpp:113-258): k_grad3c lifts and projects the gradients
// of the primitives (u, v, w, e), k_visc3c forms the stresses and heat flux
// and writes their weighted divergence and each side's normal flux, and
// k_rhs3c<P, false, double, true> adds both to the Euler residual before its
// one projection.
template <int P> struct C3V {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1;
  static constexpr int L2 = L * L, L3 = L2 * L, Q2 = Q * Q, Q3 = Q2 * Q, QF2 = QF * QF;
  static constexpr int NT = C3<P>::NT, NW = NT / 64, NPT = C3<P>::NPT;
  static constexpr int mx(int x, int y) { return x > y ? x : y; }
  // k_grad3c solves NG = 3 NPR gradient components at once: two primitives
  // per pass (6 components, 2 passes) where the LDS holds them, else one
  // (3 components, 4 passes)
  static constexpr int sz(int ng) {
    return mx(mx(5 * Q2 * L, 6 * ng * Q2), ng * Q3);   // state transform, face blocks, mass operator
  }
  // k_grad3c (doubles): Pq [4][Q3] primitives at the nodes (the first 5 Q3
  // of Pq and Fw hold the nodal state before); Fw [NG][Q3] the pass's
  // weighted gradients, then the CG direction; S1, S2 [SZ]; Jq [4][6][QF2]
  // the jumps; Bv [NG][L3]; the preconditioner [3][L][L]; reductions; totals
  static constexpr int grad_lds(int ng) {
    return ((4 + ng) * Q3 + 2 * sz(ng) + 24 * QF2 + ng * L3 + 3 * L2 + 16 * NW + 32) * 8;
  }
  static constexpr int NPR = grad_lds(6) <= 96 * 1024 ? 2 : 1, NG = 3 * NPR;
  static constexpr int SZ = sz(NG), grad = grad_lds(NG);
  static constexpr int ML = (NG * L3 + NT - 1) / NT;      // mass-solve entries per thread
  // k_visc3c: Un [5][Q3] the nodal state; Fc [3][Q3] one field's viscous
  // flux (first the gradients at the nodes); S1, S2 [SV]
  static constexpr int SV = mx(5 * Q2 * L, 18 * Q2);
  static constexpr int visc = (8 * Q3 + 2 * SV) * 8;
};

bool curved3_ns_supported(int p) { return p >= 1 && p <= 7; }

// Lifted, projected gradients of (u, v, w, e) of a curved hex (the
// Euler3d::gradient_lift path, src/Euler3d.cpp:1144-1199): at the nodes
// wJ dq/dx_x = sum_d (wJ dxi_d/dx_x) D_d q, plus at the face nodes
// (wJ / w_end) G2GL(J_side/J_vol ({q} - q-) n_x) with the side's own normal,
// then M^-1 B^T of that.  G [ne][12][L3] in the order (u_x, u_y, u_z, v_x,
// ..., e_z).  A boundary side's ghost primitives are the isothermal wall's or
// the imposed state's (ghost_prim3).
template <int P>
__global__ void __launch_bounds__(C3<P>::NT)
k_grad3c(CurvedArgs a) {
  using S = C3V<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L3 = S::L3, Q2 = S::Q2, Q3 = S::Q3, QF2 = S::QF2;
  constexpr int NT = S::NT, NPT = S::NPT, ML = S::ML, NPR = S::NPR, NG = S::NG;
  extern __shared__ double sm[];
  double *Pq = sm, *Fw = Pq + 4 * Q3, *S1 = Fw + NG * Q3, *S2 = S1 + S::SZ, *Jq = S2 + S::SZ, *Bv = Jq + 24 * QF2;
  double *Pre = Bv + NG * L3, *red = Pre + 3 * S::L2, *Tot = red + 16 * S::NW;
  static_assert(5 * Q * S::L2 <= S::SZ && NG * Q3 <= S::SZ && 6 * NG * Q * QF <= S::SZ && 5 * NG <= 32, "scratch");
  static_assert(S::grad <= 160 * 1024, "LDS");
  const int ne = a.ne, e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;
  const Geo &geo = a.geo;
  const double *ge = a.gm + (size_t)e * 10 * Q3;
  for (int i = tid; i < 3 * S::L2; i += NT) Pre[i] = a.pre[(size_t)e * 3 * S::L2 + i];
  // jumps J_side/J_vol ({q} - q-) of the primitives at the Gauss points
  for (int i = tid; i < 6 * QF2; i += NT) {
    const int s = i / QF2, q = i - s * QF2;
    const int es = e * 6 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const double jr = a.fgeo[((size_t)es * QF2 + q) * 4 + 3];
    const double *tl = a.T + ((size_t)es) * 5 * QF2;
    double pl[4], pr[4];
    {
      const double r = tl[q], u = tl[QF2 + q] / r, v = tl[2 * QF2 + q] / r, w = tl[3 * QF2 + q] / r;
      double en = tl[4 * QF2 + q] / r;
      double wk = u * u; wk = v * v + wk; wk = w * w + wk;
      pl[0] = u; pl[1] = v; pl[2] = w; pl[3] = -0.5 * wk + en;
    }
    if (nb < 0) {
      ghost_prim3(-1 - nb, tl[q], a.phys, pr);
    } else {
      const int qn = geo.perm[geo.permid[es] * QF2 + q];
      const double *tr = a.T + ((size_t)nb * 6 + ns) * 5 * QF2;
      const double r = tr[qn], u = tr[QF2 + qn] / r, v = tr[2 * QF2 + qn] / r, w = tr[3 * QF2 + qn] / r;
      double en = tr[4 * QF2 + qn] / r;
      double wk = u * u; wk = v * v + wk; wk = w * w + wk;
      pr[0] = u; pr[1] = v; pr[2] = w; pr[3] = -0.5 * wk + en;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) Jq[(c * 6 + s) * QF2 + q] = jr * (0.5 * (pl[c] + pr[c]) - pl[c]);
  }
  // modal state -> nodal [5][Q3] across Pq and Fw, then the primitives in place
  for (int i = tid; i < 5 * L3; i += NT) {
    const int v = i / L3;
    S1[i] = a.U[((size_t)v * ne + e) * L3 + (i - v * L3)];
  }
  __syncthreads();
  chain3<5, L, Q, NT>(CT<P>::BT(), S1, S2, S1, Pq, tid);
  double mg[NPT][9];
#pragma unroll
  for (int rr = 0; rr < NPT; ++rr) {
    const int n = tid + rr * NT;
#pragma unroll
    for (int c = 0; c < 9; ++c) mg[rr][c] = n < Q3 ? ge[c * Q3 + n] : 0.0;
    if (n < Q3) {
      const double r = Pq[n], u = Pq[Q3 + n] / r, v = Pq[2 * Q3 + n] / r, w = Pq[3 * Q3 + n] / r;
      double en = Pq[4 * Q3 + n] / r;
      double wk = u * u; wk = v * v + wk; wk = w * w + wk;
      Pq[n] = u; Pq[Q3 + n] = v; Pq[2 * Q3 + n] = w; Pq[3 * Q3 + n] = -0.5 * wk + en;
    }
  }
  const double *Dn = CT<P>::DN();
  const double wend = CT<P>::BW()[0] / CT<P>::BT()[0];   // GLL end weight
  for (int c0 = 0; c0 < 4; c0 += NPR) {
    __syncthreads();
    // weighted nodal gradients of primitives c0 .. c0 + NPR - 1 into Fw [NG][Q3]
#pragma unroll 1
    for (int pc = 0; pc < NPR; ++pc) {
      const double *F = Pq + (c0 + pc) * Q3;
#pragma unroll
      for (int rr = 0; rr < NPT; ++rr) {
        const int n = tid + rr * NT;
        if (n < Q3) {
          const int i = n / Q2, j = (n / Q) % Q, k = n % Q;
          double dr = 0.0, ds = 0.0, dt = 0.0;
#pragma unroll
          for (int m = 0; m < Q; ++m) {
            dr = fma(Dn[i * Q + m], F[(m * Q + j) * Q + k], dr);
            ds = fma(Dn[j * Q + m], F[(i * Q + m) * Q + k], ds);
            dt = fma(Dn[k * Q + m], F[(i * Q + j) * Q + m], dt);
          }
#pragma unroll
          for (int x = 0; x < 3; ++x)
            Fw[(3 * pc + x) * Q3 + n] = fma(mg[rr][6 + x], dt, fma(mg[rr][3 + x], ds, mg[rr][x] * dr));
        }
      }
    }
    // face blocks n_x J_side/J_vol ({q} - q-) [NG][6][QF2] -> GLL face nodes
    for (int i = tid; i < 6 * NG * QF2; i += NT) {
      const int b = i / QF2, q = i - b * QF2, g = b / 6, s = b - g * 6, c = c0 + g / 3, x = g - (g / 3) * 3;
      S2[i] = a.fgeo[(((size_t)e * 6 + s) * QF2 + q) * 4 + x] * Jq[(c * 6 + s) * QF2 + q];
    }
    __syncthreads();
    contract<6 * NG, QF, QF, 1, Q, NT>(CT<P>::GL(), S2, S1, tid);   // [6 NG][Q][QF]
    __syncthreads();
    contract<6 * NG, Q, QF, 2, Q, NT>(CT<P>::GL(), S1, S2, tid);    // [6 NG][Q][Q]
    __syncthreads();
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
      if (n < Q3) {
        const int i = n / Q2, j = (n / Q) % Q, k = n % Q;
        const double cw = ge[9 * Q3 + n] / wend;
        auto add = [&](int f, int a1, int a2) {
#pragma unroll
          for (int g = 0; g < NG; ++g)
            Fw[g * Q3 + n] = fma(cw, S2[((g * 6 + f) * Q + a1) * Q + a2], Fw[g * Q3 + n]);
        };
        if (k == 0) add(0, i, j);
        if (k == Q - 1) add(5, i, j);
        if (j == 0) add(1, i, k);
        if (j == Q - 1) add(3, i, k);
        if (i == 0) add(4, j, k);
        if (i == Q - 1) add(2, j, k);
      }
    }
    __syncthreads();
    // b = B^T (wJ g), then M x = b
    chain3<NG, Q, L, NT>(CT<P>::BL(), Fw, S1, S2, Bv, tid);
    double bvec[ML], x[ML];
#pragma unroll
    for (int k = 0; k < ML; ++k) {
      const int g = tid + k * NT;
      bvec[k] = g < NG * L3 ? Bv[g] : 0.0;
    }
    const int it = pcg_solve<P, NG>(bvec, Bv, Fw, S1, S2, Pre, ge, red, Tot, x, tid);
#pragma unroll
    for (int k = 0; k < ML; ++k) {
      const int g = tid + k * NT;
      if (g < NG * L3) a.G[((size_t)e * 12 + 3 * c0) * L3 + g] = x[k];
    }
    if (tid == 0 && a.iters) atomicMax(a.iters, it);
  }
}

// Viscous terms of a curved hex from the state and the lifted gradients G
// (src/Navier_Stokes3d.cpp:113-258): the stresses and heat flux at the nodes;
// V [ne][4][Q3] = wJ div F_v for (rho u, rho v, rho w, rho E), and each side's
// normal flux F_v.n with its own normal at its Gauss points into TV
// [ne][6][4][QF2] (the three Cartesian components interpolated to the side
// first, then dotted, as the reference forms the side flux).
template <int P>
__global__ void __launch_bounds__(C3<P>::NT)
k_visc3c(CurvedArgs a) {
  using S = C3V<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L3 = S::L3, Q2 = S::Q2, Q3 = S::Q3, QF2 = S::QF2;
  constexpr int NT = S::NT, NPT = S::NPT;
  extern __shared__ double sm[];
  double *Un = sm, *Fc = Un + 5 * Q3, *S1 = Fc + 3 * Q3, *S2 = S1 + S::SV;
  static_assert(S::visc <= 160 * 1024, "LDS");
  const int ne = a.ne, e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;
  const double *ge = a.gm + (size_t)e * 10 * Q3;
  for (int i = tid; i < 5 * L3; i += NT) {
    const int v = i / L3;
    S1[i] = a.U[((size_t)v * ne + e) * L3 + (i - v * L3)];
  }
  __syncthreads();
  chain3<5, L, Q, NT>(CT<P>::BT(), S1, S2, S1, Un, tid);
  // the 12 gradient components at this thread's nodes
  double gq[NPT][12];
  for (int c = 0; c < 4; ++c) {
    for (int i = tid; i < 3 * L3; i += NT) S1[i] = a.G[((size_t)e * 12 + 3 * c) * L3 + i];
    __syncthreads();
    chain3<3, L, Q, NT>(CT<P>::BT(), S1, S2, S1, Fc, tid);
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
#pragma unroll
      for (int x = 0; x < 3; ++x) gq[rr][3 * c + x] = n < Q3 ? Fc[x * Q3 + n] : 0.0;
    }
    __syncthreads();
  }
  double mg[NPT][9];
#pragma unroll
  for (int rr = 0; rr < NPT; ++rr) {
    const int n = tid + rr * NT;
#pragma unroll
    for (int c = 0; c < 9; ++c) mg[rr][c] = n < Q3 ? ge[c * Q3 + n] : 0.0;
  }
  const double mu = a.phys.vis, gp = a.phys.gamma / a.phys.Pr;
  const double *Dn = CT<P>::DN();
  for (int f = 0; f < 4; ++f) {
    // the field's viscous flux (F_x, F_y, F_z) at the nodes
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
      if (n >= Q3) continue;
      const double *g = gq[rr];
      const double ux = g[0], uy = g[1], uz = g[2], vx = g[3], vy = g[4], vz = g[5];
      const double wx = g[6], wy = g[7], wz = g[8];
      const double t12 = mu * (uy + vx), t13 = mu * (uz + wx), t23 = mu * (vz + wy);
      double fx, fy, fz;
      if (f == 0) {
        fx = mu * ((2.0 / 3.0) * ((2.0 * ux - vy) - wz)); fy = t12; fz = t13;
      } else if (f == 1) {
        fx = t12; fy = mu * ((2.0 / 3.0) * ((2.0 * vy - ux) - wz)); fz = t23;
      } else if (f == 2) {
        fx = t13; fy = t23; fz = mu * ((2.0 / 3.0) * ((2.0 * wz - ux) - vy));
      } else {
        const double t11 = mu * ((2.0 / 3.0) * ((2.0 * ux - vy) - wz));
        const double t22 = mu * ((2.0 / 3.0) * ((2.0 * vy - ux) - wz));
        const double t33 = mu * ((2.0 / 3.0) * ((2.0 * wz - ux) - vy));
        const double r = Un[n], u = Un[Q3 + n] / r, v = Un[2 * Q3 + n] / r, w = Un[3 * Q3 + n] / r;
        fx = fma(gp, mu * g[9], fma(t13, w, fma(t12, v, t11 * u)));
        fy = fma(gp, mu * g[10], fma(t23, w, fma(t22, v, t12 * u)));
        fz = fma(gp, mu * g[11], fma(t33, w, fma(t23, v, t13 * u)));
      }
      Fc[n] = fx; Fc[Q3 + n] = fy; Fc[2 * Q3 + n] = fz;
    }
    __syncthreads();
    // wJ div F_v at the nodes
#pragma unroll
    for (int rr = 0; rr < NPT; ++rr) {
      const int n = tid + rr * NT;
      if (n < Q3) {
        const int i = n / Q2, j = (n / Q) % Q, k = n % Q;
        double acc = 0.0;
#pragma unroll
        for (int x = 0; x < 3; ++x) {
          const double *F = Fc + x * Q3;
          double dr = 0.0, ds = 0.0, dt = 0.0;
#pragma unroll
          for (int m = 0; m < Q; ++m) {
            dr = fma(Dn[i * Q + m], F[(m * Q + j) * Q + k], dr);
            ds = fma(Dn[j * Q + m], F[(i * Q + m) * Q + k], ds);
            dt = fma(Dn[k * Q + m], F[(i * Q + j) * Q + m], dt);
          }
          acc = fma(mg[rr][x], dr, acc);
          acc = fma(mg[rr][3 + x], ds, acc);
          acc = fma(mg[rr][6 + x], dt, acc);
        }
        a.V[((size_t)e * 4 + f) * Q3 + n] = acc;
      }
    }
    // the face nodes of (F_x, F_y, F_z) [6][3][Q][Q] -> the Gauss points
    for (int i = tid; i < 18 * Q2; i += NT) {
      const int b = i / Q2, t = i - b * Q2, s = b / 3, x = b - s * 3, t1 = t / Q, t2 = t - t1 * Q;
      int n;
      switch (s) {
        case 0: n = (t1 * Q + t2) * Q; break;
        case 5: n = (t1 * Q + t2) * Q + Q - 1; break;
        case 1: n = (t1 * Q) * Q + t2; break;
        case 3: n = (t1 * Q + Q - 1) * Q + t2; break;
        case 4: n = t1 * Q + t2; break;
        default: n = ((Q - 1) * Q + t1) * Q + t2; break;
      }
      S1[i] = Fc[x * Q3 + n];
    }
    __syncthreads();
    contract<18, Q, Q, 1, QF, NT>(CT<P>::Ig(), S1, S2, tid);    // [18][QF][Q]
    __syncthreads();
    contract<18, QF, Q, 2, QF, NT>(CT<P>::Ig(), S2, S1, tid);   // [18][QF][QF]
    __syncthreads();
    for (int i = tid; i < 6 * QF2; i += NT) {
      const int s = i / QF2, q = i - s * QF2;
      const double *fg = a.fgeo + (((size_t)e * 6 + s) * QF2 + q) * 4;
      const double *W = S1 + (s * 3) * QF2 + q;
      a.TV[(((size_t)e * 6 + s) * 4 + f) * QF2 + q] = fma(fg[2], W[2 * QF2], fma(fg[1], W[QF2], fg[0] * W[0]));
    }
    __syncthreads();
  }
}

template <int P>
static hipError_t launch_c3(const CurvedArgs &a) {
  const int sh = C3<P>::lds;
  if (a.project) {   // the projection of given nodal data stays FP64 on both paths
    (void)hipFuncSetAttribute((const void *)k_rhs3c<P, true>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
    hipLaunchKernelGGL((k_rhs3c<P, true>), dim3(a.ne), dim3(C3<P>::NT), sh, a.stream, a);
  } else if (a.ns) {
    if constexpr (P <= 7) {
      if (a.op == 1) {
        (void)hipFuncSetAttribute((const void *)k_grad3c<P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  C3V<P>::grad);
        hipLaunchKernelGGL(k_grad3c<P>, dim3(a.ne), dim3(C3<P>::NT), C3V<P>::grad, a.stream, a);
      } else if (a.op == 2) {
        (void)hipFuncSetAttribute((const void *)k_visc3c<P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  C3V<P>::visc);
        hipLaunchKernelGGL(k_visc3c<P>, dim3(a.ne), dim3(C3<P>::NT), C3V<P>::visc, a.stream, a);
      } else {
        (void)hipFuncSetAttribute((const void *)k_rhs3c<P, false, double, true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, sh);
        hipLaunchKernelGGL((k_rhs3c<P, false, double, true>), dim3(a.ne), dim3(C3<P>::NT), sh, a.stream, a);
      }
    } else {
      return hipErrorInvalidValue;
    }
  } else if (a.mixed) {
    (void)hipFuncSetAttribute((const void *)k_rhs3c<P, false, float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              sh);
    hipLaunchKernelGGL((k_rhs3c<P, false, float>), dim3(a.ne), dim3(C3<P>::NT), sh, a.stream, a);
  } else {
    (void)hipFuncSetAttribute((const void *)k_rhs3c<P, false>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
    hipLaunchKernelGGL((k_rhs3c<P, false>), dim3(a.ne), dim3(C3<P>::NT), sh, a.stream, a);
  }
  return hipGetLastError();
}

bool curved3_supported(int p) { return p >= 1 && p <= 8; }

hipError_t launch_curved3(const CurvedArgs &a) {
  switch (a.p) {
#define DGX_CASE(PP) case PP: return launch_c3<PP>(a);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5) DGX_CASE(6) DGX_CASE(7) DGX_CASE(8)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t upload_tables_c3d(const double *h) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), h, sizeof(double) * kTabTotal);
  if (e != hipSuccess) return e;
  std::vector<float> f(h, h + kTabTotal);
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tabf3), f.data(), sizeof(float) * kTabTotal);
}

}  // namespace dgx
