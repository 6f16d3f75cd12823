// Fused hex stage kernels (p <= 9).
//
// One workgroup of 512 threads per element for both kernels, two workgroups
// per CU (4 waves per SIMD, at most 128 VGPRs).  Every sum-factorized
// contraction runs on all conserved fields of the element at once (pencils of
// 5 fields x q^2 fill the workgroup instead of q^2 = 81 pencils at p = 7),
// nodal values live in registers (one thread per node), and the last
// contraction of each chain accumulates modal results in registers, so a
// workgroup needs only two LDS tensors of about 5 q^3 doubles.  The last
// dimension of every LDS tensor is padded to an odd pitch, which keeps the
// 64-bit LDS accesses of a wavefront on distinct banks.
//
//   k_stage3   residual of Euler3d::convection + project (src/Euler3d.cpp:468-575,
//              584-817; src/Hex.cpp:1786-1952) plus, for Navier-Stokes, the BR1
//              viscous face flux and the viscous volume term from k_vterm3;
//              optionally fused with the LSRK14 register update
//              (src/TimeInt.cpp:1989-2013) and the face traces of the new state
//              (Hex::fill_sides, src/Hex.cpp:933-947) for the next stage.
//   k_vterm3   BR1 lifted gradients of (u, v, w, e) and the viscous stresses and
//              heat flux they give (src/Navier_Stokes3d.cpp:182-391,
//              src/Euler3d.cpp:1144-1199): viscous volume divergence (modal) and
//              normal viscous flux at the Gauss points of every face.
#include "dgx_contract.hpp"
#include "dgx_kernels.hpp"

namespace dgx {

template <int P, int NTH = kStageThreads> struct S3 {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1;
  static constexpr int L2 = L * L, L3 = L * L * L, Q2 = Q * Q, Q3 = Q * Q * Q, QF2 = QF * QF;
  static constexpr int LP = odd(L), QP = odd(Q), FP = odd(QF);   // padded pitches
  static constexpr int NT = NTH;
  static constexpr int NPT = (Q3 + NT - 1) / NT;        // nodes per thread
  // doubles per LDS tensor: the largest intermediate of any phase
  static constexpr int NA = cmax(cmax(cmax(5 * Q2 * QP, 5 * Q2 * LP),                 // nodal, backward
                                      cmax(30 * L * FP, 30 * L * LP)),                 // face projection
                                 cmax(cmax(24 * Q * QP, 24 * QF * QP),                 // viscous layers
                                      cmax(cmax(20 * L * LP + 15 * L2,                 // trace collapse
                                                10 * QF * cmax(LP, 3 * L) + 15 * QF2), // trace faces
                                           24 * QF * FP)));
  static constexpr int NSL = (5 * L2 + NT - 1) / NT;    // modal pencil slots (5 fields)
  static constexpr int NSG = (4 * L2 + NT - 1) / NT;    // modal pencil slots (4 fields)
  static constexpr int NG = 30 * QF * FP;               // face corrections
  static constexpr int lds_stage = (2 * NA + NG) * 8;
  static constexpr int lds_vterm = (2 * NA + 24 * Q * QP) * 8;
  // BR2: face-lift nodal values, metric rows, end traces (+ layer velocity)
  static constexpr bool br2_vel_in_b = 42 * Q * QP <= NA;
  static constexpr int lds_vterm_br2 = (2 * NA + 24 * Q * QP + 16 + 2 * Q + (br2_vel_in_b ? 0 : 18 * Q * QP)) * 8;
  // padded position of nodal index n = (i*Q + j)*Q + k
  static __device__ __forceinline__ int nidx(int n) { return (n / Q) * QP + (n - (n / Q) * Q); }
};

// Final contraction of a chain over the first spatial dimension of
// [NF][Q][L][L] (pitch LP), accumulated into registers:
// acc[k][l] += sum_i M[l][i] in[f][i][m][n] for the thread's pencils
// G = tid + k*NT = f*L2 + m*L + n.
template <int P, int NF, int NS, int NTH = kStageThreads>
__device__ __forceinline__ void chain_to_regs(const double *__restrict__ M, const double *in,
                                              double (&acc)[NS][P + 1], int tid) {
  using S = S3<P, NTH>;
  constexpr int L = S::L, Q = S::Q, L2 = S::L2, LP = S::LP, NT = S::NT;
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const int G = tid + k * NT;
    if (G < NF * L2) {
      const int f = G / L2, pp = G - f * L2, m = pp / L, n = pp - m * L;
      double x[Q];
#pragma unroll
      for (int i = 0; i < Q; ++i) x[i] = in[((f * Q + i) * L + m) * LP + n];
#pragma unroll
      for (int l = 0; l < L; ++l) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < Q; ++i) s = fma(M[l * Q + i], x[i], s);
        acc[k][l] += s;
      }
    }
  }
}

// Collocation derivative along reference direction D of NF nodal fields in A
// ([NF][Q][Q][Q], pitch QP), written (ACC: added) to B.  Summed over the three
// directions of the contravariant fluxes, B holds the nodal divergence, and
// projecting it once (proj_to_regs) equals the three projected-derivative
// chains: (BW D)_d (x) BW (x) BW = (BW (x) BW (x) BW)(D along d).
template <int P, int NF, int D, bool ACC, int NTH = kStageThreads>
__device__ __forceinline__ void nodal_deriv(const double *A, double *B, int tid) {
  using S = S3<P, NTH>;
  constexpr int Q = S::Q, QP = S::QP, NT = S::NT;
  static_assert(D != 2 || !ACC, "the c derivative comes first and writes B");
  if constexpr (D == 2) con_last<NF * Q * Q, Q, Q, QP, QP, NT>(CT<P>::DN(), A, B, tid);
  else if constexpr (D == 1) con_mid<NF * Q, Q, 1, Q, Q, QP, QP, NT, ACC>(CT<P>::DN(), A, B, tid);
  else con_mid<NF, Q, Q, Q, Q, QP, QP, NT, ACC>(CT<P>::DN(), A, B, tid);
}

// Projection BW (x) BW (x) BW of NF nodal fields in B, accumulated into the
// modal registers acc (the pencils of chain_to_regs); A is scratch.
template <int P, int NF, int NS, int NTH = kStageThreads>
__device__ __forceinline__ void proj_to_regs(double *A, double *B, double (&acc)[NS][P + 1], int tid) {
  using S = S3<P, NTH>;
  constexpr int L = S::L, Q = S::Q, LP = S::LP, QP = S::QP, NT = S::NT;
  con_last<NF * Q * Q, Q, L, QP, LP, NT>(CT<P>::BW(), B, A, tid);   // [NF Q Q][L]
  __syncthreads();
  con_mid<NF * Q, Q, 1, L, L, LP, LP, NT>(CT<P>::BW(), A, B, tid);  // [NF Q][L][L]
  __syncthreads();
  chain_to_regs<P, NF, NS, NTH>(CT<P>::BW(), B, acc, tid);           // [NF][L][L][L]
  __syncthreads();
}

// modal [NF][L][L][L] (pitch LP) in A -> nodal [NF][Q][Q][Q] (pitch QP) in B
template <int P, int NF, int NTH = kStageThreads>
__device__ __forceinline__ void backward_batch(double *A, double *B, int tid) {
  using S = S3<P, NTH>;
  constexpr int L = S::L, Q = S::Q, LP = S::LP, QP = S::QP, NT = S::NT;
  con_mid<NF, L, L, L, Q, LP, LP, NT>(CT<P>::BT(), A, B, tid);   // [NF][Q][L][L]
  __syncthreads();
  con_mid<NF * Q, L, 1, L, Q, LP, LP, NT>(CT<P>::BT(), B, A, tid);   // [NF Q][Q][L]
  __syncthreads();
  con_last<NF * Q * Q, L, Q, LP, QP, NT>(CT<P>::BT(), A, B, tid);   // [NF Q Q][Q]
  __syncthreads();
}

// Face blocks g[NFB][QF][QF] (pitch FP) -> H[NFB][L][L] (pitch LP) in B, via W in A.
template <int P, int NFB, int NTH = kStageThreads>
__device__ __forceinline__ void face_project(const double *g, double *A, double *B, int tid) {
  using S = S3<P, NTH>;
  constexpr int L = S::L, QF = S::QF, LP = S::LP, FP = S::FP, NT = S::NT;
  con_mid<NFB, QF, 1, QF, L, FP, FP, NT>(CT<P>::PW(), g, A, tid);     // [NFB][L][QF]
  __syncthreads();
  con_last<NFB * L, QF, L, FP, LP, NT>(CT<P>::PW(), A, B, tid);     // [NFB][L][L]
  __syncthreads();
}

// Contribution at modal (l,m,n) of the six face blocks H[s] ([L][L], pitch LP)
// weighted by c[s].  Face normal dims: 4,2 -> a; 1,3 -> b; 0,5 -> c.
template <int P>
__device__ __forceinline__ double face_sum(const double *H, int l, int m, int n, const double c[6]) {
  constexpr int L = P + 1, LP = odd(L), HB = L * LP;
  const double *B0 = CT<P>::Bend(), *B1 = CT<P>::Bend() + L;
  double s = 0.0;
  s = fma(c[4] * B0[l], H[4 * HB + m * LP + n], s);
  s = fma(c[2] * B1[l], H[2 * HB + m * LP + n], s);
  s = fma(c[1] * B0[m], H[1 * HB + l * LP + n], s);
  s = fma(c[3] * B1[m], H[3 * HB + l * LP + n], s);
  s = fma(c[0] * B0[n], H[0 * HB + l * LP + m], s);
  s = fma(c[5] * B1[n], H[5 * HB + l * LP + m], s);
  return s;
}

struct StageArgs {
  const double *T;       // face traces of the input state
  const double *V;       // viscous volume term [ne][4][L3] (Navier-Stokes)
  const double *TV;      // viscous normal flux [ne][6][4][QF2] (Navier-Stokes)
  double *U;             // modal state (read; written when updating)
  double *Ut;            // LSRK register (update mode)
  double *R;             // residual output (residual mode)
  double *Tn;            // traces of the updated state (update mode)
  double ca, cb, dt;     // Ut = ca*Ut - dt*R ; U += cb*Ut
  int ne;
  int *flag;
};

// The stage kernel (125 VGPRs, 76 KB LDS at p = 7) runs 8 waves per element:
// every contraction phase is then a single pass, and a CU holds 4 waves per
// SIMD instead of 2 within the same LDS.
constexpr int kStage3Threads = 2 * kStageThreads;
#ifndef DGX_VTERM_THREADS
#define DGX_VTERM_THREADS 512
#endif
constexpr int kVterm3Threads = DGX_VTERM_THREADS;

template <int P, bool NS, bool UPDATE>
__global__ void __launch_bounds__(kStage3Threads)
k_stage3(StageArgs a, Tab tb, Geo geo, Phys ph) {
  using S = S3<P, kStage3Threads>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, L3 = S::L3, Q3 = S::Q3, QF2 = S::QF2,
                LP = S::LP, FP = S::FP, NT = S::NT, NPT = S::NPT, NSL = S::NSL;
  extern __shared__ double sm[];
  double *A = sm, *B = A + S::NA, *gF = B + S::NA;
  const int ne = a.ne, e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;
  const double gm1 = ph.gamma - 1.0;

  // the element's metric terms and modal state are loaded first: the metric
  // into registers before any barrier, the modal state landing in LDS after
  // the face loop, whose dependent loads (links, orientation, neighbour traces)
  // it overlaps
  double mt[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) mt[c] = geo.metric[9 * (size_t)e + c];
  constexpr int NML = (5 * L3 + NT - 1) / NT;
  double um[NML];
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int i = tid + k * NT;
    const int v = i / L3, m = i - v * L3;
    um[k] = i < 5 * L3 ? a.U[((size_t)v * ne + e) * L3 + m] : 0.0;
  }

  // face corrections g = (J_side/J_vol) (Fhat - F(U-).n) [- BR1 viscous part]
  bool bad = false;
  for (int i = tid; i < 6 * QF2; i += NT) {
    const int s = i / QF2, q = i - s * QF2;
    const int es = e * 6 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int qn = geo.perm[geo.permid[es] * QF2 + q];
    const double nx = geo.fnorm[3 * es], ny = geo.fnorm[3 * es + 1], nz = geo.fnorm[3 * es + 2];
    const double jr = geo.fjr[es];
    double UL[5], UR[5], FL[5];
    const double *tl = a.T + ((size_t)e * 6 + s) * 5 * QF2;
#pragma unroll
    for (int v = 0; v < 5; ++v) UL[v] = tl[v * QF2 + q];
    if (nb >= 0) {
      const double *tr = a.T + ((size_t)nb * 6 + ns) * 5 * QF2;
#pragma unroll
      for (int v = 0; v < 5; ++v) UR[v] = tr[v * QF2 + qn];
    } else {
      ghost3(-1 - nb, UL, nx, ny, nz, ph.bcU, UR);
    }
    double Fh[5];
    riemann3(UL, UR, nx, ny, nz, ph.gamma, ph.flux, FL, Fh, bad);
    const int qa = q / QF, qb = q - qa * QF;
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      double g = jr * (Fh[v] - FL[v]);
      if (NS && v > 0 && nb >= 0) {   // boundary: ghost flux -S, so the face flux is the own one
        const double *vl = a.TV + ((size_t)e * 6 + s) * 4 * QF2;
        const double *vr = a.TV + ((size_t)nb * 6 + ns) * 4 * QF2;
        g += 0.5 * jr * (vl[(v - 1) * QF2 + q] + vr[(v - 1) * QF2 + qn]);
      }
      gF[((v * 6 + s) * QF + qa) * FP + qb] = g;
    }
  }
  if (bad) atomicOr(a.flag, 1);

  // modal state -> nodal registers
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int i = tid + k * NT;
    const int v = i / L3, m = i - v * L3;
    if (i < 5 * L3) A[(m / L + v * L2) * LP + (m % L)] = um[k];
  }
  __syncthreads();
  backward_batch<P, 5, kStage3Threads>(A, B, tid);
  constexpr int FS = Q * Q * S::QP;   // padded nodal field size
  double un[NPT][5];
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int n = tid + r * NT;
    const int pn = S::nidx(n < Q3 ? n : 0);
#pragma unroll
    for (int v = 0; v < 5; ++v) un[r][v] = n < Q3 ? B[v * FS + pn] : 1.0;
  }

  // volume: projected divergence of the contravariant Euler fluxes
  double acc[NSL][L];
#pragma unroll
  for (int k = 0; k < NSL; ++k)
#pragma unroll
    for (int l = 0; l < L; ++l) acc[k][l] = 0.0;
  double pr[NPT], ux[NPT], uy[NPT], uz[NPT];
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const double ir = 1.0 / un[r][0];
    ux[r] = un[r][1] * ir; uy[r] = un[r][2] * ir; uz[r] = un[r][3] * ir;
    pr[r] = gm1 * (un[r][4] - 0.5 * (un[r][1] * ux[r] + un[r][2] * uy[r] + un[r][3] * uz[r]));
  }
  auto vol_dir = [&](auto Dc) {
    constexpr int d = decltype(Dc)::value;
    const double m0 = mt[3 * d], m1 = mt[3 * d + 1], m2 = mt[3 * d + 2];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int n = tid + r * NT;
      if (n < Q3) {
        const int pn = S::nidx(n);
        const double vn = m0 * ux[r] + m1 * uy[r] + m2 * uz[r];   // contravariant velocity
        A[pn] = un[r][0] * vn;
        A[FS + pn] = un[r][1] * vn + m0 * pr[r];
        A[2 * FS + pn] = un[r][2] * vn + m1 * pr[r];
        A[3 * FS + pn] = un[r][3] * vn + m2 * pr[r];
        A[4 * FS + pn] = (un[r][4] + pr[r]) * vn;
      }
    }
    __syncthreads();
    nodal_deriv<P, 5, d, d != 2, kStage3Threads>(A, B, tid);
    __syncthreads();
  };
  // nodal divergence of the contravariant fluxes (c first: it writes B), then
  // one projection to modes
  vol_dir(IC<2>());
  vol_dir(IC<1>());
  vol_dir(IC<0>());
  proj_to_regs<P, 5, NSL, kStage3Threads>(A, B, acc, tid);

  // the update's inputs (viscous volume term, RK register, state) are loaded
  // before the face projection, whose barriers then cover their latency
  double vin[NSL][L], utin[NSL][L], uin[NSL][L];
#pragma unroll
  for (int k = 0; k < NSL; ++k) {
    const int G = tid + k * NT;
    const int v = G / L2, pp = G - v * L2;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const bool in = G < 5 * L2;
      const size_t gi = ((size_t)v * ne + e) * L3 + l * L2 + pp;
      vin[k][l] = (NS && in && v > 0) ? a.V[((size_t)e * 4 + (v - 1)) * L3 + l * L2 + pp] : 0.0;
      utin[k][l] = (UPDATE && in) ? a.Ut[gi] : 0.0;
      uin[k][l] = (UPDATE && in) ? a.U[gi] : 0.0;
    }
  }

  // faces: H = (PW (x) PW) gF for 5 fields x 6 faces, in B
  face_project<P, 30, kStage3Threads>(gF, A, B, tid);
  const double one6[6] = {1.0, 1.0, 1.0, 1.0, 1.0, 1.0};
#pragma unroll
  for (int k = 0; k < NSL; ++k) {
    const int G = tid + k * NT;
    if (G < 5 * L2) {
      const int v = G / L2, pp = G - v * L2, m = pp / L, nn = pp - m * L;
#pragma unroll
      for (int l = 0; l < L; ++l) {
        double r = acc[k][l] + face_sum<P>(B + v * 6 * L * LP, l, m, nn, one6);
        if (NS && v > 0) r -= vin[k][l];
        const size_t gi = ((size_t)v * ne + e) * L3 + l * L2 + pp;
        if (!UPDATE) {
          a.R[gi] = r;
        } else {
          const double ut = a.ca * utin[k][l] - a.dt * r;
          const double u = fma(a.cb, ut, uin[k][l]);
          a.Ut[gi] = ut;
          a.U[gi] = u;
          acc[k][l] = u;
        }
      }
    }
  }
  if (!UPDATE) return;

  // face traces of the updated state for the next stage
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NSL; ++k) {
    const int G = tid + k * NT;
    if (G < 5 * L2) {
      const int v = G / L2, pp = G - v * L2, m = pp / L, nn = pp - m * L;
#pragma unroll
      for (int l = 0; l < L; ++l) A[((v * L + l) * L + m) * LP + nn] = acc[k][l];
    }
  }
  __syncthreads();
  constexpr int KP = 3;   // pitch of the [..][2] collapse
  double *K0 = B, *K1 = B + 10 * L * LP, *K2 = B + 20 * L * LP;
  con_mid<5, L, L, L, 2, LP, LP, NT>(CT<P>::Bend(), A, K0, tid);       // [5][2][L][L]
  con_mid<5 * L, L, 1, L, 2, LP, LP, NT>(CT<P>::Bend(), A, K1, tid);   // [5][L][2][L]
  con_last<5 * L2, L, 2, LP, KP, NT>(CT<P>::Bend(), A, K2, tid);       // [5][L][L][2]
  __syncthreads();
  double *W = A, *O = A + 10 * QF * cmax(LP, KP * L);
  double *Te = a.Tn + (size_t)e * 6 * 5 * QF2;
  // faces 4 / 2: [5][2][QF][QF]
  con_mid<10, L, 1, L, QF, LP, LP, NT>(CT<P>::Bg(), K0, W, tid);      // [10][QF][L]
  __syncthreads();
  con_last<10 * QF, L, QF, LP, FP, NT>(CT<P>::Bg(), W, O, tid);       // [10][QF][QF]
  __syncthreads();
  for (int i = tid; i < 10 * QF2; i += NT) {
    const int ve = i / QF2, q = i - ve * QF2, v = ve >> 1, end = ve & 1;
    Te[((end ? 2 : 4) * 5 + v) * QF2 + q] = O[(ve * QF + q / QF) * FP + q % QF];
  }
  __syncthreads();
  // faces 1 / 3: [5][QF][2][QF]
  con_mid<5, L, 2, L, QF, LP, LP, NT>(CT<P>::Bg(), K1, W, tid);       // [5][QF][2][L]
  __syncthreads();
  con_last<10 * QF, L, QF, LP, FP, NT>(CT<P>::Bg(), W, O, tid);       // [5][QF][2][QF]
  __syncthreads();
  for (int i = tid; i < 10 * QF2; i += NT) {
    const int v = i / (2 * QF2), r = i - v * 2 * QF2, end = r / QF2, q = r - end * QF2;
    const int ia = q / QF, ic = q - ia * QF;
    Te[((end ? 3 : 1) * 5 + v) * QF2 + q] = O[((v * QF + ia) * 2 + end) * FP + ic];
  }
  __syncthreads();
  // faces 0 / 5: [5][QF][QF][2]
  con_mid<5, L, L, 2, QF, KP, KP, NT>(CT<P>::Bg(), K2, W, tid);       // [5][QF][L][2]
  __syncthreads();
  con_mid<5 * QF, L, 1, 2, QF, KP, KP, NT>(CT<P>::Bg(), W, O, tid);   // [5][QF][QF][2]
  __syncthreads();
  for (int i = tid; i < 10 * QF2; i += NT) {
    const int v = i / (2 * QF2), r = i - v * 2 * QF2, end = r / QF2, q = r - end * QF2;
    Te[((end ? 5 : 0) * 5 + v) * QF2 + q] = O[(v * QF2 + q) * KP + end];
  }
}

// Last contraction of the lifted gradient along reference direction D for the
// 4 primitives: out = NPD x along D plus clo KE_lo Hn[face lo] + chi KE_hi
// Hn[face hi] (the two face lifts across D at their own GLL points; faces 4/2
// for a, 1/3 for b, 0/5 for c).  One thread per pencil of 4 Q^2, all Q outputs.
template <int P, int D, int NT>
__device__ __forceinline__ void lifted_dir(const double *in, double *out, const double *Hn, double clo, double chi,
                                           int tid) {
  constexpr int Q = P + 2, QP = odd(Q), HB = Q * QP, NP = 4 * Q * Q;
  constexpr int flo = D == 0 ? 4 : (D == 1 ? 1 : 0), fhi = D == 0 ? 2 : (D == 1 ? 3 : 5);
  const double *M = CT<P>::NPD(), *KE = CT<P>::KE();
  for (int pc = tid; pc < NP; pc += NT) {
    const int f = pc / (Q * Q), t = pc - f * Q * Q, t1 = t / Q, t2 = t - t1 * Q;
    // (t1, t2): the two other indices in order; base and stride of the pencil
    int base, stride;
    if (D == 0) { base = (f * Q * Q + t1) * QP + t2; stride = Q * QP; }          // (j, k)
    else if (D == 1) { base = (f * Q * Q + t1 * Q) * QP + t2; stride = QP; }    // (i, k)
    else { base = (f * Q * Q + t1 * Q + t2) * QP; stride = 1; }                 // (i, j)
    double x[Q];
#pragma unroll
    for (int m = 0; m < Q; ++m) x[m] = in[base + m * stride];
    const double hl = clo * Hn[(f * 6 + flo) * HB + t1 * QP + t2];
    const double hh = chi * Hn[(f * 6 + fhi) * HB + t1 * QP + t2];
#pragma unroll 3   // about 3 coefficient rows live in scalar registers at a time
    for (int o = 0; o < Q; ++o) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < Q; ++m) s = fma(M[o * Q + m], x[m], s);
      s = fma(KE[o], hl, s);
      s = fma(KE[Q + o], hh, s);
      out[base + o * stride] = s;
    }
  }
}

// Viscous flux of equation q (1..3 momentum, 4 energy) contracted with the
// vector (n0, n1, n2), from the stress tensor t = (t11 t12 t13 t22 t23 t33),
// heat flux h = kappa grad e and the velocity (u, v, w).
__device__ __forceinline__ double visc_flux(const double t[6], const double h[3], double u, double v,
                                           double w, int q, double n0, double n1, double n2) {
  const double r0 = t[0] * n0 + t[1] * n1 + t[2] * n2;   // row 1 of tau . n
  const double r1 = t[1] * n0 + t[3] * n1 + t[4] * n2;
  const double r2 = t[2] * n0 + t[4] * n1 + t[5] * n2;
  if (q == 0) return r0;
  if (q == 1) return r1;
  if (q == 2) return r2;
  return (r0 * u + r1 * v + r2 * w) + (h[0] * n0 + h[1] * n1 + h[2] * n2);
}

// BR1 viscous terms: V[e][eq][L3] = modal divergence of the viscous flux,
// TV[e][s][eq][QF2] = viscous flux . n at the Gauss points of face s.
// With BR2, TV holds the BR2 face flux (stresses of grad_h q + eta r_s on face
// s) and V also carries the lift of (BR2 - BR1) flux . n on the element's own
// faces, so that the stage kernel's face correction, 1/2 (TV- + TV+) lifted,
// completes the strong form -div F(G) - lift({F_BR2}.n - F(G)-.n).
// The lifted gradient is assembled one reference direction at a time: for an
// affine hex the normals of the two faces across reference direction d are
// -+ grad(xi_d)/|grad(xi_d)|, so the physical gradient is
// sum_d dxi_d/dx (D_d + lift of those two faces / |grad xi_d|), and each
// direction's nodal contribution goes straight into the stress tensor and heat
// flux at the thread's nodes.
// Two 512-thread workgroups per CU (4 waves per SIMD) where their LDS fits,
// which needs at most 128 VGPRs.
template <int P, bool BR2>
__global__ void __launch_bounds__(kVterm3Threads)
__attribute__((amdgpu_waves_per_eu((BR2 ? S3<P, kVterm3Threads>::lds_vterm_br2 : S3<P, kVterm3Threads>::lds_vterm) <=
                                           80 * 1024 ? 4 : 1)))
k_vterm3(const double *__restrict__ U, const double *__restrict__ T, double *__restrict__ V,
         double *__restrict__ TV, Tab tb, Geo geo, Phys ph, int ne) {
  using S = S3<P, kVterm3Threads>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, L3 = S::L3, Q3 = S::Q3, QF2 = S::QF2,
                LP = S::LP, QP = S::QP, FP = S::FP, NT = S::NT, NPT = S::NPT, NSG = S::NSG;
  extern __shared__ double sm[];
  double *A = sm, *B = A + S::NA, *Jq = B + S::NA;
  const int e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;
  constexpr int FS = Q * Q * QP;

  // the modal state's loads are issued first and land in LDS after the jump
  // loop, whose dependent loads they overlap
  constexpr int NML = (5 * L3 + NT - 1) / NT;
  double um[NML];
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int i = tid + k * NT;
    const int v = i / L3, m = i - v * L3;
    um[k] = i < 5 * L3 ? U[((size_t)v * ne + e) * L3 + m] : 0.0;
  }
  // primitive jumps J_side/J_vol ({q} - q-) at the Gauss points (average_flux '+')
  for (int i = tid; i < 6 * QF2; i += NT) {
    const int s = i / QF2, q = i - s * QF2;
    const int es = e * 6 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int qn = geo.perm[geo.permid[es] * QF2 + q];
    const double jr = geo.fjr[es];
    const double *tl = T + ((size_t)e * 6 + s) * 5 * QF2;
    double pl[4], pq[4];
    {
      const double r = tl[q], u = tl[QF2 + q] / r, v = tl[2 * QF2 + q] / r, w = tl[3 * QF2 + q] / r;
      double en = tl[4 * QF2 + q] / r;
      double wk = u * u; wk = v * v + wk; wk = w * w + wk;
      pl[0] = u; pl[1] = v; pl[2] = w; pl[3] = -0.5 * wk + en;
    }
    if (nb < 0) {
      ghost_prim3(-1 - nb, tl[q], ph, pq);
    } else {
      const double *tr = T + ((size_t)nb * 6 + ns) * 5 * QF2;
      const double r = tr[qn], u = tr[QF2 + qn] / r, v = tr[2 * QF2 + qn] / r, w = tr[3 * QF2 + qn] / r;
      double en = tr[4 * QF2 + qn] / r;
      double wk = u * u; wk = v * v + wk; wk = w * w + wk;
      pq[0] = u; pq[1] = v; pq[2] = w; pq[3] = -0.5 * wk + en;
    }
    const int qa = q / QF, qb = q - qa * QF;
#pragma unroll
    for (int c = 0; c < 4; ++c) Jq[((c * 6 + s) * QF + qa) * FP + qb] = jr * (0.5 * (pl[c] + pq[c]) - pl[c]);
  }
#pragma unroll
  for (int k = 0; k < NML; ++k) {
    const int i = tid + k * NT;
    const int v = i / L3, m = i - v * L3;
    if (i < 5 * L3) A[(m / L + v * L2) * LP + (m % L)] = um[k];
  }
  __syncthreads();
  backward_batch<P, 5, NT>(A, B, tid);
  double pv[NPT][4];   // u, v, w, e at the thread's nodes
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int n = tid + r * NT;
    const int pn = S::nidx(n < Q3 ? n : 0);
    const double rho = n < Q3 ? B[pn] : 1.0;
    const double u = n < Q3 ? B[FS + pn] / rho : 0.0, v = n < Q3 ? B[2 * FS + pn] / rho : 0.0;
    const double w = n < Q3 ? B[3 * FS + pn] / rho : 0.0;
    const double en = n < Q3 ? B[4 * FS + pn] / rho : 0.0;
    double wk = u * u; wk = v * v + wk; wk = w * w + wk;
    pv[r][0] = u; pv[r][1] = v; pv[r][2] = w; pv[r][3] = -0.5 * wk + en;
  }
  __syncthreads();
  // face lift of the jumps: (PW (x) PW) Jq -> H[4][6][L][L], then its values at
  // each face's own GLL points Hn[4][6][Q][Q] = (BT (x) BT) H, kept in the Jq region
  con_mid<24, QF, 1, QF, L, FP, FP, NT>(CT<P>::PW(), Jq, A, tid);     // [24][L][QF]
  __syncthreads();
  con_last<24 * L, QF, L, FP, LP, NT>(CT<P>::PW(), A, Jq, tid);     // [24][L][L]
  __syncthreads();
  con_mid<24, L, 1, L, Q, LP, LP, NT>(CT<P>::BT(), Jq, A, tid);     // [24][Q][L]
  __syncthreads();
  con_last<24 * Q, L, Q, LP, QP, NT>(CT<P>::BT(), A, Jq, tid);      // [24][Q][Q]
  __syncthreads();
  const double *Hn = Jq;
  constexpr int HB = Q * QP;   // one nodal face block

  const double *mt = geo.metric + 9 * (size_t)e;
  const double mu = ph.vis, kap = ph.vis * ph.gamma / ph.Pr;
  double tq[NPT][9];   // t11 t12 t13 t22 t23 t33, kappa grad e
#pragma unroll
  for (int r = 0; r < NPT; ++r)
#pragma unroll
    for (int j = 0; j < 9; ++j) tq[r][j] = 0.0;
  // The lifted gradient along reference direction d, at the GLL nodes: the
  // projected derivative (BT BW along the two other directions, then BT QD
  // along d: three contractions on nodal tensors) plus the nodal values of the
  // two face lifts across d, folded into the last contraction as two more
  // inputs with the end-trace columns of KE.  Equal to projecting D_d q and the
  // face terms to modes and evaluating the result back at the nodes.
  auto grad_dir = [&](auto Dc) {
    constexpr int d = decltype(Dc)::value;
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int n = tid + r * NT;
      if (n < Q3) {
        const int pn = S::nidx(n);
#pragma unroll
        for (int c = 0; c < 4; ++c) A[c * FS + pn] = pv[r][c];
      }
    }
    __syncthreads();
    if constexpr (d == 2) con_mid<4 * Q, Q, 1, Q, Q, QP, QP, NT>(CT<P>::NP1(), A, B, tid);   // along b
    else con_last<4 * Q * Q, Q, Q, QP, QP, NT>(CT<P>::NP1(), A, B, tid);                    // along c
    __syncthreads();
    if constexpr (d == 0) con_mid<4 * Q, Q, 1, Q, Q, QP, QP, NT>(CT<P>::NP1(), B, A, tid);  // along b
    else con_mid<4, Q, Q, Q, Q, QP, QP, NT>(CT<P>::NP1(), B, A, tid);                       // along a
    __syncthreads();
    const double g0 = mt[3 * d], g1 = mt[3 * d + 1], g2 = mt[3 * d + 2];
    const double inv = 1.0 / sqrt(g0 * g0 + g1 * g1 + g2 * g2);
    lifted_dir<P, d, NT>(A, B, Hn, -inv, inv, tid);                                        // along d
    __syncthreads();
    // nodal d/dxi_d of (u, v, w, e) -> stress and heat flux
    const double c43 = mu * (4.0 / 3.0), c23 = mu * (2.0 / 3.0);
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int n = tid + r * NT;
      const int pn = S::nidx(n < Q3 ? n : 0);
      const double au = B[pn], av = B[FS + pn], aw = B[2 * FS + pn], ae = B[3 * FS + pn];
      const double ux = g0 * au, uy = g1 * au, uz = g2 * au;
      const double vx = g0 * av, vy = g1 * av, vz = g2 * av;
      const double wx = g0 * aw, wy = g1 * aw, wz = g2 * aw;
      tq[r][0] += c43 * ux - c23 * (vy + wz);
      tq[r][1] += mu * (uy + vx);
      tq[r][2] += mu * (uz + wx);
      tq[r][3] += c43 * vy - c23 * (ux + wz);
      tq[r][4] += mu * (vz + wy);
      tq[r][5] += c43 * wz - c23 * (ux + vy);
      tq[r][6] += kap * (g0 * ae);
      tq[r][7] += kap * (g1 * ae);
      tq[r][8] += kap * (g2 * ae);
    }
  };
  grad_dir(IC<0>());
  grad_dir(IC<1>());
  grad_dir(IC<2>());

  double acc[NSG][L];   // modal viscous volume term of the 4 equations
  // BR2 scratch: Hn [24][Q][Q] and gs[9], KE[2][Q] in the Jq region; the
  // (BR2 - BR1) flux difference Df [4][6][Q][Q] in B, the face-layer velocity
  // Vel [6][3][Q][Q] after it (or after KE when B is too small)
  constexpr int HN = 24 * Q * QP, VOFF = S::br2_vel_in_b ? 0 : HN + 16 + 2 * Q;
  double *Df = B, *Vel = S::br2_vel_in_b ? B + HN : Jq + VOFF;
  if constexpr (BR2) {
    // the face lifts' nodal values Hn[24][Q][Q] are in place; after them the
    // unit rows of the metric and the end-trace table
    if (tid < 9) {
      const int d = tid / 3;
      Jq[HN + tid] = mt[tid] / sqrt(mt[3 * d] * mt[3 * d] + mt[3 * d + 1] * mt[3 * d + 1] +
                                    mt[3 * d + 2] * mt[3 * d + 2]);
    }
    if (tid >= 16 && tid < 16 + 2 * Q) Jq[HN + tid] = CT<P>::KE()[tid - 16];
    __syncthreads();
  }

  // normal viscous flux on the GLL face layers [6][4][Q][Q] (pitch QP) -> Gauss
  // points (with BR2, also the velocity on the layers)
  double *Lay = A;
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int n = tid + r * NT;
    if (n < Q3) {
      const int i = n / (Q * Q), j = (n / Q) % Q, k = n % Q;
      // face f at layer position (a1, a2); one call per face the node lies on,
      // each with a constant face (no indexed arrays, which would live in scratch)
      auto put = [&](int f, int a1, int a2) {
        const double *nn = geo.fnorm + 3 * (e * 6 + f);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          Lay[((f * 4 + q) * Q + a1) * QP + a2] =
              visc_flux(tq[r], tq[r] + 6, pv[r][0], pv[r][1], pv[r][2], q, nn[0], nn[1], nn[2]);
        if constexpr (BR2) {
#pragma unroll
          for (int q = 0; q < 3; ++q) Vel[((f * 3 + q) * Q + a1) * QP + a2] = pv[r][q];
        }
      };
      if (k == 0) put(0, i, j);
      if (k == Q - 1) put(5, i, j);
      if (j == 0) put(1, i, k);
      if (j == Q - 1) put(3, i, k);
      if (i == 0) put(4, j, k);
      if (i == Q - 1) put(2, j, k);
    }
  }
  __syncthreads();
  if constexpr (BR2) {
    // BR2 correction, one thread per face point: flux difference Df and Lay += Df
    for (int idx = tid; idx < 6 * Q * Q; idx += NT) {
      const int f = idx / (Q * Q), rem = idx - f * (Q * Q), l1 = rem / Q, l2 = rem - l1 * Q;
      int i, j, k;
      if (f == 0 || f == 5) { i = l1; j = l2; k = f == 0 ? 0 : Q - 1; }
      else if (f == 1 || f == 3) { i = l1; k = l2; j = f == 1 ? 0 : Q - 1; }
      else { j = l1; k = l2; i = f == 4 ? 0 : Q - 1; }
      const double *nn = geo.fnorm + 3 * (e * 6 + f);
      const double *vl = Vel + (f * 3 * Q + l1) * QP + l2;
      double fd[4];
      br2_flux<Q, QP>(fd, Jq, Jq + HN + 16, f, i, j, k, Jq + HN, mu, kap, ph.eta, vl[0], vl[Q * QP],
                      vl[2 * Q * QP], nn[0], nn[1], nn[2]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        Lay[((f * 4 + q) * Q + l1) * QP + l2] += fd[q];
        Df[((q * 6 + f) * Q + l1) * QP + l2] = fd[q];
      }
    }
    __syncthreads();
  }
  double *Gs = B;   // scratch of the GLL -> Gauss interpolation
  if constexpr (BR2) {
    // lift of the flux difference on the element's own faces, into the volume term:
    // (PW Ig (x) PW Ig) difference -> modal face blocks [4][6][L][L] in B
    con_mid<24, Q, 1, Q, L, QP, QP, NT>(CT<P>::PI(), B, Jq, tid);   // [24][L][Q]
    __syncthreads();
    con_last<24 * L, Q, L, QP, LP, NT>(CT<P>::PI(), Jq, B, tid);    // [24][L][L]
    __syncthreads();
    double jr6[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) jr6[s] = geo.fjr[e * 6 + s];
#pragma unroll
    for (int k = 0; k < NSG; ++k) {
      const int Gi = tid + k * NT;
      const int q = Gi / L2, pp = Gi - q * L2, m = pp / L, nn = pp - m * L;
#pragma unroll
      for (int l = 0; l < L; ++l) acc[k][l] = Gi < 4 * L2 ? face_sum<P>(B + q * 6 * L * LP, l, m, nn, jr6) : 0.0;
    }
    Gs = Jq;
  } else {
#pragma unroll
    for (int k = 0; k < NSG; ++k)
#pragma unroll
      for (int l = 0; l < L; ++l) acc[k][l] = 0.0;
  }
  con_mid<24, Q, 1, Q, QF, QP, QP, NT>(CT<P>::Ig(), Lay, Gs, tid);   // [24][QF][Q]
  __syncthreads();
  con_last<24 * QF, Q, QF, QP, FP, NT>(CT<P>::Ig(), Gs, A, tid);     // [24][QF][QF]
  __syncthreads();
  for (int i = tid; i < 24 * QF2; i += NT) {
    const int sq = i / QF2, q = i - sq * QF2;
    TV[(size_t)e * 24 * QF2 + i] = A[(sq * QF + q / QF) * FP + q % QF];
  }
  __syncthreads();

  // viscous volume divergence of the 4 equations
  auto visc_dir = [&](auto Dc) {
    constexpr int d = decltype(Dc)::value;
    const double m0 = mt[3 * d], m1 = mt[3 * d + 1], m2 = mt[3 * d + 2];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int n = tid + r * NT;
      if (n < Q3) {
        const int pn = S::nidx(n);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          A[q * FS + pn] = visc_flux(tq[r], tq[r] + 6, pv[r][0], pv[r][1], pv[r][2], q, m0, m1, m2);
      }
    }
    __syncthreads();
    nodal_deriv<P, 4, d, d != 2, NT>(A, B, tid);
    __syncthreads();
  };
  visc_dir(IC<2>());
  visc_dir(IC<1>());
  visc_dir(IC<0>());
  proj_to_regs<P, 4, NSG, NT>(A, B, acc, tid);
#pragma unroll
  for (int k = 0; k < NSG; ++k) {
    const int Gi = tid + k * NT;
    if (Gi < 4 * L2) {
      const int q = Gi / L2, pp = Gi - q * L2;
#pragma unroll
      for (int l = 0; l < L; ++l) V[((size_t)e * 4 + q) * L3 + l * L2 + pp] = acc[k][l];
    }
  }
}

template <int P>
static hipError_t launch_stage3_p(const StageLaunch &s) {
  StageArgs a;
  a.T = s.T; a.V = s.V; a.TV = s.TV; a.U = s.U; a.Ut = s.Ut; a.R = s.R; a.Tn = s.Tn;
  a.ca = s.ca; a.cb = s.cb; a.dt = s.dt; a.ne = s.ne; a.flag = s.flag;
  if (s.op == 0) {
    if (s.phys.visc == 1) {
      const int sh = S3<P, kVterm3Threads>::lds_vterm_br2;
      (void)hipFuncSetAttribute((const void *)k_vterm3<P, true>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
      hipLaunchKernelGGL((k_vterm3<P, true>), dim3(s.ne), dim3(kVterm3Threads), sh, s.stream, s.U, s.T, s.Vw,
                         s.TVw, s.tab, s.geo, s.phys, s.ne);
    } else {
      const int sh = S3<P, kVterm3Threads>::lds_vterm;
      (void)hipFuncSetAttribute((const void *)k_vterm3<P, false>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
      hipLaunchKernelGGL((k_vterm3<P, false>), dim3(s.ne), dim3(kVterm3Threads), sh, s.stream, s.U, s.T, s.Vw,
                         s.TVw, s.tab, s.geo, s.phys, s.ne);
    }
    return hipGetLastError();
  }
  const int sh = S3<P, kStage3Threads>::lds_stage;
#define DGX_STAGE(NSV, UPD)                                                                          \
  do {                                                                                               \
    (void)hipFuncSetAttribute((const void *)k_stage3<P, NSV, UPD>,                                   \
                              hipFuncAttributeMaxDynamicSharedMemorySize, sh);                       \
    hipLaunchKernelGGL((k_stage3<P, NSV, UPD>), dim3(s.ne), dim3(kStage3Threads), sh, s.stream, a, s.tab, s.geo, \
                       s.phys);                                                                      \
  } while (0)
  if (s.ns) { if (s.update) DGX_STAGE(true, true); else DGX_STAGE(true, false); }
  else { if (s.update) DGX_STAGE(false, true); else DGX_STAGE(false, false); }
#undef DGX_STAGE
  return hipGetLastError();
}

bool stage3_supported(int p) { return p >= 1 && p <= 9; }

hipError_t launch_stage3(const StageLaunch &s) {
  switch (s.p) {
#define DGX_CASE(PP) case PP: return launch_stage3_p<PP>(s);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t upload_tables_s3d(const double *h) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), h, sizeof(double) * kTabTotal);
}

}  // namespace dgx
