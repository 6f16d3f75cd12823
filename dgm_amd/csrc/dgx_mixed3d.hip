// Mixed-precision fused hex stage kernels (p <= 8).
//
// Same operator, data flow and element-per-workgroup structure as the FP64
// kernels of dgx_stage3d.hip, with the arithmetic split by what each part
// needs:
//
//   FP64  conserved state U and the LSRK register Ut (HBM), face traces T
//         (HBM, evaluated in FP64: the Lax-Friedrichs dissipation
//         lambda (U- - U+) scales absolute trace errors by the sound speed, so
//         FP32 traces would cost ~1e-5 of the momentum residual at Ma = 0.1),
//         the element's constant mode, pointwise fluxes and primitives, the
//         face Riemann fluxes, metrics, the register update.
//   FP32  every sum-factorized contraction, on MFMA (v_mfma_f32_4x4x1_16b_f32,
//         one pencil per lane), applied to deviations from an element
//         constant; LDS tiles; the viscous terms V and TV (HBM).
//
// Deviations: the modal state is split into its constant mode (kept in FP64)
// and the rest (FP32), so nodal values are c + u' with u' small.  Before each
// divergence or gradient chain, the flux (or primitive) of the constant state
// is subtracted in FP64; the chains contain a derivative along one direction,
// which maps constants to zero, so the subtraction is exact and the FP32
// contraction only sees the variation within the element.  This keeps the
// large background of a low-Mach flow (p ~ 1/(gamma Ma^2)) out of FP32.
//
// MFMA layout (v_mfma_f32_4x4x1_16b_f32, 16 blocks of 4x4x1): result register
// r of lane l is sum_k A[4*(l/4) + r] * B[l].  With B = the lane's own pencil
// input x[k] and A = M[o0 + l%4][k], the four result registers of a lane are
// outputs o0..o0+3 of its own pencil; K steps accumulate in the result.
#include <vector>

#include "dgx_contract.hpp"
#include "dgx_kernels.hpp"

namespace dgx {

namespace {

typedef float f4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// FP32 MFMA coefficient tables in constant memory.  A matrix M[NO][N]
// (row-major) is stored as [NOB][4][NPAD] with NOB = ceil(NO/4) row blocks,
// NPAD = N rounded up to 4, zero padded: lane l of a wave loads row
// 4*ob + l%4 of block ob as float4s.
// ---------------------------------------------------------------------------
constexpr int mq_size(int no, int n) { return ((no + 3) / 4) * 4 * ((n + 3) / 4 * 4); }
constexpr int mq_order_size(int p) {
  return mq_size(p + 2, p + 1) + 2 * mq_size(p + 1, p + 2) + mq_size(p + 1, p + 1) + mq_size(2, p + 1) +
         mq_size(p + 1, p + 1) + mq_size(p + 1, p + 2) + mq_size(p + 1, p + 2);
}
constexpr int mq_order_offset(int p) {
  int s = 0;
  for (int k = 1; k < p; ++k) s += mq_order_size(k);
  return s;
}
constexpr int kMaxMixed = 8;
constexpr int kMqTotal = mq_order_offset(kMaxMixed + 1);
__constant__ __attribute__((aligned(16))) float c_mq[kMqTotal];

template <int P> struct MT {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1;
  static constexpr int oBT = mq_order_offset(P);            // [Q][L]   modal -> GLL
  static constexpr int oBW = oBT + mq_size(Q, L);           // [L][Q]   GLL inner product
  static constexpr int oQD = oBW + mq_size(L, Q);           // [L][Q]   inner product of D
  static constexpr int oBg = oQD + mq_size(L, Q);           // [QF][L]  modal -> Gauss
  static constexpr int oBend = oBg + mq_size(QF, L);        // [2][L]   modal -> +-1
  static constexpr int oPW = oBend + mq_size(2, L);         // [L][QF]  Gauss inner product
  static constexpr int oIg = oPW + mq_size(L, QF);          // [QF][Q]  GLL -> Gauss
  static constexpr int oPI = oIg + mq_size(QF, Q);          // [L][Q]   GLL face layer -> modal face
};

// sqrt(1/2)^3: the constant normalized Legendre mode of a hex
constexpr double kPhi0 = 0.35355339059327376220;

__device__ __forceinline__ f4 mfma4(float a, float b, f4 c) {
  return __builtin_amdgcn_mfma_f32_4x4x1f32(a, b, c, 0, 0, 0);
}

// Per-lane coefficients of M[NO][N] at c_mq + off: cf[ob][k] = M[4*ob + lane%4][k].
template <int NO, int N>
__device__ __forceinline__ void load_coef(int off, int tid, float (&cf)[(NO + 3) / 4][N]) {
  constexpr int NOB = (NO + 3) / 4, NPAD = (N + 3) / 4 * 4;
  const int r = tid & 3;
#pragma unroll
  for (int ob = 0; ob < NOB; ++ob) {
    const f4 *row = reinterpret_cast<const f4 *>(c_mq + off + (ob * 4 + r) * NPAD);
#pragma unroll
    for (int kk = 0; kk < NPAD / 4; ++kk) {
      const f4 v = row[kk];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (4 * kk + j < N) cf[ob][4 * kk + j] = v[j];
    }
  }
}

// Waves whose first pencil of pass k lies beyond NP skip the pass as a whole,
// so the MFMAs always run with every lane of the wave active.
__device__ __forceinline__ bool wave_has(int tid, int k, int NT, int NP) {
  return (tid & ~63) + k * NT < NP;
}

// [A][N][B][C] (last-dim pitch PI) -> [A][NO][B][C] (pitch PO), contracting N.
template <int A, int N, int B, int C, int NO, int PI, int PO, int NT, bool ACC = false>
__device__ __forceinline__ void mcon_mid(int off, const float *in, float *out, int tid) {
  constexpr int NP = A * B * C, KP = (NP + NT - 1) / NT, NOB = (NO + 3) / 4;
  float cf[NOB][N];
  load_coef<NO, N>(off, tid, cf);
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    if (!wave_has(tid, k, NT, NP)) continue;
    const int pc0 = tid + k * NT;
    const bool ok = pc0 < NP;
    const int pc = ok ? pc0 : 0;
    const int c = pc % C, t = pc / C, b = t % B, a = t / B;
    float x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = in[((a * N + i) * B + b) * PI + c];
#pragma unroll
    for (int ob = 0; ob < NOB; ++ob) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int i = 0; i < N; ++i) acc = mfma4(cf[ob][i], x[i], acc);
      if (ok) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int o = 4 * ob + r;
          if (o < NO) {
            float *dst = out + ((a * NO + o) * B + b) * PO + c;
            if (ACC) *dst += acc[r]; else *dst = acc[r];
          }
        }
      }
    }
  }
}

// [R][NI] (pitch PI) -> [R][NO] (pitch PO), contracting the last dimension.
template <int R, int NI, int NO, int PI, int PO, int NT>
__device__ __forceinline__ void mcon_last(int off, const float *in, float *out, int tid) {
  constexpr int KP = (R + NT - 1) / NT, NOB = (NO + 3) / 4;
  float cf[NOB][NI];
  load_coef<NO, NI>(off, tid, cf);
#pragma unroll
  for (int k = 0; k < KP; ++k) {
    if (!wave_has(tid, k, NT, R)) continue;
    const int r0 = tid + k * NT;
    const bool ok = r0 < R;
    const int rr = ok ? r0 : 0;
    float x[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) x[i] = in[rr * PI + i];
#pragma unroll
    for (int ob = 0; ob < NOB; ++ob) {
      f4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int i = 0; i < NI; ++i) acc = mfma4(cf[ob][i], x[i], acc);
      if (ok) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          if (4 * ob + r < NO) out[rr * PO + 4 * ob + r] = acc[r];
      }
    }
  }
}

template <int P, int NTH = kStageThreads> struct M3 {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1;
  static constexpr int L2 = L * L, L3 = L * L * L, Q2 = Q * Q, Q3 = Q * Q * Q, QF2 = QF * QF;
  static constexpr int LP = odd(L), QP = odd(Q), FP = odd(QF);
  static constexpr int NT = NTH;
  static constexpr int NPT = (Q3 + NT - 1) / NT;
  static constexpr int NOB = (L + 3) / 4;               // output blocks of L modes
  static constexpr int NA = cmax(cmax(cmax(5 * Q2 * QP, 5 * Q2 * LP), cmax(30 * L * FP, 30 * L * LP)),
                                 cmax(cmax(24 * Q * QP, 24 * QF * QP),
                                      cmax(cmax(20 * L * LP + 15 * L2, 10 * QF * cmax(LP, 3 * L) + 15 * QF2),
                                           24 * QF * FP)));
  static constexpr int NSL = (5 * L2 + NT - 1) / NT;
  static constexpr int NSG = (4 * L2 + NT - 1) / NT;
  static constexpr int NG = 30 * QF * FP;
  // per-element constants of the stage kernel (doubles, after gF): constant
  // state cn[5], linear modes lin[5][3], flux of the constant state cF[3][5],
  // its linearization H[3][3][5] (direction, reference dimension, field),
  // B_1 at the GLL points b1[Q]
  static constexpr int EC_CN = 0, EC_LIN = 5, EC_CF = 20, EC_H = 35, EC_B1 = 80, EC_N = 80 + Q;
  static constexpr int lds_stage = (2 * NA + NG) * 4 + EC_N * 8 + 64;
  static constexpr int lds_vterm = (2 * NA + 24 * QF * FP) * 4 + 64;
  // BR2: face-lift nodal values, metric rows, end traces (+ layer velocity)
  static constexpr bool br2_vel_in_b = 42 * Q * QP <= NA;
  static constexpr int lds_vterm_br2 = (2 * NA + 24 * Q * QP + 16 + 2 * Q + (br2_vel_in_b ? 0 : 18 * Q * QP)) * 4 + 64;
  static constexpr int FS = Q * Q * QP;
  static __device__ __forceinline__ int nidx(int n) { return (n / Q) * QP + (n - (n / Q) * Q); }
  // FP64 face traces of the updated state, in batches of TB fields over the
  // LDS of the stage kernel (as doubles): per field the modal block (later
  // reused for the Gauss-point passes W, O) and the three Bend collapses.
  static constexpr int KP = 3;
  static constexpr int TW = QF * cmax(2 * LP, KP * L), TO = cmax(2 * QF * FP, KP * QF2);
  static constexpr int TM = cmax(L * L * LP, TW + TO);
  static constexpr int TK = 2 * L * LP + 2 * L * LP + L2 * KP;
  static constexpr int TB0 = ((2 * NA + NG) / 2) / (TM + TK);
  static constexpr int TB = TB0 < 5 ? TB0 : 5;
};

// Final contraction of a chain over the first spatial dimension of
// [NF][Q][L][L] (pitch LP), accumulated into the MFMA result registers:
// acc[k][ob][r] += sum_i M[4ob+r][i] in[f][i][m][n] for the pencils
// G = tid + k*NT = f*L2 + m*L + n.
template <int P, int NF, int NS, int NTH = kStageThreads>
__device__ __forceinline__ void mchain_to_regs(int off, const float *in, f4 (&acc)[NS][M3<P>::NOB],
                                               int tid) {
  using S = M3<P, NTH>;
  constexpr int L = S::L, Q = S::Q, L2 = S::L2, LP = S::LP, NT = S::NT, NOB = S::NOB;
  float cf[NOB][Q];
  load_coef<L, Q>(off, tid, cf);
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    if (!wave_has(tid, k, NT, NF * L2)) continue;
    const int G0 = tid + k * NT;
    const int G = G0 < NF * L2 ? G0 : 0;
    const int f = G / L2, pp = G - f * L2, m = pp / L, n = pp - m * L;
    float x[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i) x[i] = in[((f * Q + i) * L + m) * LP + n];
#pragma unroll
    for (int ob = 0; ob < NOB; ++ob)
#pragma unroll
      for (int i = 0; i < Q; ++i) acc[k][ob] = mfma4(cf[ob][i], x[i], acc[k][ob]);
  }
}

// Projected divergence chain for reference direction D of NF nodal fields in
// A ([NF][Q][Q][Q], pitch QP); result accumulated into registers.
template <int P, int NF, int NS, int D, int NTH = kStageThreads>
__device__ __forceinline__ void mdiv_chain(float *A, float *B, f4 (&acc)[NS][M3<P>::NOB], int tid) {
  using S = M3<P, NTH>;
  using T = MT<P>;
  constexpr int L = S::L, Q = S::Q, LP = S::LP, QP = S::QP, NT = S::NT;
  constexpr int oc = D == 2 ? T::oQD : T::oBW, ob = D == 1 ? T::oQD : T::oBW, oa = D == 0 ? T::oQD : T::oBW;
  mcon_last<NF * Q * Q, Q, L, QP, LP, NT>(oc, A, B, tid);        // [NF Q Q][L]
  __syncthreads();
  mcon_mid<NF * Q, Q, 1, L, L, LP, LP, NT>(ob, B, A, tid);       // [NF Q][L][L]
  __syncthreads();
  mchain_to_regs<P, NF, NS, NTH>(oa, A, acc, tid);               // [NF][L][L][L]
  __syncthreads();
}

// modal [NF][L][L][L] (pitch LP) in A -> nodal [NF][Q][Q][Q] (pitch QP) in B
template <int P, int NF, int NTH = kStageThreads>
__device__ __forceinline__ void mbackward(float *A, float *B, int tid) {
  using S = M3<P, NTH>;
  constexpr int L = S::L, Q = S::Q, LP = S::LP, QP = S::QP, NT = S::NT, o = MT<P>::oBT;
  mcon_mid<NF, L, L, L, Q, LP, LP, NT>(o, A, B, tid);            // [NF][Q][L][L]
  __syncthreads();
  mcon_mid<NF * Q, L, 1, L, Q, LP, LP, NT>(o, B, A, tid);        // [NF Q][Q][L]
  __syncthreads();
  mcon_last<NF * Q * Q, L, Q, LP, QP, NT>(o, A, B, tid);         // [NF Q Q][Q]
  __syncthreads();
}

// Face blocks g[NFB][QF][QF] (pitch FP) -> H[NFB][L][L] (pitch LP) in B, via A.
template <int P, int NFB, int NTH = kStageThreads>
__device__ __forceinline__ void mface_project(const float *g, float *A, float *B, int tid) {
  using S = M3<P, NTH>;
  constexpr int L = S::L, QF = S::QF, LP = S::LP, FP = S::FP, NT = S::NT, o = MT<P>::oPW;
  mcon_mid<NFB, QF, 1, QF, L, FP, FP, NT>(o, g, A, tid);         // [NFB][L][QF]
  __syncthreads();
  mcon_last<NFB * L, QF, L, FP, LP, NT>(o, A, B, tid);           // [NFB][L][L]
  __syncthreads();
}

// Contribution at modal (l, m, n) of the six face blocks H[s] ([L][L], pitch
// LP) weighted by c[s].  Face normal dims: 4,2 -> a; 1,3 -> b; 0,5 -> c.
template <int P>
__device__ __forceinline__ float mface_sum(const float *H, int l, int m, int n, const float c[6]) {
  constexpr int L = P + 1, LP = odd(L), HB = L * LP;
  const double *B0 = CT<P>::Bend(), *B1 = CT<P>::Bend() + L;
  float s = 0.f;
  s = fmaf(c[4] * (float)B0[l], H[4 * HB + m * LP + n], s);
  s = fmaf(c[2] * (float)B1[l], H[2 * HB + m * LP + n], s);
  s = fmaf(c[1] * (float)B0[m], H[1 * HB + l * LP + n], s);
  s = fmaf(c[3] * (float)B1[m], H[3 * HB + l * LP + n], s);
  s = fmaf(c[0] * (float)B0[n], H[0 * HB + l * LP + m], s);
  s = fmaf(c[5] * (float)B1[n], H[5 * HB + l * LP + m], s);
  return s;
}

// This thread's entries (index tid + k NTH of [5][L3]) of element e's modal
// state, as registers.  The loads are issued here and consumed by store_split
// after independent work (the face loop), whose own loads they overlap.
template <int P, int NTH> struct ModalRegs {
  static constexpr int N = (5 * (P + 1) * (P + 1) * (P + 1) + NTH - 1) / NTH;
};
template <int P, int NTH, int N>
__device__ __forceinline__ void load_modal(const double *__restrict__ U, int e, int ne, double (&um)[N], int tid) {
  constexpr int L3 = (P + 1) * (P + 1) * (P + 1);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int i = tid + k * NTH;
    const int v = i / L3, m = i - v * L3;
    um[k] = i < 5 * L3 ? U[((size_t)v * ne + e) * L3 + m] : 0.0;
  }
}

// The FP32 part of the modal state into A ([5][L][L][L], pitch LP): every
// mode but the constant one, which stays FP64 (cn), and with LIN also but the
// three linear ones (lin).
template <int P, int NTH, bool LIN, int N>
__device__ __forceinline__ void store_split(const double (&um)[N], float *A, int tid) {
  using S = M3<P, NTH>;
  constexpr int L = S::L, L2 = S::L2, L3 = S::L3, LP = S::LP;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int i = tid + k * NTH;
    if (i < 5 * L3) {
      const int v = i / L3, m = i - v * L3;
      const bool low = m == 0 || (LIN && (m == 1 || m == L || m == L2));
      A[(m / L + v * L2) * LP + (m % L)] = low ? 0.f : (float)um[k];
    }
  }
}

// The constant modes' nodal values cn[v] (FP64) and, with lin, the linear
// parts: lin[v][x] = (1/2) u_x with u_x the coefficient of the mode linear
// along reference dimension x (a: index L^2, b: L, c: 1), so that the nodal
// value of the linear part is sum_x lin[v][x] B_1(xi_x) (B_0^2 = 1/2).
template <int P>
__device__ __forceinline__ void low_modes(const double *__restrict__ U, int e, int ne, double (&cn)[5],
                                          double (*lin)[3]) {
  constexpr int L = P + 1, L2 = L * L, L3 = L2 * L;
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const double *u = U + ((size_t)v * ne + e) * L3;
    cn[v] = u[0] * kPhi0;
    if (lin) {
      lin[v][0] = 0.5 * u[L2];
      lin[v][1] = 0.5 * u[L];
      lin[v][2] = 0.5 * u[1];
    }
  }
}

// Directional derivative of the contravariant Euler flux F(U).n at the state
// c along dU: dF = F'(c)[dU] (all arguments uniform across the workgroup).
__device__ __forceinline__ void euler_flux_jvp(const double c[5], double n0, double n1, double n2,
                                               double gm1, const double dU[5], double dF[5]) {
  const double ir = 1.0 / c[0];
  const double u = c[1] * ir, v = c[2] * ir, w = c[3] * ir;
  const double p = gm1 * (c[4] - 0.5 * (c[1] * u + c[2] * v + c[3] * w));
  const double vn = n0 * u + n1 * v + n2 * w;
  const double du = (dU[1] - u * dU[0]) * ir, dv = (dU[2] - v * dU[0]) * ir, dw = (dU[3] - w * dU[0]) * ir;
  const double dvn = n0 * du + n1 * dv + n2 * dw;
  const double dp = gm1 * (dU[4] - (u * dU[1] + v * dU[2] + w * dU[3]) + 0.5 * (u * u + v * v + w * w) * dU[0]);
  dF[0] = n0 * dU[1] + n1 * dU[2] + n2 * dU[3];
  dF[1] = dU[1] * vn + c[1] * dvn + n0 * dp;
  dF[2] = dU[2] * vn + c[2] * dvn + n1 * dp;
  dF[3] = dU[3] * vn + c[3] * dvn + n2 * dp;
  dF[4] = (dU[4] + dp) * vn + (c[4] + p) * dvn;
}

}  // namespace

#ifndef DGX_VTERM_M_THREADS
#define DGX_VTERM_M_THREADS 512
#endif
// Threads per element of the mixed viscous kernel.
constexpr int kVterm3mThreads = DGX_VTERM_M_THREADS;

struct MixedArgs {
  const double *T;       // face traces of the input state (FP64)
  const float *V;        // viscous volume term [ne][4][L3] (Navier-Stokes)
  const float *TV;       // viscous normal flux [ne][6][4][QF2] (Navier-Stokes)
  double *U;             // modal state (read; written when updating)
  double *Ut;            // LSRK register (update mode)
  double *R;             // residual output (residual mode)
  double *Tn;            // traces of the updated state (update mode)
  double ca, cb, dt;     // Ut = ca*Ut - dt*R ; U += cb*Ut
  int ne;
  int *flag;
};

// Residual (or fused LSRK14 stage) of Euler3d / Navier_Stokes3d, mixed precision.
template <int P, bool NS, bool UPDATE>
__global__ void __launch_bounds__(kStageThreads)
k_stage3m(MixedArgs a, Geo geo, Phys ph) {
  using S = M3<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, L3 = S::L3, Q3 = S::Q3, QF2 = S::QF2,
                LP = S::LP, FP = S::FP, NT = S::NT, NPT = S::NPT, NSL = S::NSL, NOB = S::NOB, FS = S::FS;
  extern __shared__ __attribute__((aligned(16))) float smf[];
  float *A = smf, *B = A + S::NA, *gF = B + S::NA;
  const int ne = a.ne, e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;
  const double gm1 = ph.gamma - 1.0;

  double *ec = reinterpret_cast<double *>(gF + S::NG);   // 8-byte aligned: 2 NA + NG is even
  // the element's metric terms into registers before any barrier
  double mt[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) mt[c] = geo.metric[9 * (size_t)e + c];
  double um[ModalRegs<P, NT>::N];
  load_modal<P, NT>(a.U, e, ne, um, tid);
  {
    double cn[5], lin[5][3];
    low_modes<P>(a.U, e, ne, cn, lin);
    if (tid < 5) {
      ec[S::EC_CN + tid] = cn[tid];
#pragma unroll
      for (int x = 0; x < 3; ++x) ec[S::EC_LIN + 3 * tid + x] = lin[tid][x];
    } else if (tid >= 16 && tid < 25) {
      // H[d][x] = F_d'(c)[lin_x]
      const int d = (tid - 16) / 3, x = (tid - 16) % 3;
      double dU[5], dF[5];
#pragma unroll
      for (int v = 0; v < 5; ++v) dU[v] = lin[v][x];
      const double *mg = geo.metric + 9 * (size_t)e + 3 * d;   // d varies by thread: not the register copy
      euler_flux_jvp(cn, mg[0], mg[1], mg[2], gm1, dU, dF);
#pragma unroll
      for (int v = 0; v < 5; ++v) ec[S::EC_H + (d * 3 + x) * 5 + v] = dF[v];
    } else if (tid >= 32 && tid < 35) {
      // cF[d] = F_d(c)
      const int d = tid - 32;
      const double *mg = geo.metric + 9 * (size_t)e + 3 * d;
      const double m0 = mg[0], m1 = mg[1], m2 = mg[2];
      const double ir = 1.0 / cn[0];
      const double u = cn[1] * ir, v = cn[2] * ir, w = cn[3] * ir;
      const double pc = gm1 * (cn[4] - 0.5 * (cn[1] * u + cn[2] * v + cn[3] * w));
      const double vn = m0 * u + m1 * v + m2 * w;
      ec[S::EC_CF + 5 * d] = cn[0] * vn;
      ec[S::EC_CF + 5 * d + 1] = cn[1] * vn + m0 * pc;
      ec[S::EC_CF + 5 * d + 2] = cn[2] * vn + m1 * pc;
      ec[S::EC_CF + 5 * d + 3] = cn[3] * vn + m2 * pc;
      ec[S::EC_CF + 5 * d + 4] = (cn[4] + pc) * vn;
    } else if (tid >= 64 && tid < 64 + Q) {
      ec[S::EC_B1 + tid - 64] = CT<P>::BT()[(tid - 64) * L + 1];
    }
  }
  // face corrections g = (J_side/J_vol) (Fhat - F(U-).n) [- BR1 viscous part], FP64
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
      if (NS && v > 0) {
        const float *vl = a.TV + ((size_t)e * 6 + s) * 4 * QF2;
        const float *vr = a.TV + ((size_t)nb * 6 + ns) * 4 * QF2;
        g += 0.5 * jr * ((double)vl[(v - 1) * QF2 + q] + (double)vr[(v - 1) * QF2 + qn]);
      }
      gF[((v * 6 + s) * QF + qa) * FP + qb] = (float)g;
    }
  }
  if (bad) atomicOr(a.flag, 1);
  store_split<P, NT, true>(um, A, tid);

  __syncthreads();
  mbackward<P, 5>(A, B, tid);
  // nodal state: constant and linear modes in FP64, the rest from the FP32 transform
  double un[NPT][5];
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int n = tid + r * NT;
    const int nc = n < Q3 ? n : 0;
    const int pn = S::nidx(nc);
    const double ba = ec[S::EC_B1 + nc / (Q * Q)], bb = ec[S::EC_B1 + (nc / Q) % Q], bc = ec[S::EC_B1 + nc % Q];
#pragma unroll
    for (int v = 0; v < 5; ++v)
      un[r][v] = ec[S::EC_CN + v] + ec[S::EC_LIN + 3 * v] * ba + ec[S::EC_LIN + 3 * v + 1] * bb +
                 ec[S::EC_LIN + 3 * v + 2] * bc + (n < Q3 ? (double)B[v * FS + pn] : 0.0);
  }


  // volume: projected divergence of the contravariant flux deviations
  f4 acc[NSL][NOB];
#pragma unroll
  for (int k = 0; k < NSL; ++k)
#pragma unroll
    for (int ob = 0; ob < NOB; ++ob) acc[k][ob] = f4{0.f, 0.f, 0.f, 0.f};
  // The flux is linearized about the constant state: F(c) + F'(c)[U_lin - c]
  // is linear in the reference coordinates, so it is subtracted at the nodes
  // (the FP32 chains see only the second-order remainder) and its projected
  // divergence, 2 sqrt(2) times its slope along the chain's direction at mode
  // (0,0,0) only, is added back in FP64 at the output.
  auto vol_dir = [&](auto Dc) {
    constexpr int d = decltype(Dc)::value;
    const double m0 = mt[3 * d], m1 = mt[3 * d + 1], m2 = mt[3 * d + 2];
    const double *cF = ec + S::EC_CF + 5 * d, *H = ec + S::EC_H + 15 * d;
    double vn[NPT], pr[NPT];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const double ir = 1.0 / un[r][0];
      const double ux = un[r][1] * ir, uy = un[r][2] * ir, uz = un[r][3] * ir;
      pr[r] = gm1 * (un[r][4] - 0.5 * (un[r][1] * ux + un[r][2] * uy + un[r][3] * uz));
      vn[r] = m0 * ux + m1 * uy + m2 * uz;
    }
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double c = cF[v], h0 = H[v], h1 = H[5 + v], h2 = H[10 + v];
      const double mv = v == 1 ? m0 : (v == 2 ? m1 : m2);
#pragma unroll
      for (int r = 0; r < NPT; ++r) {
        const int n = tid + r * NT;
        if (n < Q3) {
          const int pn = S::nidx(n);
          const double ba = ec[S::EC_B1 + n / (Q * Q)], bb = ec[S::EC_B1 + (n / Q) % Q];
          const double bc = ec[S::EC_B1 + n % Q];
          const double f = v == 0 ? un[r][0] * vn[r]
                         : (v == 4 ? (un[r][4] + pr[r]) * vn[r] : un[r][v] * vn[r] + mv * pr[r]);
          A[v * FS + pn] = (float)(((f - c) - h0 * ba) - h1 * bb - h2 * bc);
        }
      }
    }
    __syncthreads();
    mdiv_chain<P, 5, NSL, d>(A, B, acc, tid);
  };
  vol_dir(IC<0>());
  vol_dir(IC<1>());
  vol_dir(IC<2>());

  // the update's HBM inputs (viscous volume term, RK register) are loaded
  // before the face projection, whose barriers then cover their latency
  double vin[NSL][L], utin[NSL][L];
#pragma unroll
  for (int k = 0; k < NSL; ++k) {
    const int G = tid + k * NT;
    const int v = G / L2, pp = G - v * L2;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const bool in = G < 5 * L2;
      vin[k][l] = (NS && in && v > 0) ? (double)a.V[((size_t)e * 4 + (v - 1)) * L3 + l * L2 + pp] : 0.0;
      utin[k][l] = (UPDATE && in) ? a.Ut[((size_t)v * ne + e) * L3 + l * L2 + pp] : 0.0;
    }
  }

  // faces: H = (PW (x) PW) gF for 5 fields x 6 faces, in B
  mface_project<P, 30>(gF, A, B, tid);
  const float one6[6] = {1.f, 1.f, 1.f, 1.f, 1.f, 1.f};
  double unew[NSL][L];
#pragma unroll
  for (int k = 0; k < NSL; ++k) {
    const int G = tid + k * NT;
    if (G < 5 * L2) {
      const int v = G / L2, pp = G - v * L2, m = pp / L, nn = pp - m * L;
#pragma unroll
      for (int l = 0; l < L; ++l) {
        double r = (double)(acc[k][l / 4][l % 4] + mface_sum<P>(B + v * 6 * L * LP, l, m, nn, one6));
        if (l == 0 && pp == 0)   // 2 sqrt(3) sum_d H[d][d][v]
          r += 3.4641016151377545870 * (ec[S::EC_H + 0 * 15 + 0 * 5 + v] + ec[S::EC_H + 1 * 15 + 1 * 5 + v] +
                                        ec[S::EC_H + 2 * 15 + 2 * 5 + v]);
        if (NS && v > 0) r -= vin[k][l];
        const size_t gi = ((size_t)v * ne + e) * L3 + l * L2 + pp;
        if (!UPDATE) {
          a.R[gi] = r;
        } else {
          const double ut = a.ca * utin[k][l] - a.dt * r;
          const double u = fma(a.cb, ut, a.U[gi]);
          a.Ut[gi] = ut;
          a.U[gi] = u;
          unew[k][l] = u;
        }
      }
    }
  }
  if (!UPDATE) return;

  // FP64 face traces of the updated state for the next stage, TB fields at a time
  auto batch = [&](auto V0) {
    constexpr int v0 = decltype(V0)::value;
    constexpr int F = (5 - v0) < S::TB ? (5 - v0) : S::TB;
    double *Md = reinterpret_cast<double *>(smf);
    double *K0 = Md + F * S::TM, *K1 = K0 + F * 2 * L * LP, *K2 = K1 + F * 2 * L * LP;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NSL; ++k) {
      const int G = tid + k * NT;
      if (G < 5 * L2) {
        const int v = G / L2, pp = G - v * L2, m = pp / L, nn = pp - m * L;
        if (v >= v0 && v < v0 + F) {
#pragma unroll
          for (int l = 0; l < L; ++l) Md[(((v - v0) * L + l) * L + m) * LP + nn] = unew[k][l];
        }
      }
    }
    __syncthreads();
    constexpr int KP = S::KP;
    con_mid<F, L, L, L, 2, LP, LP, NT>(CT<P>::Bend(), Md, K0, tid);       // [F][2][L][L]
    con_mid<F * L, L, 1, L, 2, LP, LP, NT>(CT<P>::Bend(), Md, K1, tid);   // [F][L][2][L]
    con_last<F * L2, L, 2, LP, KP, NT>(CT<P>::Bend(), Md, K2, tid);       // [F][L][L][2]
    __syncthreads();
    double *W = Md, *O = Md + F * S::TW;
    double *Te = a.Tn + (size_t)e * 6 * 5 * QF2;
    // faces 4 / 2: [F][2][QF][QF]
    con_mid<2 * F, L, 1, L, QF, LP, LP, NT>(CT<P>::Bg(), K0, W, tid);     // [2F][QF][L]
    __syncthreads();
    con_last<2 * F * QF, L, QF, LP, FP, NT>(CT<P>::Bg(), W, O, tid);      // [2F][QF][QF]
    __syncthreads();
    for (int i = tid; i < 2 * F * QF2; i += NT) {
      const int ve = i / QF2, q = i - ve * QF2, v = ve >> 1, end = ve & 1;
      Te[((end ? 2 : 4) * 5 + v0 + v) * QF2 + q] = O[(ve * QF + q / QF) * FP + q % QF];
    }
    __syncthreads();
    // faces 1 / 3: [F][QF][2][QF]
    con_mid<F, L, 2, L, QF, LP, LP, NT>(CT<P>::Bg(), K1, W, tid);         // [F][QF][2][L]
    __syncthreads();
    con_last<2 * F * QF, L, QF, LP, FP, NT>(CT<P>::Bg(), W, O, tid);      // [F][QF][2][QF]
    __syncthreads();
    for (int i = tid; i < 2 * F * QF2; i += NT) {
      const int v = i / (2 * QF2), r = i - v * 2 * QF2, end = r / QF2, q = r - end * QF2;
      const int ia = q / QF, ic = q - ia * QF;
      Te[((end ? 3 : 1) * 5 + v0 + v) * QF2 + q] = O[((v * QF + ia) * 2 + end) * FP + ic];
    }
    __syncthreads();
    // faces 0 / 5: [F][QF][QF][2]
    con_mid<F, L, L, 2, QF, KP, KP, NT>(CT<P>::Bg(), K2, W, tid);         // [F][QF][L][2]
    __syncthreads();
    con_mid<F * QF, L, 1, 2, QF, KP, KP, NT>(CT<P>::Bg(), W, O, tid);     // [F][QF][QF][2]
    __syncthreads();
    for (int i = tid; i < 2 * F * QF2; i += NT) {
      const int v = i / (2 * QF2), r = i - v * 2 * QF2, end = r / QF2, q = r - end * QF2;
      Te[((end ? 5 : 0) * 5 + v0 + v) * QF2 + q] = O[(v * QF2 + q) * KP + end];
    }
  };
  batch(IC<0>());
  if constexpr (S::TB < 5) batch(IC<S::TB>());
  if constexpr (2 * S::TB < 5) batch(IC<2 * S::TB>());
  if constexpr (3 * S::TB < 5) batch(IC<3 * S::TB>());
  if constexpr (4 * S::TB < 5) batch(IC<4 * S::TB>());
}

// Face-lift contribution of the two faces normal to reference direction D at
// modal (l, m, n) of block set H ([6][L][L], pitch LP), with weights clo, chi
// for the face at -1 and +1 (faces 4/2 for a, 1/3 for b, 0/5 for c), FP32.
template <int P, int D>
__device__ __forceinline__ float mface_pair(const float *H, int l, int m, int n, float clo, float chi) {
  constexpr int L = P + 1, LP = odd(L), HB = L * LP;
  const double *B0 = CT<P>::Bend(), *B1 = CT<P>::Bend() + L;
  if (D == 0) return fmaf(clo * (float)B0[l], H[4 * HB + m * LP + n], chi * (float)B1[l] * H[2 * HB + m * LP + n]);
  if (D == 1) return fmaf(clo * (float)B0[m], H[1 * HB + l * LP + n], chi * (float)B1[m] * H[3 * HB + l * LP + n]);
  return fmaf(clo * (float)B0[n], H[0 * HB + l * LP + m], chi * (float)B1[n] * H[5 * HB + l * LP + m]);
}

// Viscous flux of equation q contracted with (n0, n1, n2) from the stress
// tensor t (t11 t12 t13 t22 t23 t33), heat flux h and velocity, FP32.
__device__ __forceinline__ float mvisc_flux(const float t[6], const float h[3], float u, float v, float w,
                                            int q, float n0, float n1, float n2) {
  const float r0 = t[0] * n0 + t[1] * n1 + t[2] * n2;
  const float r1 = t[1] * n0 + t[3] * n1 + t[4] * n2;
  const float r2 = t[2] * n0 + t[4] * n1 + t[5] * n2;
  if (q == 0) return r0;
  if (q == 1) return r1;
  if (q == 2) return r2;
  return (r0 * u + r1 * v + r2 * w) + (h[0] * n0 + h[1] * n1 + h[2] * n2);
}

// BR1 viscous terms, mixed precision: V[e][eq][L3] (FP32) = modal divergence of
// the viscous flux, TV[e][s][eq][QF2] (FP32) = viscous flux . n at the Gauss
// points of face s.  The lifted gradient is assembled one reference direction
// at a time (affine faces: normals -+ grad xi_d / |grad xi_d|) straight into
// the stress tensor and heat flux at the thread's nodes.
// At p = 7 the BR2 variant needs a few registers more than the 128 that keep
// two 512-thread workgroups per CU: capped there (4 registers spill).
template <int P, bool BR2>
__global__ void __launch_bounds__(kVterm3mThreads) __attribute__((amdgpu_waves_per_eu((BR2 && P == 7) ? 4 : 1)))
k_vterm3m(const double *__restrict__ U, const double *__restrict__ Tr, float *__restrict__ V,
          float *__restrict__ TV, Geo geo, Phys ph, int ne) {
  using S = M3<P, kVterm3mThreads>;
  using T = MT<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, L3 = S::L3, Q3 = S::Q3, QF2 = S::QF2,
                LP = S::LP, QP = S::QP, FP = S::FP, NT = S::NT, NPT = S::NPT, NSG = S::NSG, NOB = S::NOB,
                FS = S::FS;
  extern __shared__ __attribute__((aligned(16))) float smf[];
  float *A = smf, *B = A + S::NA, *Jq = B + S::NA;
  const int e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;

  double cn[5];
  double um[ModalRegs<P, NT>::N];
  load_modal<P, NT>(U, e, ne, um, tid);
  low_modes<P>(U, e, ne, cn, nullptr);
  // primitive jumps J_side/J_vol ({q} - q-) at the Gauss points (average_flux '+'), FP64
  for (int i = tid; i < 6 * QF2; i += NT) {
    const int s = i / QF2, q = i - s * QF2;
    const int es = e * 6 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int qn = geo.perm[geo.permid[es] * QF2 + q];
    const double jr = geo.fjr[es];
    const double *tl = Tr + ((size_t)e * 6 + s) * 5 * QF2;
    const double *tr = Tr + ((size_t)nb * 6 + ns) * 5 * QF2;
    double pl[4], pq[4];
    {
      const double r = tl[q], u = tl[QF2 + q] / r, v = tl[2 * QF2 + q] / r, w = tl[3 * QF2 + q] / r;
      double en = tl[4 * QF2 + q] / r;
      double wk = u * u; wk = v * v + wk; wk = w * w + wk;
      pl[0] = u; pl[1] = v; pl[2] = w; pl[3] = -0.5 * wk + en;
    }
    {
      const double r = tr[qn], u = tr[QF2 + qn] / r, v = tr[2 * QF2 + qn] / r, w = tr[3 * QF2 + qn] / r;
      double en = tr[4 * QF2 + qn] / r;
      double wk = u * u; wk = v * v + wk; wk = w * w + wk;
      pq[0] = u; pq[1] = v; pq[2] = w; pq[3] = -0.5 * wk + en;
    }
    const int qa = q / QF, qb = q - qa * QF;
#pragma unroll
    for (int c = 0; c < 4; ++c)
      Jq[((c * 6 + s) * QF + qa) * FP + qb] = (float)(jr * (0.5 * (pl[c] + pq[c]) - pl[c]));
  }
  store_split<P, NT, false>(um, A, tid);
  __syncthreads();
  mbackward<P, 5, NT>(A, B, tid);
  // primitives: deviations from the constant state (FP32) for the gradient
  // chains, full velocities (FP32) for the viscous work
  const double icr = 1.0 / cn[0];
  const double cu = cn[1] * icr, cv = cn[2] * icr, cw = cn[3] * icr;
  const double ce = cn[4] * icr - 0.5 * ((cu * cu + cv * cv) + cw * cw);
  float pv[NPT][4], uw[NPT][3];
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int n = tid + r * NT;
    const int pn = S::nidx(n < Q3 ? n : 0);
    const bool ok = n < Q3;
    const double rho = cn[0] + (ok ? (double)B[pn] : 0.0);
    const double u = (cn[1] + (ok ? (double)B[FS + pn] : 0.0)) / rho;
    const double v = (cn[2] + (ok ? (double)B[2 * FS + pn] : 0.0)) / rho;
    const double w = (cn[3] + (ok ? (double)B[3 * FS + pn] : 0.0)) / rho;
    const double en = (cn[4] + (ok ? (double)B[4 * FS + pn] : 0.0)) / rho;
    double wk = u * u; wk = v * v + wk; wk = w * w + wk;
    pv[r][0] = (float)(u - cu); pv[r][1] = (float)(v - cv); pv[r][2] = (float)(w - cw);
    pv[r][3] = (float)((-0.5 * wk + en) - ce);
    uw[r][0] = (float)u; uw[r][1] = (float)v; uw[r][2] = (float)w;
  }
  __syncthreads();
  // face lift of the jumps: (PW (x) PW) Jq -> H[4][6][L][L], kept in the Jq region
  mcon_mid<24, QF, 1, QF, L, FP, FP, NT>(T::oPW, Jq, A, tid);     // [24][L][QF]
  __syncthreads();
  mcon_last<24 * L, QF, L, FP, LP, NT>(T::oPW, A, Jq, tid);     // [24][L][L]
  __syncthreads();
  const float *H = Jq;

  const double *mt = geo.metric + 9 * (size_t)e;
  const float mu = (float)ph.vis, kap = (float)(ph.vis * ph.gamma / ph.Pr);
  float tq[NPT][9];   // t11 t12 t13 t22 t23 t33, kappa grad e
#pragma unroll
  for (int r = 0; r < NPT; ++r)
#pragma unroll
    for (int j = 0; j < 9; ++j) tq[r][j] = 0.f;
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
    f4 D[NSG][NOB];
#pragma unroll
    for (int k = 0; k < NSG; ++k)
#pragma unroll
      for (int ob = 0; ob < NOB; ++ob) D[k][ob] = f4{0.f, 0.f, 0.f, 0.f};
    mdiv_chain<P, 4, NSG, d, NT>(A, B, D, tid);
    const double g0d = mt[3 * d], g1d = mt[3 * d + 1], g2d = mt[3 * d + 2];
    const float inv = (float)(1.0 / sqrt(g0d * g0d + g1d * g1d + g2d * g2d));
#pragma unroll
    for (int k = 0; k < NSG; ++k) {
      const int Gi = tid + k * NT;
      if (Gi < 4 * L2) {
        const int c = Gi / L2, pp = Gi - c * L2, m = pp / L, nn = pp - m * L;
#pragma unroll
        for (int l = 0; l < L; ++l)
          A[((c * L + l) * L + m) * LP + nn] =
              D[k][l / 4][l % 4] + mface_pair<P, d>(H + c * 6 * L * LP, l, m, nn, -inv, inv);
      }
    }
    __syncthreads();
    mbackward<P, 4, NT>(A, B, tid);
    const float g0 = (float)g0d, g1 = (float)g1d, g2 = (float)g2d;
    const float c43 = mu * (4.f / 3.f), c23 = mu * (2.f / 3.f);
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int n = tid + r * NT;
      const int pn = S::nidx(n < Q3 ? n : 0);
      const float au = B[pn], av = B[FS + pn], aw = B[2 * FS + pn], ae = B[3 * FS + pn];
      const float ux = g0 * au, uy = g1 * au, uz = g2 * au;
      const float vx = g0 * av, vy = g1 * av, vz = g2 * av;
      const float wx = g0 * aw, wy = g1 * aw, wz = g2 * aw;
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
    __syncthreads();
  };
  grad_dir(IC<0>());
  grad_dir(IC<1>());
  grad_dir(IC<2>());

  f4 acc[NSG][NOB];   // modal viscous volume term of the 4 equations
  // BR2 scratch: Hn [24][Q][Q] and gs[9], KE[2][Q] in the Jq region; the
  // (BR2 - BR1) flux difference Df [4][6][Q][Q] in B, the face-layer velocity
  // Vel [6][3][Q][Q] after it (or after KE when B is too small)
  constexpr int HN = 24 * Q * QP, VOFF = S::br2_vel_in_b ? 0 : HN + 16 + 2 * Q;
  float *Df = B, *Vel = S::br2_vel_in_b ? B + HN : Jq + VOFF;
  if constexpr (BR2) {
    // nodal values of the face lifts at each face's own GLL points: Hn[24][Q][Q]
    mcon_mid<24, L, 1, L, Q, LP, LP, NT>(T::oBT, Jq, A, tid);   // [24][Q][L]
    if (tid < 9) {   // unit rows of the metric
      const int d = tid / 3;
      Jq[HN + tid] = (float)(mt[tid] / sqrt(mt[3 * d] * mt[3 * d] + mt[3 * d + 1] * mt[3 * d + 1] +
                                            mt[3 * d + 2] * mt[3 * d + 2]));
    }
    if (tid >= 16 && tid < 16 + 2 * Q) Jq[HN + tid] = (float)CT<P>::KE()[tid - 16];
    __syncthreads();
    mcon_last<24 * Q, L, Q, LP, QP, NT>(T::oBT, A, Jq, tid);    // [24][Q][Q]
    __syncthreads();
  }

  // normal viscous flux on the GLL face layers [6][4][Q][Q] (pitch QP) -> Gauss
  // points (with BR2, also the velocity on the layers)
  float *Lay = A;
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int n = tid + r * NT;
    if (n < Q3) {
      const int i = n / (Q * Q), j = (n / Q) % Q, k = n % Q;
      // face f at layer position (a1, a2); one call per face the node lies on,
      // each with a constant face (no indexed arrays, which would live in scratch)
      auto put = [&](int f, int a1, int a2) {
        const double *nn = geo.fnorm + 3 * (e * 6 + f);
        const float n0 = (float)nn[0], n1 = (float)nn[1], n2 = (float)nn[2];
#pragma unroll
        for (int q = 0; q < 4; ++q)
          Lay[((f * 4 + q) * Q + a1) * QP + a2] =
              mvisc_flux(tq[r], tq[r] + 6, uw[r][0], uw[r][1], uw[r][2], q, n0, n1, n2);
        if constexpr (BR2) {
#pragma unroll
          for (int q = 0; q < 3; ++q) Vel[((f * 3 + q) * Q + a1) * QP + a2] = uw[r][q];
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
    const float eta = (float)ph.eta;
    for (int idx = tid; idx < 6 * Q * Q; idx += NT) {
      const int f = idx / (Q * Q), rem = idx - f * (Q * Q), l1 = rem / Q, l2 = rem - l1 * Q;
      int i, j, k;
      if (f == 0 || f == 5) { i = l1; j = l2; k = f == 0 ? 0 : Q - 1; }
      else if (f == 1 || f == 3) { i = l1; k = l2; j = f == 1 ? 0 : Q - 1; }
      else { j = l1; k = l2; i = f == 4 ? 0 : Q - 1; }
      const double *nn = geo.fnorm + 3 * (e * 6 + f);
      const float *vl = Vel + (f * 3 * Q + l1) * QP + l2;
      float fd[4];
      br2_flux<Q, QP>(fd, Jq, Jq + HN + 16, f, i, j, k, Jq + HN, mu, kap, eta, vl[0], vl[Q * QP],
                      vl[2 * Q * QP], (float)nn[0], (float)nn[1], (float)nn[2]);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        Lay[((f * 4 + q) * Q + l1) * QP + l2] += fd[q];
        Df[((q * 6 + f) * Q + l1) * QP + l2] = fd[q];
      }
    }
    __syncthreads();
  }
  float *Gs = B;   // scratch of the GLL -> Gauss interpolation
  if constexpr (BR2) {
    // lift of the flux difference on the element's own faces, into the volume term
    mcon_mid<24, Q, 1, Q, L, QP, QP, NT>(T::oPI, B, Jq, tid);   // [24][L][Q]
    __syncthreads();
    mcon_last<24 * L, Q, L, QP, LP, NT>(T::oPI, Jq, B, tid);    // [24][L][L]
    __syncthreads();
    float jr6[6];
#pragma unroll
    for (int s = 0; s < 6; ++s) jr6[s] = (float)geo.fjr[e * 6 + s];
#pragma unroll
    for (int k = 0; k < NSG; ++k) {
      const int Gi = tid + k * NT;
      const int q = Gi / L2, pp = Gi - q * L2, m = pp / L, nn = pp - m * L;
#pragma unroll
      for (int ob = 0; ob < NOB; ++ob) acc[k][ob] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int l = 0; l < L; ++l)
        acc[k][l / 4][l % 4] = Gi < 4 * L2 ? mface_sum<P>(B + q * 6 * L * LP, l, m, nn, jr6) : 0.f;
    }
    Gs = Jq;
  } else {
#pragma unroll
    for (int k = 0; k < NSG; ++k)
#pragma unroll
      for (int ob = 0; ob < NOB; ++ob) acc[k][ob] = f4{0.f, 0.f, 0.f, 0.f};
  }
  mcon_mid<24, Q, 1, Q, QF, QP, QP, NT>(T::oIg, Lay, Gs, tid);   // [24][QF][Q]
  __syncthreads();
  mcon_last<24 * QF, Q, QF, QP, FP, NT>(T::oIg, Gs, A, tid);     // [24][QF][QF]
  __syncthreads();
  for (int i = tid; i < 24 * QF2; i += NT) {
    const int sq = i / QF2, q = i - sq * QF2;
    TV[(size_t)e * 24 * QF2 + i] = A[(sq * QF + q / QF) * FP + q % QF];
  }
  __syncthreads();

  // viscous volume divergence of the 4 equations
  auto visc_dir = [&](auto Dc) {
    constexpr int d = decltype(Dc)::value;
    const float m0 = (float)mt[3 * d], m1 = (float)mt[3 * d + 1], m2 = (float)mt[3 * d + 2];
#pragma unroll
    for (int r = 0; r < NPT; ++r) {
      const int n = tid + r * NT;
      if (n < Q3) {
        const int pn = S::nidx(n);
#pragma unroll
        for (int q = 0; q < 4; ++q)
          A[q * FS + pn] = mvisc_flux(tq[r], tq[r] + 6, uw[r][0], uw[r][1], uw[r][2], q, m0, m1, m2);
      }
    }
    __syncthreads();
    mdiv_chain<P, 4, NSG, d, NT>(A, B, acc, tid);
  };
  visc_dir(IC<0>());
  visc_dir(IC<1>());
  visc_dir(IC<2>());
#pragma unroll
  for (int k = 0; k < NSG; ++k) {
    const int Gi = tid + k * NT;
    if (Gi < 4 * L2) {
      const int q = Gi / L2, pp = Gi - q * L2;
#pragma unroll
      for (int l = 0; l < L; ++l) V[((size_t)e * 4 + q) * L3 + l * L2 + pp] = acc[k][l / 4][l % 4];
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

// one matrix M[no][n] (row-major doubles) into the MFMA layout [NOB][4][NPAD]
void pack_mq(const double *M, int no, int n, float *out) {
  const int nob = (no + 3) / 4, npad = (n + 3) / 4 * 4;
  for (int ob = 0; ob < nob; ++ob)
    for (int r = 0; r < 4; ++r)
      for (int k = 0; k < npad; ++k) {
        const int o = 4 * ob + r;
        out[(ob * 4 + r) * npad + k] = (o < no && k < n) ? (float)M[o * n + k] : 0.f;
      }
}

template <int P>
hipError_t launch_mixed3_p(const StageLaunch &s) {
  using S = M3<P>;
  if (s.op == 0) {
    if (s.phys.visc == 1) {
      const int sh = M3<P, kVterm3mThreads>::lds_vterm_br2;
      (void)hipFuncSetAttribute((const void *)k_vterm3m<P, true>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
      hipLaunchKernelGGL((k_vterm3m<P, true>), dim3(s.ne), dim3(kVterm3mThreads), sh, s.stream, s.U, s.T, s.Vfw,
                         s.TVfw, s.geo, s.phys, s.ne);
    } else {
      const int sh = M3<P, kVterm3mThreads>::lds_vterm;
      (void)hipFuncSetAttribute((const void *)k_vterm3m<P, false>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
      hipLaunchKernelGGL((k_vterm3m<P, false>), dim3(s.ne), dim3(kVterm3mThreads), sh, s.stream, s.U, s.T, s.Vfw,
                         s.TVfw, s.geo, s.phys, s.ne);
    }
    return hipGetLastError();
  }
  MixedArgs a;
  a.T = s.T; a.V = s.Vf; a.TV = s.TVf; a.U = s.U; a.Ut = s.Ut; a.R = s.R; a.Tn = s.Tn;
  a.ca = s.ca; a.cb = s.cb; a.dt = s.dt; a.ne = s.ne; a.flag = s.flag;
  const int sh = S::lds_stage;
#define DGX_MSTAGE(NSV, UPD)                                                                         \
  do {                                                                                               \
    (void)hipFuncSetAttribute((const void *)k_stage3m<P, NSV, UPD>,                                  \
                              hipFuncAttributeMaxDynamicSharedMemorySize, sh);                       \
    hipLaunchKernelGGL((k_stage3m<P, NSV, UPD>), dim3(s.ne), dim3(kStageThreads), sh, s.stream, a, s.geo,      \
                       s.phys);                                                                      \
  } while (0)
  if (s.ns) { if (s.update) DGX_MSTAGE(true, true); else DGX_MSTAGE(true, false); }
  else { if (s.update) DGX_MSTAGE(false, true); else DGX_MSTAGE(false, false); }
#undef DGX_MSTAGE
  return hipGetLastError();
}

}  // namespace

bool mixed3_supported(int p) { return p >= 1 && p <= kMaxMixed; }

hipError_t launch_mixed3(const StageLaunch &s) {
  switch (s.p) {
    case 1: return launch_mixed3_p<1>(s);
    case 2: return launch_mixed3_p<2>(s);
    case 3: return launch_mixed3_p<3>(s);
    case 4: return launch_mixed3_p<4>(s);
    case 5: return launch_mixed3_p<5>(s);
    case 6: return launch_mixed3_p<6>(s);
    case 7: return launch_mixed3_p<7>(s);
    case 8: return launch_mixed3_p<8>(s);
    default: return hipErrorInvalidValue;
  }
}

// FP64 tables of every order (packed as in dgx_device.hpp) -> this unit's
// constant memory: the FP64 tables for the trace constants and face sums, and
// the FP32 MFMA tables.
hipError_t upload_tables_m3d(const double *h) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), h, sizeof(double) * kTabTotal);
  if (e != hipSuccess) return e;
  std::vector<float> mq(kMqTotal, 0.f);
  for (int p = 1; p <= kMaxMixed; ++p) {
    const int L = p + 1, Q = p + 2, QF = p + 1;
    const double *t = h + tab_offset(p);
    const double *BT = t, *BW = BT + Q * L, *QD = BW + L * Q, *Bg = QD + L * Q, *Bend = Bg + QF * L;
    const double *PW = Bend + 2 * L, *Ig = PW + L * QF;
    float *o = mq.data() + mq_order_offset(p);
    pack_mq(BT, Q, L, o); o += mq_size(Q, L);
    pack_mq(BW, L, Q, o); o += mq_size(L, Q);
    pack_mq(QD, L, Q, o); o += mq_size(L, Q);
    pack_mq(Bg, QF, L, o); o += mq_size(QF, L);
    pack_mq(Bend, 2, L, o); o += mq_size(2, L);
    pack_mq(PW, L, QF, o); o += mq_size(L, QF);
    pack_mq(Ig, QF, Q, o); o += mq_size(QF, Q);
    pack_mq(Ig + QF * Q, L, Q, o); o += mq_size(L, Q);   // PI follows Ig in the FP64 tables
    if (o != mq.data() + mq_order_offset(p + 1)) return hipErrorInvalidValue;
  }
  return hipMemcpyToSymbol(HIP_SYMBOL(c_mq), mq.data(), sizeof(float) * kMqTotal);
}

}  // namespace dgx
