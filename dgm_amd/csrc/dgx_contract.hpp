// FP64 sum-factorization contractions over LDS tensors, shared by the fused
// hex kernels.  One thread owns one pencil at a time and computes a chunk of
// outputs from its inputs; chunks are outermost so that only a chunk of
// coefficient rows (constant memory, scalar registers) is live at a time.
#pragma once
#include "dgx_device.hpp"

#ifndef DGX_STAGE_THREADS
#define DGX_STAGE_THREADS 256
#endif

namespace dgx {

// Threads per workgroup (one element per workgroup) of the fused hex kernels.
constexpr int kStageThreads = DGX_STAGE_THREADS;

constexpr int odd(int n) { return n | 1; }
template <int V> struct IC { static constexpr int value = V; };
constexpr int cmax(int a, int b) { return a > b ? a : b; }


// Coefficient rows per pass (about 40 doubles live in scalar registers).
template <int NI> constexpr int out_chunk() { return NI * 3 <= 40 ? 3 : 2; }

// [A][N][B][C] (last-dim pitch PI) -> [A][NO][B][C] (pitch PO), contracting N.
// When every thread owns at most one pencil (NP <= NT), the pencil's inputs
// are read from LDS once and all output chunks are computed from registers;
// otherwise the chunks are outermost (inputs re-read per chunk) so that only
// one chunk of coefficient rows is live in scalar registers across the
// pencil loop.  Both orders give bitwise the same sums.
template <int A, int N, int B, int C, int NO, int PI, int PO, int NT, bool ACC = false>
__device__ __forceinline__ void con_mid(const double *__restrict__ M, const double *in,
                                        double *out, int tid) {
  constexpr int NP = A * B * C, OC = out_chunk<N>();
  auto pencil = [&](int pc, int o0, int o1) {
    const int c = pc % C, t = pc / C, b = t % B, a = t / B;
    double x[N];
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = in[((a * N + i) * B + b) * PI + c];
#pragma unroll
    for (int o = o0; o < o1; ++o) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < N; ++i) s = fma(M[o * N + i], x[i], s);
      double *dst = out + ((a * NO + o) * B + b) * PO + c;
      if (ACC) *dst += s; else *dst = s;
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

// [R][NI] (pitch PI) -> [R][NO] (pitch PO), contracting the last dimension
// (loop order as in con_mid).
template <int R, int NI, int NO, int PI, int PO, int NT>
__device__ __forceinline__ void con_last(const double *__restrict__ M, const double *in, double *out,
                                         int tid) {
  constexpr int OC = out_chunk<NI>();
  auto pencil = [&](int r, int o0, int o1) {
    double x[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) x[i] = in[r * PI + i];
#pragma unroll
    for (int o = o0; o < o1; ++o) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < NI; ++i) s = fma(M[o * NI + i], x[i], s);
      out[r * PO + o] = s;
    }
  };
  if constexpr (R <= NT) {
    if (tid < R) pencil(tid, 0, NO);
  } else {
#pragma unroll
    for (int o0 = 0; o0 < NO; o0 += OC)
      for (int r = tid; r < R; r += NT) pencil(r, o0, o0 + OC < NO ? o0 + OC : NO);
  }
}

// BR2 correction of the normal viscous flux at GLL node (i, j, k) on face f of
// an affine hex (faces: 0 z-, 1 y-, 2 x+, 3 y+, 4 x-, 5 z+).  The nodal
// stresses come from the gradient lifted over all six faces,
// G = grad_h q + sum_d r_d; BR2 evaluates the face flux with grad_h q + eta r_f
// instead, so fd receives the flux (equations 1..4, contracted with n) of the
// velocity / energy gradient increment
//   -sum_d r_d + eta (lifting of face f alone).
// Hn[(c*6 + s)][Q][Q] (pitch QP) holds the nodal values, at face s's GLL
// points, of the modal face lift of primitive c (u, v, w, e) on face s, without
// the 1/|grad xi_d| factor; KE[2][Q] = sum_l phi_l(-+1) phi_l(xi_i).  The
// lifting across reference direction d at a node is then
//   r_d = (KE[1][xi_d] Hn[face d+] - KE[0][xi_d] Hn[face d-]) / |grad xi_d|,
// with physical components grad(xi_d) r_d; gs[3d + x] = grad(xi_d)_x / |grad xi_d|.
// A gradient increment g (x) a (du_i/dx_j = g_j a_i) has stress . n
//   mu (a_i (g.n) + g_i (a.n)) - 2/3 mu (g.a) n_i
// and heat flux . n = kappa a_e (g.n).
template <int Q, int QP, typename R, typename RT>
__device__ __forceinline__ void br2_flux(R (&fd)[4], const R *Hn, const RT *KE, int f, int i, int j, int k,
                                         const R *gs, R mu, R kap, R eta, R u, R v, R w, R n0, R n1, R n2) {
  const int df = (f == 2 || f == 4) ? 0 : ((f == 1 || f == 3) ? 1 : 2);
  const bool hi = (f == 2 || f == 3 || f == 5);
  const R c23 = mu * R(2.0 / 3.0);
  R r0 = R(0), r1 = R(0), r2 = R(0), he = R(0);
#pragma unroll 1
  for (int d = 0; d < 3; ++d) {   // one direction at a time: fewer live LDS loads
    const int slo = d == 0 ? 4 : (d == 1 ? 1 : 0), shi = d == 0 ? 2 : (d == 1 ? 3 : 5);
    const int x = d == 0 ? i : (d == 1 ? j : k);
    const int t1 = d == 0 ? j : i, t2 = d == 2 ? j : k;
    const R klo = R(KE[x]), khi = R(KE[Q + x]);
    R a[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const R hl = Hn[((c * 6 + slo) * Q + t1) * QP + t2], hh = Hn[((c * 6 + shi) * Q + t1) * QP + t2];
      R val = klo * hl - khi * hh;                                // -r_d |grad xi_d|
      if (d == df) val += eta * (hi ? khi * hh : -(klo * hl));   // + eta * lifting of face f
      a[c] = val;
    }
    const R g0 = gs[3 * d], g1 = gs[3 * d + 1], g2 = gs[3 * d + 2];
    const R gn = g0 * n0 + g1 * n1 + g2 * n2;
    const R an = a[0] * n0 + a[1] * n1 + a[2] * n2;
    const R ga = g0 * a[0] + g1 * a[1] + g2 * a[2];
    r0 += mu * (a[0] * gn + g0 * an) - c23 * ga * n0;
    r1 += mu * (a[1] * gn + g1 * an) - c23 * ga * n1;
    r2 += mu * (a[2] * gn + g2 * an) - c23 * ga * n2;
    he += kap * a[3] * gn;
  }
  fd[0] = r0;
  fd[1] = r1;
  fd[2] = r2;
  fd[3] = (r0 * u + r1 * v + r2 * w) + he;
}

}  // namespace dgx
