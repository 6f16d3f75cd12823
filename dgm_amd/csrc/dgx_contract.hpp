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
template <int A, int N, int B, int C, int NO, int PI, int PO, int NT, bool ACC = false>
__device__ __forceinline__ void con_mid(const double *__restrict__ M, const double *in,
                                        double *out, int tid) {
  constexpr int NP = A * B * C, OC = out_chunk<N>();
#pragma unroll
  for (int o0 = 0; o0 < NO; o0 += OC) {
    for (int pc = tid; pc < NP; pc += NT) {
      const int c = pc % C, t = pc / C, b = t % B, a = t / B;
      double x[N];
#pragma unroll
      for (int i = 0; i < N; ++i) x[i] = in[((a * N + i) * B + b) * PI + c];
#pragma unroll
      for (int o = o0; o < (o0 + OC < NO ? o0 + OC : NO); ++o) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) s = fma(M[o * N + i], x[i], s);
        double *dst = out + ((a * NO + o) * B + b) * PO + c;
        if (ACC) *dst += s; else *dst = s;
      }
    }
  }
}

// [R][NI] (pitch PI) -> [R][NO] (pitch PO), contracting the last dimension.
template <int R, int NI, int NO, int PI, int PO, int NT>
__device__ __forceinline__ void con_last(const double *__restrict__ M, const double *in, double *out,
                                         int tid) {
  constexpr int OC = out_chunk<NI>();
#pragma unroll
  for (int o0 = 0; o0 < NO; o0 += OC) {
    for (int r = tid; r < R; r += NT) {
      double x[NI];
#pragma unroll
      for (int i = 0; i < NI; ++i) x[i] = in[r * PI + i];
#pragma unroll
      for (int o = o0; o < (o0 + OC < NO ? o0 + OC : NO); ++o) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NI; ++i) s = fma(M[o * NI + i], x[i], s);
        out[r * PO + o] = s;
      }
    }
  }
}

}  // namespace dgx
