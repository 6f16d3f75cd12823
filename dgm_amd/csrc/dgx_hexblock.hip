// Element-block seam: batched hex inner product and backward transform over a
// block of elements that share one basis (GpGpu::hex_kernels,
// src/GpGpu.hpp:24-57, src/GpGpu.cu:236-262 and 505-750; driven by
// HexBlock::forward_transform / backward_transform, src/HexBlock.cpp:35-130).
//
// Per element, as Hex::inner_product (src/Hex.cpp:703-747):
//   f *= wJ;  H[i][b][c] = sum_a Ba[i][a] f[a][b][c];
//   Ha[i][j][c] = sum_b Bb[j][b] H[i][b][c];  fh[i][j][k] = sum_c Bc[k][c] Ha[i][j][c]
// and as Hex::backward_transform (src/Hex.cpp:796-812):
//   H[a][j][k] = sum_i Ba[i][a] fh[i][j][k];  Ha[a][b][k] = sum_j Bb[j][b] H[a][j][k];
//   f[a][b][c] = sum_k Bc[k][c] Ha[a][b][k].
// Each element occupies a slot of qa*qb*qc doubles; the modal coefficients
// are the first nm^3 of the slot (the layout hex_kernels uses).  Both
// transforms work in place: a workgroup reads its element's whole slot into
// LDS before it writes any of it.
//
// The sums run in ascending index order with separate multiplies and adds, so
// the results are the same bits as a plain loop in that order.  The work is
// HBM-bound: per element qtot doubles in and nm^3 out (or the reverse) against
// qtot + 2 (nm qtot + nm^2 qb qc + nm^3 qc) flops, 3.2 flop/B at nm = 8,
// q = 9, below MI355X's FP64 ridge of 9.8 flop/B.  The kernels stream elements
// through a grid of resident workgroups (as many as fit on the chip), each
// holding the three basis matrices in LDS for all of its elements.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/dgx.h"
#include "dgx_kernels.hpp"

namespace dgx {

constexpr int kHbThreads = 256;
constexpr int kHbMax = 16;   // modes and points per direction

struct HbArgs {
  double *u;            // [ne][qa*qb*qc] slots
  const double *B;      // Ba [nm][qa], Bb [nm][qb], Bc [nm][qc], then their transposes
  const double *wJ;     // [qa*qb*qc] shared, or [ne][qa*qb*qc]
  long long ne;
  int nm, qa, qb, qc;
  int wj_stride;        // 0: one wJ for the block; qa*qb*qc: one per element
};

// NMC, QC > 0: sizes known at compile time (nm = NMC, qa = qb = qc = QC), so
// the contraction loops unroll; 0: sizes from the arguments.
template <int NMC, int QC>
__global__ void __launch_bounds__(kHbThreads) k_hb_inner(HbArgs a) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  const int nm = NMC ? NMC : a.nm, qa = QC ? QC : a.qa, qb = QC ? QC : a.qb, qc = QC ? QC : a.qc;
  const int qt = qa * qb * qc, tid = threadIdx.x;
  double *Ba = sm, *Bb = Ba + nm * qa, *Bc = Bb + nm * qb;
  double *X = Bc + nm * qc;          // [qa][qb][qc], later Ha [nm][nm][qc]
  double *H = X + qt;                // [nm][qb][qc]
  for (int i = tid; i < nm * (qa + qb + qc); i += kHbThreads) Ba[i] = a.B[i];
  const int bc = qb * qc;
  // compile-time shapes load the next element's slot into registers while the
  // current one is contracted (the barriers wait on LDS only)
  constexpr int NPT = QC ? (QC * QC * QC + kHbThreads - 1) / kHbThreads : 1;
  double pre[NPT];
  auto fetch = [&](long long e) {
    if constexpr (QC > 0) {
      const double *ue = a.u + e * qt;
#pragma unroll
      for (int r = 0; r < NPT; ++r) {
        const int i = tid + r * kHbThreads;
        pre[r] = (e < a.ne && i < qt) ? ue[i] : 0.0;
      }
    }
  };
  fetch(blockIdx.x);
  for (long long e = blockIdx.x; e < a.ne; e += gridDim.x) {
    double *ue = a.u + e * qt;
    const double *w = a.wJ + e * a.wj_stride;
    __syncthreads();   // basis loaded / previous element's Ha read
    if constexpr (QC > 0) {
#pragma unroll
      for (int r = 0; r < NPT; ++r) {
        const int i = tid + r * kHbThreads;
        if (i < qt) X[i] = pre[r] * w[i];
      }
      fetch(e + gridDim.x);
    } else {
      for (int i = tid; i < qt; i += kHbThreads) X[i] = ue[i] * w[i];
    }
    __syncthreads();
    for (int o = tid; o < nm * bc; o += kHbThreads) {
      const int i = o / bc, r = o - i * bc;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < qa; ++k) s = s + Ba[i * qa + k] * X[k * bc + r];
      H[o] = s;
    }
    __syncthreads();
    for (int o = tid; o < nm * nm * qc; o += kHbThreads) {
      const int i = o / (nm * qc), r = o - i * (nm * qc), j = r / qc, c = r - j * qc;
      const double *h = H + i * bc + c;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < qb; ++k) s = s + Bb[j * qb + k] * h[k * qc];
      X[o] = s;
    }
    __syncthreads();
    for (int o = tid; o < nm * nm * nm; o += kHbThreads) {
      const int ij = o / nm, k = o - ij * nm;
      const double *h = X + ij * qc;
      double s = 0.0;
#pragma unroll
      for (int c = 0; c < qc; ++c) s = s + h[c] * Bc[k * qc + c];
      ue[o] = s;
    }
  }
}

template <int NMC, int QC>
__global__ void __launch_bounds__(kHbThreads) k_hb_backward(HbArgs a) {
#pragma clang fp contract(off)
  extern __shared__ double sm[];
  const int nm = NMC ? NMC : a.nm, qa = QC ? QC : a.qa, qb = QC ? QC : a.qb, qc = QC ? QC : a.qc;
  const int qt = qa * qb * qc, tid = threadIdx.x;
  const int n2 = nm * nm, n3 = n2 * nm;
  double *Ba = sm, *Bb = Ba + nm * qa, *Bc = Bb + nm * qb;
  double *X = Bc + nm * qc;                  // fh [nm][nm][nm], later Ha [qa][qb][nm]
  double *H = X + std::max(n3, qa * qb * nm);   // [qa][nm][nm]
  for (int i = tid; i < nm * (qa + qb + qc); i += kHbThreads) Ba[i] = a.B[i];
  constexpr int NPT = NMC ? (NMC * NMC * NMC + kHbThreads - 1) / kHbThreads : 1;
  double pre[NPT];
  auto fetch = [&](long long e) {
    if constexpr (NMC > 0) {
      const double *ue = a.u + e * qt;
#pragma unroll
      for (int r = 0; r < NPT; ++r) {
        const int i = tid + r * kHbThreads;
        pre[r] = (e < a.ne && i < n3) ? ue[i] : 0.0;
      }
    }
  };
  fetch(blockIdx.x);
  for (long long e = blockIdx.x; e < a.ne; e += gridDim.x) {
    double *ue = a.u + e * qt;
    __syncthreads();
    if constexpr (NMC > 0) {
#pragma unroll
      for (int r = 0; r < NPT; ++r) {
        const int i = tid + r * kHbThreads;
        if (i < n3) X[i] = pre[r];
      }
      fetch(e + gridDim.x);
    } else {
      for (int i = tid; i < n3; i += kHbThreads) X[i] = ue[i];
    }
    __syncthreads();
    for (int o = tid; o < qa * n2; o += kHbThreads) {
      const int p = o / n2, r = o - p * n2;
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < nm; ++i) s = s + Ba[i * qa + p] * X[i * n2 + r];
      H[o] = s;
    }
    __syncthreads();
    for (int o = tid; o < qa * qb * nm; o += kHbThreads) {
      const int p = o / (qb * nm), r = o - p * (qb * nm), b = r / nm, k = r - b * nm;
      const double *h = H + p * n2 + k;
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < nm; ++j) s = s + Bb[j * qb + b] * h[j * nm];
      X[o] = s;
    }
    __syncthreads();
    for (int o = tid; o < qt; o += kHbThreads) {
      const int ab = o / qc, c = o - ab * qc;
      const double *h = X + ab * nm;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < nm; ++k) s = s + h[k] * Bc[k * qc + c];
      ue[o] = s;
    }
  }
}

// One wavefront per element for the GLL shapes (NM modes, Q = NM + 1 points):
// each lane owns whole pencils and computes all NM (or Q) outputs of a pencil
// from one pass over its inputs, which cuts the LDS reads of the contractions
// about NM-fold against one output per thread, and a wavefront needs no
// workgroup barrier (its LDS operations complete in order).  The basis is read
// with uniform addresses (scalar loads).  Per wavefront LDS: X and H of
// max(Q^3, NM Q^2) doubles each.  Same sums, in the same order, as the
// generic kernels.
constexpr int kHbWaveThreads = 128;

template <int NM, int Q> struct HbW {
  static constexpr int QT = Q * Q * Q, BC = Q * Q, N2 = NM * NM, N3 = N2 * NM;
  static constexpr int XS = QT > NM * BC ? QT : NM * BC;   // doubles per buffer
};

// read-only global data through the scalar (constant) path: uniform loads of
// the basis become s_load into SGPRs instead of one vector load per lane
typedef const __attribute__((address_space(4))) double cdouble;

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NM, int Q>
__global__ void __launch_bounds__(kHbWaveThreads) k_hb_inner_w(HbArgs a) {
#pragma clang fp contract(off)
  using W = HbW<NM, Q>;
  constexpr int QT = W::QT, BC = W::BC;
  extern __shared__ double sm[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = kHbWaveThreads / 64;
  double *X = sm + wv * 2 * W::XS, *H = X + W::XS;
  // the next element's slot is loaded into registers while this one is
  // contracted, so a wavefront waits for HBM only on its first element
  constexpr int NPL = (QT + 63) / 64;
  const long long step = (long long)gridDim.x * nw;
  double pre[NPL];
  auto fetch = [&](long long e) {
    const double *ue = a.u + e * QT;
#pragma unroll
    for (int r = 0; r < NPL; ++r) {
      const int i = lane + 64 * r;
      pre[r] = (e < a.ne && i < QT) ? ue[i] : 0.0;
    }
  };
  fetch((long long)blockIdx.x * nw + wv);
  for (long long e = (long long)blockIdx.x * nw + wv; e < a.ne; e += step) {
    // the basis is reloaded (scalar cache) for every element rather than held
    // across the loop, which would need far more SGPRs than there are
    const double *bp = a.B;
    asm volatile("" : "+s"(bp));
    cdouble *Ba = (cdouble *)bp, *Bb = Ba + NM * Q, *Bc = Bb + NM * Q;
    double *ue = a.u + e * QT;
    const double *w = a.wJ + e * a.wj_stride;
#pragma unroll
    for (int r = 0; r < NPL; ++r) {
      const int i = lane + 64 * r;
      if (i < QT) X[i] = pre[r] * w[i];
    }
    fetch(e + step);
    wave_lds_sync();
    for (int r = lane; r < BC; r += 64) {          // pencils along a: H[i][r]
      double x[Q];
#pragma unroll
      for (int k = 0; k < Q; ++k) x[k] = X[k * BC + r];
#pragma unroll 2
      for (int i = 0; i < NM; ++i) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < Q; ++k) s = s + Ba[i * Q + k] * x[k];
        H[i * BC + r] = s;
      }
    }
    wave_lds_sync();
    for (int p = lane; p < NM * Q; p += 64) {      // pencils along b: Ha[i][j][c] into X
      const int i = p / Q, c = p - i * Q;
      double h[Q];
#pragma unroll
      for (int k = 0; k < Q; ++k) h[k] = H[i * BC + k * Q + c];
#pragma unroll 2
      for (int j = 0; j < NM; ++j) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < Q; ++k) s = s + Bb[j * Q + k] * h[k];
        X[(i * NM + j) * Q + c] = s;
      }
    }
    wave_lds_sync();
    for (int ij = lane; ij < W::N2; ij += 64) {    // pencils along c: fh[i][j][k] staged in H
      double h[Q];
#pragma unroll
      for (int c = 0; c < Q; ++c) h[c] = X[ij * Q + c];
#pragma unroll 2
      for (int k = 0; k < NM; ++k) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < Q; ++c) s = s + h[c] * Bc[k * Q + c];
        H[ij * NM + k] = s;
      }
    }
    wave_lds_sync();
    for (int i = lane; i < W::N3; i += 64) ue[i] = H[i];
    wave_lds_sync();
  }
}

template <int NM, int Q>
__global__ void __launch_bounds__(kHbWaveThreads) k_hb_backward_w(HbArgs a) {
#pragma clang fp contract(off)
  using W = HbW<NM, Q>;
  constexpr int QT = W::QT, N2 = W::N2, N3 = W::N3;
  extern __shared__ double sm[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = kHbWaveThreads / 64;
  double *X = sm + wv * 2 * W::XS, *H = X + W::XS;
  constexpr int NPL = (N3 + 63) / 64;
  const long long step = (long long)gridDim.x * nw;
  double pre[NPL];
  auto fetch = [&](long long e) {
    const double *ue = a.u + e * QT;
#pragma unroll
    for (int r = 0; r < NPL; ++r) {
      const int i = lane + 64 * r;
      pre[r] = (e < a.ne && i < N3) ? ue[i] : 0.0;
    }
  };
  fetch((long long)blockIdx.x * nw + wv);
  for (long long e = (long long)blockIdx.x * nw + wv; e < a.ne; e += step) {
    // the basis is reloaded (scalar cache) for every element rather than held
    // across the loop, which would need far more SGPRs than there are
    const double *bp = a.B;
    asm volatile("" : "+s"(bp));
    // transposed copies (rows of Q x NM), so that each output reads a
    // contiguous run of the basis
    cdouble *BaT = (cdouble *)bp + 3 * NM * Q, *BbT = BaT + NM * Q, *BcT = BbT + NM * Q;
    double *ue = a.u + e * QT;
#pragma unroll
    for (int r = 0; r < NPL; ++r) {
      const int i = lane + 64 * r;
      if (i < N3) X[i] = pre[r];
    }
    fetch(e + step);
    wave_lds_sync();
    for (int r = lane; r < N2; r += 64) {          // pencils along i: H[a][j][k]
      double x[NM];
#pragma unroll
      for (int i = 0; i < NM; ++i) x[i] = X[i * N2 + r];
#pragma unroll 2
      for (int p = 0; p < Q; ++p) {
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i) s = s + BaT[p * NM + i] * x[i];
        H[p * N2 + r] = s;
      }
    }
    wave_lds_sync();
    for (int pk = lane; pk < Q * NM; pk += 64) {   // pencils along j: Ha[a][b][k] into X
      const int p = pk / NM, k = pk - p * NM;
      double h[NM];
#pragma unroll
      for (int j = 0; j < NM; ++j) h[j] = H[p * N2 + j * NM + k];
#pragma unroll 2
      for (int b = 0; b < Q; ++b) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NM; ++j) s = s + BbT[b * NM + j] * h[j];
        X[(p * Q + b) * NM + k] = s;
      }
    }
    wave_lds_sync();
    for (int ab = lane; ab < Q * Q; ab += 64) {    // pencils along k: f[a][b][c] staged in H
      double h[NM];
#pragma unroll
      for (int k = 0; k < NM; ++k) h[k] = X[ab * NM + k];
#pragma unroll 2
      for (int c = 0; c < Q; ++c) {
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < NM; ++k) s = s + h[k] * BcT[c * NM + k];
        H[ab * Q + c] = s;
      }
    }
    wave_lds_sync();
    for (int i = lane; i < QT; i += 64) ue[i] = H[i];
    wave_lds_sync();
  }
}

}  // namespace dgx

struct dgx_hexblock {
  long long ne = 0;
  int nm = 0, qa = 0, qb = 0, qc = 0, qt = 0, device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false, have_basis = false, have_wj = false;
  int wj_stride = 0;
  double *d_u = nullptr, *d_B = nullptr, *d_wJ = nullptr;
  int ncu = 1;
};

namespace {

int fail(int code, const std::string &msg) { return dgx::set_error(code, msg); }

int hip_fail(const char *what, hipError_t e) {
  return fail(DGX_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

size_t lds_inner(const dgx_hexblock *h) {
  return sizeof(double) * ((size_t)h->nm * (h->qa + h->qb + h->qc) + h->qt + (size_t)h->nm * h->qb * h->qc);
}

size_t lds_backward(const dgx_hexblock *h) {
  const size_t n3 = (size_t)h->nm * h->nm * h->nm;
  return sizeof(double) * ((size_t)h->nm * (h->qa + h->qb + h->qc) + std::max(n3, (size_t)h->qa * h->qb * h->nm) +
                           (size_t)h->qa * h->nm * h->nm);
}

// Run one transform on the caller's device buffer or, for host memory, on the
// block's own device copy (the reference copies in and out on every call).
int run(dgx_hexblock *h, double *u, int on_device, bool inner) {
  if (!h || !u) return fail(DGX_ERR_ARGUMENT, "null argument");
  if (!h->have_basis) return fail(DGX_ERR_ARGUMENT, "hexblock: set the basis (dgx_hexblock_set_basis) first");
  if (inner && !h->have_wj) return fail(DGX_ERR_ARGUMENT, "hexblock: set wJ (dgx_hexblock_set_wJ) first");
  if (on_device != 0 && on_device != 1) return fail(DGX_ERR_ARGUMENT, "on_device must be 0 or 1");
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess) return fail(DGX_ERR_DEVICE, "hipGetDevice failed");
  if (cur != h->device && hipSetDevice(h->device) != hipSuccess) return fail(DGX_ERR_DEVICE, "hipSetDevice failed");
  const size_t bytes = sizeof(double) * (size_t)h->ne * h->qt;
  double *d = on_device ? u : h->d_u;
  hipError_t e = hipSuccess;
  if (!on_device) e = hipMemcpyAsync(d, u, bytes, hipMemcpyHostToDevice, h->stream);
  dgx::HbArgs a{d, h->d_B, h->d_wJ, h->ne, h->nm, h->qa, h->qb, h->qc, h->wj_stride};
  const size_t sh = inner ? lds_inner(h) : lds_backward(h);
  if (e == hipSuccess) {
    // the GLL shapes of orders 1..10 (q = nm + 1 points in every direction)
    // run one wavefront per element; any other shape the generic kernels
    void (*kern)(dgx::HbArgs) = inner ? dgx::k_hb_inner<0, 0> : dgx::k_hb_backward<0, 0>;
    int threads = dgx::kHbThreads, per_wg = 1;   // elements a workgroup holds at once
    size_t lds = sh;
    if (h->qa == h->nm + 1 && h->qb == h->qa && h->qc == h->qa) {
      switch (h->nm) {
#define DGX_HB(NM)                                                                                   \
  case NM:                                                                                           \
    kern = inner ? dgx::k_hb_inner_w<NM, NM + 1> : dgx::k_hb_backward_w<NM, NM + 1>;                   \
    per_wg = dgx::kHbWaveThreads / 64;                                                               \
    lds = sizeof(double) * 2 * dgx::HbW<NM, NM + 1>::XS * per_wg;                                    \
    break;
        DGX_HB(2) DGX_HB(3) DGX_HB(4) DGX_HB(5) DGX_HB(6) DGX_HB(7) DGX_HB(8) DGX_HB(9) DGX_HB(10) DGX_HB(11)
#undef DGX_HB
        default: break;
      }
      if (per_wg > 1) threads = dgx::kHbWaveThreads;
    }
    (void)hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    // resident workgroups only, so that every one of them streams the same
    // share of elements; no more than the elements need
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) != hipSuccess || per_cu < 1)
      per_cu = 1;
    const int grid = (int)std::min<long long>((h->ne + per_wg - 1) / per_wg, (long long)per_cu * h->ncu);
    hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, h->stream, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess && !on_device) e = hipMemcpyAsync(u, d, bytes, hipMemcpyDeviceToHost, h->stream);
  if (e == hipSuccess && !on_device) e = hipStreamSynchronize(h->stream);
  if (cur != h->device) (void)hipSetDevice(cur);
  return e == hipSuccess ? DGX_OK : hip_fail(inner ? "hexblock inner_product" : "hexblock backward_transform", e);
}

}  // namespace

extern "C" {

int dgx_hexblock_create(long long ne, int nmodes, int qa, int qb, int qc, void *stream, dgx_hexblock_handle *out) {
  if (!out) return fail(DGX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (ne < 1) return fail(DGX_ERR_ARGUMENT, "hexblock: ne must be positive");
  if (nmodes < 1 || nmodes > dgx::kHbMax || qa < 1 || qa > dgx::kHbMax || qb < 1 || qb > dgx::kHbMax || qc < 1 ||
      qc > dgx::kHbMax)
    return fail(DGX_ERR_ARGUMENT, "hexblock: modes and points per direction must be in 1.." +
                                      std::to_string(dgx::kHbMax));
  if (nmodes > qa || nmodes > qb || nmodes > qc)
    return fail(DGX_ERR_ARGUMENT, "hexblock: nmodes must not exceed qa, qb or qc (the modes share the element's slot)");
  auto *h = new dgx_hexblock;
  h->ne = ne; h->nm = nmodes; h->qa = qa; h->qb = qb; h->qc = qc; h->qt = qa * qb * qc;
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) {
    if (stream) {
      h->stream = static_cast<hipStream_t>(stream);
    } else {
      e = hipStreamCreate(&h->stream);
      h->own_stream = e == hipSuccess;
    }
  }
  int ncu = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, h->device);
  if (e == hipSuccess) e = hipMalloc(&h->d_B, 2 * sizeof(double) * (size_t)nmodes * (qa + qb + qc));
  if (e == hipSuccess) e = hipMalloc(&h->d_u, sizeof(double) * (size_t)ne * h->qt);
  if (e != hipSuccess) {
    dgx_hexblock_destroy(h);
    return fail(e == hipErrorOutOfMemory ? DGX_ERR_MEMORY : DGX_ERR_DEVICE,
                std::string("hexblock create: ") + hipGetErrorString(e));
  }
  h->ncu = std::max(ncu, 1);
  *out = h;
  return DGX_OK;
}

int dgx_hexblock_destroy(dgx_hexblock_handle h) {
  if (!h) return DGX_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->d_u) (void)hipFree(h->d_u);
  if (h->d_B) (void)hipFree(h->d_B);
  if (h->d_wJ) (void)hipFree(h->d_wJ);
  if (h->own_stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return DGX_OK;
}

int dgx_hexblock_set_wJ(dgx_hexblock_handle h, const double *wJ, int per_element) {
  if (!h || !wJ) return fail(DGX_ERR_ARGUMENT, "null argument");
  if (per_element != 0 && per_element != 1) return fail(DGX_ERR_ARGUMENT, "per_element must be 0 or 1");
  const size_t n = (size_t)h->qt * (per_element ? (size_t)h->ne : 1);
  if (h->d_wJ) { (void)hipStreamSynchronize(h->stream); (void)hipFree(h->d_wJ); h->d_wJ = nullptr; }
  h->have_wj = false;
  hipError_t e = hipMalloc(&h->d_wJ, sizeof(double) * n);
  if (e == hipSuccess) e = hipMemcpy(h->d_wJ, wJ, sizeof(double) * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail("hexblock set_wJ", e);
  h->wj_stride = per_element ? h->qt : 0;
  h->have_wj = true;
  return DGX_OK;
}

int dgx_hexblock_set_basis(dgx_hexblock_handle h, const double *Ba, const double *Bb, const double *Bc) {
  if (!h || !Ba || !Bb || !Bc) return fail(DGX_ERR_ARGUMENT, "null argument");
  (void)hipStreamSynchronize(h->stream);
  const size_t na = (size_t)h->nm * h->qa, nb = (size_t)h->nm * h->qb, nc = (size_t)h->nm * h->qc;
  std::vector<double> all(2 * (na + nb + nc));
  size_t o = 0;
  for (const double *B : {Ba, Bb, Bc}) {
    const int q = B == Ba ? h->qa : B == Bb ? h->qb : h->qc;
    std::copy(B, B + (size_t)h->nm * q, all.begin() + o);
    o += (size_t)h->nm * q;
  }
  for (const double *B : {Ba, Bb, Bc}) {   // transposes [q][nm]
    const int q = B == Ba ? h->qa : B == Bb ? h->qb : h->qc;
    for (int p = 0; p < q; ++p)
      for (int i = 0; i < h->nm; ++i) all[o + (size_t)p * h->nm + i] = B[(size_t)i * q + p];
    o += (size_t)h->nm * q;
  }
  hipError_t e = hipMemcpy(h->d_B, all.data(), sizeof(double) * all.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_fail("hexblock set_basis", e);
  h->have_basis = true;
  return DGX_OK;
}

int dgx_hexblock_inner_product(dgx_hexblock_handle h, double *u, int on_device) { return run(h, u, on_device, true); }

int dgx_hexblock_backward_transform(dgx_hexblock_handle h, double *u, int on_device) {
  return run(h, u, on_device, false);
}

int dgx_hexblock_synchronize(dgx_hexblock_handle h) {
  if (!h) return fail(DGX_ERR_ARGUMENT, "null argument");
  const hipError_t e = hipStreamSynchronize(h->stream);
  return e == hipSuccess ? DGX_OK : hip_fail("hexblock synchronize", e);
}

}  // extern "C"
