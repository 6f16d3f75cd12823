// Variable-order quads: one polynomial order per element (root.ord, read by
// Field::Field, src/Field.cpp:2290-2320), 2D Euler on affine quads.
//
// The two sides of a link carry max(L_left, L_right) Gauss points
// (Side::match, src/Side.cpp:106-117, via Field::adjust_side_order,
// src/Field.cpp:3208-3230).  Per element of order P:
//   traces    the modal state evaluated at the side's n Gauss points
//             (Quad::fill_sides' GL2G interpolation of the GLL edge values,
//             src/Quad.cpp:845-869: exact, since the element's trace is a
//             polynomial of degree P <= n - 1);
//   flux      the Riemann flux at those n points, neighbour points reversed
//             on flipped sides (sideMap, src/BaseQuad.hpp:53-57);
//   lift      Quad::local_lift (src/Quad.cpp:1971-2038): G2GL interpolation
//             (Q x n) of J_side/J_vol (F^ - F(U-).n) to the element's GLL
//             edge nodes, divided by the end weight, then projected with the
//             element's own basis.  Projected, a side's nodal row is the
//             modal face block PWv U_side with PWv = (B diag(w)) G2GL (L x n),
//             times the end value of the normal-direction modes.  For
//             n = L this is the uniform kernels' Gauss inner product; for
//             n > L the GLL rule is not exact for the product, so PWv is the
//             matrix product itself (built on the host in long double).
// The reference's default build takes this path for every order pair: its
// var_p_lift needs DGM_SIDE_FLUX_TO_VOLUME, which its build never defines.
//
// Layouts (padded): modal [field][elem][(pmax+1)^2], element e's (P_e+1)^2
// modes first (index l*L_e + m), zeros after; nodal [field][elem][(pmax+2)^2]
// with the element's (P_e+2)^2 nodes first; traces [elem][side][field][QFM].
// Elements are launched in groups of one order, each group over its own list.
#include "dgx_kernels.hpp"

namespace dgx {

template <int P>
__global__ void __launch_bounds__(64) k_vtraces2(VarpArgs a) {
  constexpr int L = P + 1, L2 = L * L, NT = 64, NF = 4;
  __shared__ double A[NF * L2], Kb[NF * L * 2], Ka[NF * 2 * L];
  const int e = a.elems[blockIdx.x], tid = threadIdx.x, QFM = a.QFM;
  for (int i = tid; i < NF * L2; i += NT) {
    const int v = i / L2;
    A[i] = a.U[((size_t)v * a.ne + e) * a.NM + (i - v * L2)];
  }
  __syncthreads();
  contract<NF, L, L, 2, 2, NT>(CT<P>::Bend(), A, Kb, tid);   // [NF][L][2]: edges 0 / 2
  contract<NF, L, L, 1, 2, NT>(CT<P>::Bend(), A, Ka, tid);   // [NF][2][L]: edges 3 / 1
  __syncthreads();
  double *Te = a.T + (size_t)e * 4 * NF * QFM;
  for (int s = 0; s < 4; ++s) {
    const int n = a.qs[e * 4 + s];
    const double *Bg = a.vtab + vp_block(P, n, 0);   // [n][L]
    for (int i = tid; i < NF * n; i += NT) {
      const int v = i / n, g = i - v * n;
      double acc = 0.0;
#pragma unroll
      for (int l = 0; l < L; ++l) {
        double k;
        if (s == 0) k = Kb[(v * L + l) * 2];
        else if (s == 2) k = Kb[(v * L + l) * 2 + 1];
        else if (s == 3) k = Ka[(v * 2) * L + l];
        else k = Ka[(v * 2 + 1) * L + l];
        acc = fma(Bg[g * L + l], k, acc);
      }
      Te[(s * NF + v) * QFM + g] = acc;
    }
  }
}

template <int P>
__global__ void __launch_bounds__(64) k_vrhs2(VarpArgs a) {
  constexpr int L = P + 1, Q = P + 2, L2 = L * L, Q2 = Q * Q, NT = 64, NF = 4;
  constexpr int NPT = (Q2 + NT - 1) / NT;
  constexpr int QFX = kMaxOrder + 1;   // side points, at most
  __shared__ double M[NF * L2], X1[NF * Q * L], X2[NF * Q * L], N1[NF * Q2], N2[NF * Q2];
  __shared__ double gF[NF * 4 * QFX], H[NF * 4 * L];
  const int e = a.elems[blockIdx.x], tid = threadIdx.x, ne = a.ne, QFM = a.QFM;
  const Geo &geo = a.geo;
  const Phys &ph = a.phys;
  const double gm1 = ph.gamma - 1.0;
  for (int i = tid; i < NF * L2; i += NT) {
    const int v = i / L2;
    M[i] = a.U[((size_t)v * ne + e) * a.NM + (i - v * L2)];
  }
  __syncthreads();
  contract<NF, L, L, 1, Q, NT>(CT<P>::BT(), M, X1, tid);   // [NF][Q][L]
  __syncthreads();
  contract<NF, Q, L, 2, Q, NT>(CT<P>::BT(), X1, N1, tid);  // [NF][Q][Q]
  __syncthreads();
  double un[NPT][NF];
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int k = tid + r * NT;
#pragma unroll
    for (int v = 0; v < NF; ++v) un[r][v] = k < Q2 ? N1[v * Q2 + k] : 1.0;
  }
  // face corrections g = J_side/J_vol (F^ - F(U-).n) at every side's points
  bool bad = false;
  for (int i = tid; i < 4 * QFX; i += NT) {
    const int s = i / QFX, q = i - s * QFX;
    const int es = e * 4 + s;
    const int n = a.qs[es];
    if (q >= n) continue;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int qn = a.flip[es] ? n - 1 - q : q;
    const double nx = geo.fnorm[3 * es], ny = geo.fnorm[3 * es + 1];
    const double jr = geo.fjr[es];
    double UL[4], UR[4], FL[4], FR[4], Fh[4], lL, lR;
    const double *tl = a.T + ((size_t)e * 4 + s) * NF * QFM;
#pragma unroll
    for (int v = 0; v < 4; ++v) UL[v] = tl[v * QFM + q];
    if (nb >= 0) {
      const double *tr = a.T + ((size_t)nb * 4 + ns) * NF * QFM;
#pragma unroll
      for (int v = 0; v < 4; ++v) UR[v] = tr[v * QFM + qn];
    } else {
      ghost2(-1 - nb, UL, nx, ny, ph.bcU, UR);
    }
    bool bl, br;
    euler_normal_flux2(UL, nx, ny, gm1, FL, lL, bl);
    euler_normal_flux2(UR, nx, ny, gm1, FR, lR, br);
    bad |= bl | br;
    if (ph.flux == 1) {
      roe_flux2(UL, UR, nx, ny, ph.gamma, FL, Fh);
    } else {
      const double lmax = fmax(lL, lR);
#pragma unroll
      for (int v = 0; v < 4; ++v) Fh[v] = 0.5 * (FL[v] + FR[v] + lmax * (UL[v] - UR[v]));
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) gF[(v * 4 + s) * QFX + q] = jr * (Fh[v] - FL[v]);
  }
  if (bad) atomicOr(a.flag, 1);
  // contravariant volume fluxes at the nodes (each thread its own nodes)
  const double *mt = geo.metric + 9 * (size_t)e;
  const double rx = mt[0], ry = mt[1], sx = mt[3], sy = mt[4];
#pragma unroll
  for (int rr = 0; rr < NPT; ++rr) {
    const int k = tid + rr * NT;
    if (k >= Q2) continue;
    const double r = un[rr][0], ru = un[rr][1], rv = un[rr][2], rE = un[rr][3];
    const double ir = 1.0 / r, u = ru * ir, vv = rv * ir;
    const double p = gm1 * (rE - 0.5 * (ru * u + rv * vv));
    const double h = rE + p;
    const double fx[4] = {ru, ru * u + p, rv * u, u * h}, fy[4] = {rv, ru * vv, rv * vv + p, vv * h};
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      N1[v * Q2 + k] = rx * fx[v] + ry * fy[v];
      N2[v * Q2 + k] = sx * fx[v] + sy * fy[v];
    }
  }
  __syncthreads();
  // modal face blocks H[v][s][l] = sum_g PWv(P, n_s)[l][g] gF[v][s][g]
  for (int i = tid; i < NF * 4 * L; i += NT) {
    const int vs = i / L, l = i - vs * L, s = vs & 3;
    const int n = a.qs[e * 4 + s];
    const double *PW = a.vtab + vp_block(P, n, 1);   // [L][n]
    const double *g = gF + vs * QFX;
    double acc = 0.0;
    for (int k = 0; k < n; ++k) acc = fma(PW[l * n + k], g[k], acc);
    H[i] = acc;
  }
  contract<NF, Q, Q, 2, L, NT>(CT<P>::BW(), N1, X1, tid);
  contract<NF, Q, Q, 2, L, NT>(CT<P>::QD(), N2, X2, tid);
  __syncthreads();
  contract<NF, Q, L, 1, L, NT>(CT<P>::QD(), X1, M, tid);
  __syncthreads();
  contract<NF, Q, L, 1, L, NT, true>(CT<P>::BW(), X2, M, tid);
  __syncthreads();
  const double cw[4] = {1.0, 1.0, 1.0, 1.0};
  for (int i = tid; i < NF * a.NM; i += NT) {
    const int v = i / a.NM, k = i - v * a.NM;
    double r = 0.0;
    if (k < L2) {
      const int l = k / L, m = k - l * L;
      r = M[v * L2 + k] + face_add2<P>(CT<P>::Bend(), H + v * 4 * L, l, m, cw);
    }
    a.R[((size_t)v * ne + e) * a.NM + k] = r;
  }
}

// nodal [v][e][QM] (the element's Q^2 first) <-> modal [v][e][NM] (L^2 first)
template <int P>
__global__ void __launch_bounds__(64) k_vforward2(VarpArgs a) {
  constexpr int L = P + 1, Q = P + 2, L2 = L * L, Q2 = Q * Q, NT = 64;
  __shared__ double A[Q2], X[Q * L], O[L2];
  const int e = a.elems[blockIdx.x % a.nel], v = blockIdx.x / a.nel, tid = threadIdx.x;
  const double *src = a.X + ((size_t)v * a.ne + e) * a.QM;
  for (int i = tid; i < Q2; i += NT) A[i] = src[i];
  __syncthreads();
  contract<1, Q, Q, 2, L, NT>(CT<P>::BW(), A, X, tid);
  __syncthreads();
  contract<1, Q, L, 1, L, NT>(CT<P>::BW(), X, O, tid);
  __syncthreads();
  double *dst = a.Y + ((size_t)v * a.ne + e) * a.NM;
  for (int i = tid; i < a.NM; i += NT) dst[i] = i < L2 ? O[i] : 0.0;
}

template <int P>
__global__ void __launch_bounds__(64) k_vbackward2(VarpArgs a) {
  constexpr int L = P + 1, Q = P + 2, L2 = L * L, Q2 = Q * Q, NT = 64;
  __shared__ double A[L2], X[Q * L], O[Q2];
  const int e = a.elems[blockIdx.x % a.nel], v = blockIdx.x / a.nel, tid = threadIdx.x;
  const double *src = a.X + ((size_t)v * a.ne + e) * a.NM;
  for (int i = tid; i < L2; i += NT) A[i] = src[i];
  __syncthreads();
  contract<1, L, L, 1, Q, NT>(CT<P>::BT(), A, X, tid);
  __syncthreads();
  contract<1, Q, L, 2, Q, NT>(CT<P>::BT(), X, O, tid);
  __syncthreads();
  double *dst = a.Y + ((size_t)v * a.ne + e) * a.QM;
  for (int i = tid; i < a.QM; i += NT) dst[i] = i < Q2 ? O[i] : 0.0;
}

template <int P>
static hipError_t launch_varp2_p(const VarpArgs &a) {
  switch (a.op) {
    case OP_TRACES:
      hipLaunchKernelGGL(k_vtraces2<P>, dim3(a.nel), dim3(64), 0, a.stream, a);
      break;
    case OP_RHS:
      hipLaunchKernelGGL(k_vrhs2<P>, dim3(a.nel), dim3(64), 0, a.stream, a);
      break;
    case OP_FORWARD:
      hipLaunchKernelGGL(k_vforward2<P>, dim3(a.nel * 4), dim3(64), 0, a.stream, a);
      break;
    case OP_BACKWARD:
      hipLaunchKernelGGL(k_vbackward2<P>, dim3(a.nel * 4), dim3(64), 0, a.stream, a);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_varp2(const VarpArgs &a) {
  if (a.nel < 1 || !a.elems || a.QFM != kMaxOrder + 1) return hipErrorInvalidValue;
  switch (a.p) {
#define DGX_CASE(PP) case PP: return launch_varp2_p<PP>(a);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9) DGX_CASE(10)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t upload_tables_v2d(const double *h) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), h, sizeof(double) * kTabTotal);
}

}  // namespace dgx
