// Quad (2D) kernels of the DG residual; same structure as the hex kernels:
// one workgroup per element, LDS tensors, register-blocked pencil contractions.
//   face traces      Quad::fill_sides                 src/Quad.cpp:845-869
//   volume term      Euler::convection                src/Euler.cpp:296-390
//   numerical flux   Euler::convective_flux (LF, Roe) src/Euler.cpp:399-757
//   lift + project   Quad::lift, Element::project     src/Quad.cpp:1971-2038, src/Element.hpp:427-437
//   BR1 viscous      Navier_Stokes::diffusion         src/Navier_Stokes.cpp:113-258
//   shallow water    Shallow::convection, LF flux     src/Shallow.cpp:251-525
//   linear system    LinEuler_quasi::convection       src/LinEuler_quasi.cpp:151-495
//                    and its discrete adjoint
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <vector>

#include "dgx_kernels.hpp"

namespace dgx {

// Float copies of the basis tables for the mixed quad kernels, which hold
// their LDS tensors and run their contractions in FP32 while the state, face
// traces, LSRK registers and pointwise physics stay FP64; filled beside c_tab
// at upload.  TB<R, P> gives the tables of arithmetic type R.
static __constant__ float c_tabf[kTabTotal];
template <typename R, int P> struct TB;
template <int P> struct TB<double, P> : CT<P> {};
template <int P> struct TB<float, P> {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1, b = tab_offset(P);
  __device__ static __forceinline__ const float *BT() { return c_tabf + b; }
  __device__ static __forceinline__ const float *BW() { return c_tabf + b + Q * L; }
  __device__ static __forceinline__ const float *QD() { return c_tabf + b + 2 * Q * L; }
  __device__ static __forceinline__ const float *Bg() { return c_tabf + b + 3 * Q * L; }
  __device__ static __forceinline__ const float *Bend() { return c_tabf + b + 3 * Q * L + QF * L; }
  __device__ static __forceinline__ const float *PW() { return c_tabf + b + 3 * Q * L + QF * L + 2 * L; }
};

template <int P> struct H2 {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1;
  static constexpr int L2 = L * L, Q2 = Q * Q;
  static constexpr int NT = 64;
};

template <int P>
__device__ __forceinline__ void backward2(const Tab &tb, const double *m, double *s1, double *out,
                                          int tid) {
  constexpr int L = H2<P>::L, Q = H2<P>::Q, NT = H2<P>::NT;
  contract<1, L, L, 1, Q, NT>(CT<P>::BT(), m, s1, tid);
  __syncthreads();
  contract<1, Q, L, 2, Q, NT>(CT<P>::BT(), s1, out, tid);
  __syncthreads();
}

// Gauss-point traces of the NF fields of element e on its four edges, all
// fields per contraction (one load phase, four contraction phases).
// LDS of the traces: A [NF][L2], Kb, Ka [NF][2][L], O [NF][2][QF], carved from `sm`.
template <int P, int NF>
constexpr int traces2_lds() { return NF * (P + 1) * (P + 1) + 4 * NF * (P + 1) + 2 * NF * (P + 1); }

template <int P, int NF, typename R>
__device__ __forceinline__ void traces2_core(const double *__restrict__ U, double *__restrict__ T, int ne, int e,
                                             R *sm) {
  using S = H2<P>;
  constexpr int L = S::L, QF = S::QF, L2 = S::L2, NT = S::NT;
  R *A = sm, *Kb = A + NF * L2, *Ka = Kb + NF * 2 * L, *O = Ka + NF * 2 * L;
  const int tid = threadIdx.x;
  double *Te = T + (size_t)e * 4 * NF * QF;
  for (int i = tid; i < NF * L2; i += NT) {
    const int v = i / L2;
    A[i] = (R)U[((size_t)v * ne + e) * L2 + (i - v * L2)];
  }
  __syncthreads();
  contract<NF, L, L, 2, 2, NT>(TB<R, P>::Bend(), A, Kb, tid);   // [NF][L][2]
  contract<NF, L, L, 1, 2, NT>(TB<R, P>::Bend(), A, Ka, tid);   // [NF][2][L]
  __syncthreads();
  contract<NF, L, 2, 1, QF, NT>(TB<R, P>::Bg(), Kb, O, tid);    // [NF][QF][2]: edges 0 / 2
  __syncthreads();
  for (int i = tid; i < NF * 2 * QF; i += NT) {
    const int v = i / (2 * QF), r = i - v * 2 * QF, a = r >> 1, end = r & 1;
    Te[((end ? 2 : 0) * NF + v) * QF + a] = O[i];
  }
  __syncthreads();
  contract<NF, 2, L, 2, QF, NT>(TB<R, P>::Bg(), Ka, O, tid);    // [NF][2][QF]: edges 3 / 1
  __syncthreads();
  for (int i = tid; i < NF * 2 * QF; i += NT) {
    const int v = i / (2 * QF), r = i - v * 2 * QF, end = r / QF, b = r - end * QF;
    Te[((end ? 1 : 3) * NF + v) * QF + b] = O[i];
  }
  __syncthreads();
}

template <int P, int NF, typename R = double>
__device__ __forceinline__ void traces2_elem(const double *__restrict__ U, double *__restrict__ T, int ne, int e) {
  __shared__ R sm[traces2_lds<P, NF>()];
  traces2_core<P, NF, R>(U, T, ne, e, sm);
}

template <int P, typename R = double>
__global__ void __launch_bounds__(64)
k_traces2(const double *__restrict__ U, double *__restrict__ T, Tab tb, int ne, int nf) {
  if (nf == 3) traces2_elem<P, 3, R>(U, T, ne, blockIdx.x);
  else traces2_elem<P, 4, R>(U, T, ne, blockIdx.x);
}

// Residual of element e (Euler, or with NS the BR1 viscous terms from V, TV);
// out(o, r) receives the modal residual r at offset o = (v*ne + e)*L2 + k.
// Every contraction runs on all fields of the element at once, so a pass
// keeps NF x Q pencils busy instead of Q.
template <int P, int NF>
struct H2F {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1, L2 = L * L, Q2 = Q * Q;
  // LDS (doubles): M [NF][L2] (modal, then the modal residual), X1, X2 [NF][Q][L],
  // N1, N2 [NF][Q2] (nodal state, then the contravariant fluxes), gF [NF][4][QF], H [NF][4][L]
  static constexpr int NM = NF * L2, NX = NF * Q * L, NN = NF * Q2;
};

template <int P, int NF, typename R = double>
__device__ __forceinline__ R *rhs2_smem() {
  using F = H2F<P, NF>;
  __shared__ R sm[F::NM + 2 * F::NX + 2 * F::NN + NF * 4 * F::QF + NF * 4 * F::L];
  return sm;
}

// Element state modal -> nodal for all NF fields; each thread keeps its nodes
// k = tid + r*NT in registers (un[r][v]).
template <int P, int NF, int NPT, typename R = double>
__device__ __forceinline__ void load_nodal2(const double *__restrict__ U, int ne, int e, R *M, R *X,
                                            R *N, double (&un)[NPT][NF], int tid) {
  using F = H2F<P, NF>;
  constexpr int L = F::L, Q = F::Q, L2 = F::L2, Q2 = F::Q2, NT = 64;
  for (int i = tid; i < NF * L2; i += NT) {
    const int v = i / L2;
    M[i] = (R)U[((size_t)v * ne + e) * L2 + (i - v * L2)];
  }
  __syncthreads();
  contract<NF, L, L, 1, Q, NT>(TB<R, P>::BT(), M, X, tid);   // [NF][Q][L]
  __syncthreads();
  contract<NF, Q, L, 2, Q, NT>(TB<R, P>::BT(), X, N, tid);   // [NF][Q][Q]
  __syncthreads();
#pragma unroll
  for (int r = 0; r < NPT; ++r) {
    const int k = tid + r * NT;
#pragma unroll
    for (int v = 0; v < NF; ++v) un[r][v] = k < Q2 ? N[v * Q2 + k] : 1.0;
  }
}

// Projected strong-form residual of all NF fields from the contravariant
// fluxes in N1 (r) and N2 (s) and the face terms gF; result in M [NF][L2].
template <int P, int NF, typename R = double>
__device__ __forceinline__ void divergence2(R *M, R *X1, R *X2, R *N1, R *N2, const R *gF, R *H, int tid) {
  using F = H2F<P, NF>;
  constexpr int L = F::L, Q = F::Q, QF = F::QF, NT = 64;
  contract<NF, Q, Q, 2, L, NT>(TB<R, P>::BW(), N1, X1, tid);
  contract<NF, Q, Q, 2, L, NT>(TB<R, P>::QD(), N2, X2, tid);
  contract<NF, 4, QF, 2, L, NT>(TB<R, P>::PW(), gF, H, tid);
  __syncthreads();
  contract<NF, Q, L, 1, L, NT>(TB<R, P>::QD(), X1, M, tid);
  __syncthreads();
  contract<NF, Q, L, 1, L, NT, true>(TB<R, P>::BW(), X2, M, tid);
  __syncthreads();
}

template <int P, bool NS, typename R = double, class Out>
__device__ __forceinline__ void rhs2_elem(const double *__restrict__ U, const double *__restrict__ T,
                                          const double *__restrict__ V, const double *__restrict__ TV,
                                          const Geo &geo, const Phys &ph, int ne, int e, int *__restrict__ flag,
                                          Out out) {
  using F = H2F<P, 4>;
  constexpr int L = F::L, QF = F::QF, L2 = F::L2, Q2 = F::Q2, NT = 64;
  constexpr int NPT = (Q2 + NT - 1) / NT;
  R *M = rhs2_smem<P, 4, R>(), *X1 = M + F::NM, *X2 = X1 + F::NX, *N1 = X2 + F::NX, *N2 = N1 + F::NN;
  R *gF = N2 + F::NN, *H = gF + 16 * QF;
  const int tid = threadIdx.x;
  const double gm1 = ph.gamma - 1.0;
  double un[NPT][4];
  load_nodal2<P, 4, NPT, R>(U, ne, e, M, X1, N1, un, tid);
  bool bad = false;
  for (int i = tid; i < 4 * QF; i += NT) {
    const int s = i / QF, a = i - s * QF;
    const int es = e * 4 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int an = nb >= 0 ? geo.perm[geo.permid[es] * QF + a] : a;
    const double nx = geo.fnorm[3 * es], ny = geo.fnorm[3 * es + 1];
    const double jr = geo.fjr[es];
    double UL[4], UR[4], FL[4], FR[4], Fh[4], lL, lR;
    const double *tl = T + ((size_t)e * 4 + s) * 4 * QF;
#pragma unroll
    for (int v = 0; v < 4; ++v) UL[v] = tl[v * QF + a];
    if (nb >= 0) {
      const double *tr = T + ((size_t)nb * 4 + ns) * 4 * QF;
#pragma unroll
      for (int v = 0; v < 4; ++v) UR[v] = tr[v * QF + an];
    } else {   // boundary side of type -1 - nb (Navier-Stokes meshes have none: rejected at creation)
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
    for (int v = 0; v < 4; ++v) {
      double g = jr * (Fh[v] - FL[v]);
      if (NS && v > 0 && nb >= 0) {
        const double *vl = TV + ((size_t)e * 4 + s) * 3 * QF;
        const double *vr = TV + ((size_t)nb * 4 + ns) * 3 * QF;
        g += 0.5 * jr * (vl[(v - 1) * QF + a] + vr[(v - 1) * QF + an]);
      }
      gF[(v * 4 + s) * QF + a] = (R)g;
    }
  }
  if (bad) atomicOr(flag, 1);
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
      N1[v * Q2 + k] = (R)(rx * fx[v] + ry * fy[v]);
      N2[v * Q2 + k] = (R)(sx * fx[v] + sy * fy[v]);
    }
  }
  __syncthreads();
  divergence2<P, 4, R>(M, X1, X2, N1, N2, gF, H, tid);
  const double cw[4] = {1.0, 1.0, 1.0, 1.0};
  for (int i = tid; i < 4 * L2; i += NT) {
    const int v = i / L2, k = i - v * L2, l = k / L, m = k - l * L;
    double r = M[i] + face_add2<P>(TB<R, P>::Bend(), H + v * 4 * L, l, m, cw);
    if (NS && v > 0) r -= V[((size_t)e * 3 + (v - 1)) * L2 + k];
    out(((size_t)v * ne + e) * L2 + k, r);
  }
  __syncthreads();
}

template <int P, bool NS, typename RT = double>
__global__ void __launch_bounds__(64)
k_rhs2(const double *__restrict__ U, const double *__restrict__ T, const double *__restrict__ V,
       const double *__restrict__ TV, double *__restrict__ R, Tab tb, Geo geo, Phys ph, int ne,
       int *__restrict__ flag) {
  rhs2_elem<P, NS, RT>(U, T, V, TV, geo, ph, ne, blockIdx.x, flag, [=](size_t o, double r) { R[o] = r; });
}

// One LSRK14 stage of an affine 2D domain in one kernel after the traces (and
// the viscous terms): the residual of element e fused with the register
// update Ut = A Ut - dt R, U += B Ut in its store, bitwise the same as the
// residual followed by the axpby and axpy.  An element reads its own modes
// only (its neighbours' through the traces), before its own store.
// With Tn, the element then writes the face traces of its new state into Tn
// (the next stage's traces, in the residual's LDS): it reads back its own
// modes, which its workgroup has just stored.
template <int P, bool NS, typename R = double>
__global__ void __launch_bounds__(64)
k_step2(double *U, const double *__restrict__ T, const double *__restrict__ V, const double *__restrict__ TV,
        double *__restrict__ Ut, Geo geo, Phys ph, int ne, int *__restrict__ flag, double A, double B, double mdt,
        double *__restrict__ Tn) {
  rhs2_elem<P, NS, R>(U, T, V, TV, geo, ph, ne, blockIdx.x, flag, [=](size_t o, double r) {
    const double ut = fma(mdt, r, A * Ut[o]);
    Ut[o] = ut;
    U[o] = fma(B, ut, U[o]);
  });
  if (Tn) traces2_core<P, 4, R>(U, Tn, ne, blockIdx.x, rhs2_smem<P, 4, R>());
}

// Shallow water (Shallow::time_derivative, src/Shallow.cpp:255-336): flux
// divergence at the GLL points, Lax-Friedrichs face flux, lift, projection.
// The balance term (src/Shallow.cpp:251-266): bal 1 stores this residual in
// Bal and returns R - Bal = 0, bal 2 returns R - Bal.
template <int P, typename R = double, class Out>
__device__ __forceinline__ void shallow2_elem(const double *__restrict__ U, const double *__restrict__ T,
                                              double *__restrict__ Bal, int bal, const Geo &geo, const Phys &ph,
                                              int ne, int e, int *__restrict__ flag, Out out) {
  using F = H2F<P, 3>;
  constexpr int L = F::L, QF = F::QF, L2 = F::L2, Q2 = F::Q2, NT = 64;
  constexpr int NPT = (Q2 + NT - 1) / NT;
  R *M = rhs2_smem<P, 3, R>(), *X1 = M + F::NM, *X2 = X1 + F::NX, *N1 = X2 + F::NX, *N2 = N1 + F::NN;
  R *gF = N2 + F::NN, *H = gF + 12 * QF;
  const int tid = threadIdx.x;
  double un[NPT][3];
  load_nodal2<P, 3, NPT, R>(U, ne, e, M, X1, N1, un, tid);
  bool bad = false;
  for (int i = tid; i < 4 * QF; i += NT) {
    const int s = i / QF, a = i - s * QF;
    const int es = e * 4 + s;
    const int nb = geo.nbr[es];
    const double nx = geo.fnorm[3 * es], ny = geo.fnorm[3 * es + 1];
    const double jr = geo.fjr[es];
    double UL[3], UR[3], FL[3], FR[3], lL, lR;
    const double *tl = T + ((size_t)e * 4 + s) * 3 * QF;
#pragma unroll
    for (int v = 0; v < 3; ++v) UL[v] = tl[v * QF + a];
    if (nb < 0) {
      ghost_shallow2(-1 - nb, UL, nx, ny, ph.bcU, UR);
    } else {
      const int an = geo.perm[geo.permid[es] * QF + a];
      const double *tr = T + ((size_t)nb * 4 + geo.nside[es]) * 3 * QF;
#pragma unroll
      for (int v = 0; v < 3; ++v) UR[v] = tr[v * QF + an];
    }
    bad |= shallow_normal_flux2(UL, nx, ny, FL, lL);
    bad |= shallow_normal_flux2(UR, nx, ny, FR, lR);
    const double lmax = fmax(lL, lR);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const double Fh = 0.5 * (FL[v] + FR[v] + lmax * (UL[v] - UR[v]));
      gF[(v * 4 + s) * QF + a] = (R)(jr * (Fh - FL[v]));
    }
  }
  if (bad) atomicOr(flag, 1);
  const double *mt = geo.metric + 9 * (size_t)e;
  const double rx = mt[0], ry = mt[1], sx = mt[3], sy = mt[4];
#pragma unroll
  for (int rr = 0; rr < NPT; ++rr) {
    const int k = tid + rr * NT;
    if (k >= Q2) continue;
    const double phi = un[rr][0], mx = un[rr][1], my = un[rr][2];
    const double u = mx / phi, vv = my / phi, p = 0.5 * (phi * phi);
    const double fx[3] = {mx, mx * u + p, my * u}, fy[3] = {my, mx * vv, my * vv + p};
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      N1[v * Q2 + k] = (R)(rx * fx[v] + ry * fy[v]);
      N2[v * Q2 + k] = (R)(sx * fx[v] + sy * fy[v]);
    }
  }
  __syncthreads();
  divergence2<P, 3, R>(M, X1, X2, N1, N2, gF, H, tid);
  const double cw[4] = {1.0, 1.0, 1.0, 1.0};
  for (int i = tid; i < 3 * L2; i += NT) {
    const int v = i / L2, k = i - v * L2, l = k / L, m = k - l * L;
    const double r = M[i] + face_add2<P>(TB<R, P>::Bend(), H + v * 4 * L, l, m, cw);
    const size_t o = ((size_t)v * ne + e) * L2 + k;
    if (bal == 0) {
      out(o, r);
    } else {
      double b;
      if (bal == 1) { b = r; Bal[o] = r; }
      else b = Bal[o];
      out(o, -1.0 * b + r);
    }
  }
  __syncthreads();
}

template <int P, typename RT = double>
__global__ void __launch_bounds__(64)
k_shallow2(const double *__restrict__ U, const double *__restrict__ T, double *__restrict__ R,
           double *__restrict__ Bal, int bal, Tab tb, Geo geo, Phys ph, int ne, int *__restrict__ flag) {
  shallow2_elem<P, RT>(U, T, Bal, bal, geo, ph, ne, blockIdx.x, flag, [=](size_t o, double r) { R[o] = r; });
}

template <int P, typename R = double>
__global__ void __launch_bounds__(64)
k_shallow_step2(double *U, const double *__restrict__ T, double *__restrict__ Ut, double *__restrict__ Bal, int bal,
                Geo geo, Phys ph, int ne, int *__restrict__ flag, double A, double B, double mdt,
                double *__restrict__ Tn) {
  shallow2_elem<P, R>(U, T, Bal, bal, geo, ph, ne, blockIdx.x, flag, [=](size_t o, double r) {
    const double ut = fma(mdt, r, A * Ut[o]);
    Ut[o] = ut;
    U[o] = fma(B, ut, U[o]);
  });
  if (Tn) traces2_core<P, 3, R>(U, Tn, ne, blockIdx.x, rhs2_smem<P, 3, R>());
}

// lifted, projected gradients of (u, v, e): G[e][2*prim + dir][L2]
template <int P>
__global__ void __launch_bounds__(64)
k_grad2(const double *__restrict__ U, const double *__restrict__ T, double *__restrict__ G, Tab tb,
        Geo geo, int ne) {
  using S = H2<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, Q2 = S::Q2, NT = S::NT;
  __shared__ double Un[4 * Q2], Jq[12 * QF], Fb[Q2], Xa[Q * L], Xc[Q * L], D2[2 * L2], H[4 * L];
  const int e = blockIdx.x, tid = threadIdx.x;
  for (int v = 0; v < 4; ++v) {
    copy_block<NT>(U + ((size_t)v * ne + e) * L2, D2, L2, tid);
    __syncthreads();
    backward2<P>(tb, D2, Xa, Un + v * Q2, tid);
  }
  for (int i = tid; i < 4 * QF; i += NT) {
    const int s = i / QF, a = i - s * QF;
    const int es = e * 4 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int an = geo.perm[geo.permid[es] * QF + a];
    const double jr = geo.fjr[es];
    const double *tl = T + ((size_t)e * 4 + s) * 4 * QF;
    const double *tr = T + ((size_t)nb * 4 + ns) * 4 * QF;
    double pl[3], pr[3];
    {
      const double r = tl[a], u = tl[QF + a] / r, v = tl[2 * QF + a] / r;
      double en = tl[3 * QF + a] / r;
      double wk = u * u; wk = v * v + wk;
      pl[0] = u; pl[1] = v; pl[2] = -0.5 * wk + en;
    }
    {
      const double r = tr[an], u = tr[QF + an] / r, v = tr[2 * QF + an] / r;
      double en = tr[3 * QF + an] / r;
      double wk = u * u; wk = v * v + wk;
      pr[0] = u; pr[1] = v; pr[2] = -0.5 * wk + en;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) Jq[(q * 4 + s) * QF + a] = jr * (0.5 * (pl[q] + pr[q]) - pl[q]);
  }
  __syncthreads();
  const double *mt = geo.metric + 9 * (size_t)e;
  for (int q = 0; q < 3; ++q) {
    for (int k = tid; k < Q2; k += NT) {
      const double r = Un[k], u = Un[Q2 + k] / r, v = Un[2 * Q2 + k] / r;
      double val;
      if (q == 0) val = u;
      else if (q == 1) val = v;
      else {
        double en = Un[3 * Q2 + k] / r;
        double wk = u * u; wk = v * v + wk;
        val = -0.5 * wk + en;
      }
      Fb[k] = val;
    }
    __syncthreads();
    contract<1, Q, Q, 2, L, NT>(CT<P>::BW(), Fb, Xa, tid);
    contract<1, Q, Q, 2, L, NT>(CT<P>::QD(), Fb, Xc, tid);
    contract<1, 4, QF, 2, L, NT>(CT<P>::PW(), Jq + q * 4 * QF, H, tid);
    __syncthreads();
    contract<1, Q, L, 1, L, NT>(CT<P>::QD(), Xa, D2, tid);
    contract<1, Q, L, 1, L, NT>(CT<P>::BW(), Xc, D2 + L2, tid);
    __syncthreads();
    for (int k = tid; k < L2; k += NT) {
      const int l = k / L, m = k - l * L;
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        double c[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) c[s] = geo.fnorm[3 * (e * 4 + s) + d];
        G[((size_t)e * 6 + 2 * q + d) * L2 + k] =
            mt[d] * D2[k] + mt[3 + d] * D2[L2 + k] + face_add2<P>(CT<P>::Bend(), H, l, m, c);
      }
    }
    __syncthreads();
  }
}

template <int P>
__global__ void __launch_bounds__(64)
k_visc2(const double *__restrict__ U, const double *__restrict__ G, double *__restrict__ V,
        double *__restrict__ TV, Tab tb, Geo geo, Phys ph, int ne) {
  using S = H2<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, Q2 = S::Q2, NT = S::NT;
  __shared__ double Un[4 * Q2], Gn[6 * Q2], M0[L2], S1[Q * L], Lay[12 * Q], Wf[12 * QF], X1[Q * L],
      X2[Q * L], Vo[L2];
  const int e = blockIdx.x, tid = threadIdx.x;
  for (int v = 0; v < 10; ++v) {
    const double *src = v < 4 ? U + ((size_t)v * ne + e) * L2 : G + ((size_t)e * 6 + (v - 4)) * L2;
    double *dst = v < 4 ? Un + v * Q2 : Gn + (v - 4) * Q2;
    copy_block<NT>(src, M0, L2, tid);
    __syncthreads();
    backward2<P>(tb, M0, S1, dst, tid);
  }
  const double mu = ph.vis, kap = ph.vis * ph.gamma / ph.Pr;
  auto vflux = [&](int k, double out[3][2]) {
    const double r = Un[k], u = Un[Q2 + k] / r, v = Un[2 * Q2 + k] / r;
    const double ux = Gn[k], uy = Gn[Q2 + k], vx = Gn[2 * Q2 + k], vy = Gn[3 * Q2 + k];
    const double ex = Gn[4 * Q2 + k], ey = Gn[5 * Q2 + k];
    const double t12 = mu * (uy + vx);
    const double t11 = mu * ((2.0 / 3.0) * (2.0 * ux - vy));
    const double t22 = mu * ((2.0 / 3.0) * (2.0 * vy - ux));
    out[0][0] = t11; out[0][1] = t12;
    out[1][0] = t12; out[1][1] = t22;
    out[2][0] = t11 * u + t12 * v + kap * ex;
    out[2][1] = t12 * u + t22 * v + kap * ey;
  };
  for (int i = tid; i < 4 * Q; i += NT) {
    const int s = i / Q, j = i - s * Q;
    int k;
    switch (s) {
      case 0: k = j * Q; break;
      case 1: k = (Q - 1) * Q + j; break;
      case 2: k = j * Q + Q - 1; break;
      default: k = j; break;
    }
    const double *nn = geo.fnorm + 3 * (e * 4 + s);
    double f[3][2];
    vflux(k, f);
#pragma unroll
    for (int q = 0; q < 3; ++q) Lay[(s * 3 + q) * Q + j] = f[q][0] * nn[0] + f[q][1] * nn[1];
  }
  __syncthreads();
  contract<1, 12, Q, 2, QF, NT>(CT<P>::Ig(), Lay, Wf, tid);
  __syncthreads();
  copy_block<NT>(Wf, TV + (size_t)e * 12 * QF, 12 * QF, tid);
  const double *mt = geo.metric + 9 * (size_t)e;
  for (int k = tid; k < Q2; k += NT) {
    double f[3][2];
    vflux(k, f);
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      Gn[(2 * q) * Q2 + k] = mt[0] * f[q][0] + mt[1] * f[q][1];
      Gn[(2 * q + 1) * Q2 + k] = mt[3] * f[q][0] + mt[4] * f[q][1];
    }
  }
  __syncthreads();
  for (int q = 0; q < 3; ++q) {
    contract<1, Q, Q, 2, L, NT>(CT<P>::BW(), Gn + 2 * q * Q2, X1, tid);
    contract<1, Q, Q, 2, L, NT>(CT<P>::QD(), Gn + (2 * q + 1) * Q2, X2, tid);
    __syncthreads();
    contract<1, Q, L, 1, L, NT>(CT<P>::QD(), X1, Vo, tid);
    __syncthreads();
    contract<1, Q, L, 1, L, NT, true>(CT<P>::BW(), X2, Vo, tid);
    __syncthreads();
    copy_block<NT>(Vo, V + ((size_t)e * 3 + q) * L2, L2, tid);
    __syncthreads();
  }
}

template <int P>
__global__ void __launch_bounds__(64)
k_forward2(const double *__restrict__ Fn, double *__restrict__ Fm, Tab tb, int ne) {
  using S = H2<P>;
  constexpr int L = S::L, Q = S::Q, L2 = S::L2, Q2 = S::Q2, NT = S::NT;
  __shared__ double A[Q2], X[Q * L], O[L2];
  const int e = blockIdx.x % ne, v = blockIdx.x / ne, tid = threadIdx.x;
  copy_block<NT>(Fn + ((size_t)v * ne + e) * Q2, A, Q2, tid);
  __syncthreads();
  contract<1, Q, Q, 2, L, NT>(CT<P>::BW(), A, X, tid);
  __syncthreads();
  contract<1, Q, L, 1, L, NT>(CT<P>::BW(), X, O, tid);
  __syncthreads();
  copy_block<NT>(O, Fm + ((size_t)v * ne + e) * L2, L2, tid);
}

template <int P>
__global__ void __launch_bounds__(64)
k_backward2(const double *__restrict__ Fm, double *__restrict__ Fn, Tab tb, int ne) {
  using S = H2<P>;
  constexpr int L = S::L, Q = S::Q, L2 = S::L2, Q2 = S::Q2, NT = S::NT;
  __shared__ double A[L2], X[Q * L], O[Q2];
  const int e = blockIdx.x % ne, v = blockIdx.x / ne, tid = threadIdx.x;
  copy_block<NT>(Fm + ((size_t)v * ne + e) * L2, A, L2, tid);
  __syncthreads();
  backward2<P>(tb, A, X, O, tid);
  copy_block<NT>(O, Fn + ((size_t)v * ne + e) * Q2, Q2, tid);
}

// Linear system residual, Lax-Friedrichs flux 1/2 (An UL + An UR) + 1/2 lam
// (UL - UR), lam = |u0.n| + c0 (LinEuler_quasi::convection / convective_flux,
// src/LinEuler_quasi.cpp:161-495, for a uniform mean flow), fused with its
// LSRK14 stage.  ADJ: the same residual of the adjoint system applied to
// X / J_e and multiplied by J_e is the transpose of the forward residual in
// modal coordinates (exact quadrature throughout), and the kernel performs one
// reversed stage: Ubar += -dt L^T X, dJ/dtheta_e += 2 g dt X[E][e][0],
// Ubar_t = ca X.
template <int P, bool ADJ, typename R = double>
__global__ void __launch_bounds__(64)
k_lin2(LinArgs a) {
  using F = H2F<P, 4>;
  constexpr int L = F::L, QF = F::QF, L2 = F::L2, Q2 = F::Q2, NT = 64;
  constexpr int NPT = (Q2 + NT - 1) / NT;
  R *M = rhs2_smem<P, 4, R>(), *X1 = M + F::NM, *X2 = X1 + F::NX, *N1 = X2 + F::NX, *N2 = N1 + F::NN;
  R *gF = N2 + F::NN, *H = gF + 16 * QF;
  const int e = blockIdx.x, tid = threadIdx.x, ne = a.ne;
  const Geo &geo = a.geo;
  const double *mt = geo.metric + 9 * (size_t)e;
  const double rx = mt[0], ry = mt[1], sx = mt[3], sy = mt[4];
  const double dm = rx * sy - ry * sx;   // 1 / J_e
  const double sc = ADJ ? dm : 1.0;
  double un[NPT][4];
  load_nodal2<P, 4, NPT, R>(a.X, ne, e, M, X1, N1, un, tid);
  for (int i = tid; i < 4 * QF; i += NT) {
    const int s = i / QF, q = i - s * QF;
    const int es = e * 4 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int qn = geo.perm[geo.permid[es] * QF + q];
    const double nx = geo.fnorm[3 * es], ny = geo.fnorm[3 * es + 1];
    const double jr = geo.fjr[es];
    double scn = 1.0;
    if (ADJ) {
      const double *mn = geo.metric + 9 * (size_t)nb;
      scn = mn[0] * mn[4] - mn[1] * mn[3];
    }
    const double *tl = a.T + ((size_t)e * 4 + s) * 4 * QF;
    const double *tr = a.T + ((size_t)nb * 4 + ns) * 4 * QF;
    double UL[4], UR[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) { UL[v] = sc * tl[v * QF + q]; UR[v] = scn * tr[v * QF + qn]; }
    const double lam = fabs(a.u0 * nx + a.v0 * ny) + a.c0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double fl = 0.0, fr = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double An = nx * a.Ax[4 * v + j] + ny * a.Ay[4 * v + j];
        fl = fma(An, UL[j], fl);
        fr = fma(An, UR[j], fr);
      }
      const double fh = 0.5 * (fl + fr + lam * (UL[v] - UR[v]));
      gF[(v * 4 + s) * QF + q] = (R)(jr * (fh - fl));
    }
  }
#pragma unroll
  for (int rr = 0; rr < NPT; ++rr) {
    const int k = tid + rr * NT;
    if (k >= Q2) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double fx = 0.0, fy = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const double u = sc * un[rr][j];
        fx = fma(a.Ax[4 * v + j], u, fx);
        fy = fma(a.Ay[4 * v + j], u, fy);
      }
      N1[v * Q2 + k] = (R)(rx * fx + ry * fy);
      N2[v * Q2 + k] = (R)(sx * fx + sy * fy);
    }
  }
  __syncthreads();
  divergence2<P, 4, R>(M, X1, X2, N1, N2, gF, H, tid);
  const double cw[4] = {1.0, 1.0, 1.0, 1.0};
  for (int i = tid; i < 4 * L2; i += NT) {
    const int v = i / L2, k = i - v * L2, l = k / L, m = k - l * L;
    double r = M[i] + face_add2<P>(TB<R, P>::Bend(), H + v * 4 * L, l, m, cw);
    const size_t gi = ((size_t)v * ne + e) * L2 + k;
    if (!ADJ) {
      if (v == 3 && k == 0) r -= 2.0 * a.theta[e] * a.g;   // source: mode 0 of the constant theta_e
      const double ut = a.ca * a.Yt[gi] - a.dt * r;
      a.Yt[gi] = ut;
      a.Y[gi] = fma(a.cb, ut, a.Y[gi]);
    } else {
      const double x = a.X[gi];
      a.Y[gi] = fma(-a.dt, r / dm, a.Y[gi]);
      if (v == 3 && k == 0) a.gtheta[e] += 2.0 * a.g * a.dt * x;
      a.Yt[gi] = a.ca * x;
    }
  }
}

// Observations (one workgroup per receiver) and the objective's adjoint (one
// workgroup, receivers in order, so that receivers sharing an element never
// race).
__global__ void __launch_bounds__(64)
k_observe(const double *__restrict__ U, const int *__restrict__ rec_e, const double *__restrict__ rec_w,
          const double *__restrict__ cobs, int ne, int L2, double *__restrict__ y) {
  __shared__ double red[64];
  const int r = blockIdx.x, tid = threadIdx.x, e = rec_e[r];
  double s = 0.0;
  for (int i = tid; i < 4 * L2; i += 64) {
    const int f = i / L2, m = i - f * L2;
    s = fma(cobs[f] * rec_w[(size_t)r * L2 + m], U[((size_t)f * ne + e) * L2 + m], s);
  }
  red[tid] = s;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) y[r] = red[0];
}

__global__ void __launch_bounds__(64)
k_observe_adjoint(const double *__restrict__ U, double *__restrict__ Ubar, const int *__restrict__ rec_e,
                  const double *__restrict__ rec_w, const double *__restrict__ cobs, int nrec, int ne, int L2,
                  const double *__restrict__ d, double dt, double *__restrict__ J) {
  __shared__ double red[64];
  const int tid = threadIdx.x;
  double jsum = 0.0;
  for (int r = 0; r < nrec; ++r) {
    const int e = rec_e[r];
    double s = 0.0;
    for (int i = tid; i < 4 * L2; i += 64) {
      const int f = i / L2, m = i - f * L2;
      s = fma(cobs[f] * rec_w[(size_t)r * L2 + m], U[((size_t)f * ne + e) * L2 + m], s);
    }
    red[tid] = s;
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    const double res = red[0] - d[r];
    __syncthreads();
    jsum += res * res;
    for (int i = tid; i < 4 * L2; i += 64) {
      const int f = i / L2, m = i - f * L2;
      const size_t gi = ((size_t)f * ne + e) * L2 + m;
      Ubar[gi] = fma(dt * res * cobs[f], rec_w[(size_t)r * L2 + m], Ubar[gi]);
    }
    __syncthreads();
  }
  if (tid == 0) *J += 0.5 * dt * jsum;
}

// Batched serving (dgx_server with batching): one launch per stage runs every
// member request of a group (same order and physics).  A fixed grid of
// workgroups walks the flat list of (request, element) work items; members
// are in work-item order, b[r].off their first item.  The step kernel fuses
// the LSRK14 register update of the direct path (vField axpby, axpy:
// Ut = A Ut - dt R, U += B Ut) into the residual's store.
template <int P, int NF, typename R = double>
__global__ void __launch_bounds__(64)
k_batch_traces2(const Batch2 *__restrict__ b, int nreq, int nwork) {
  int r = 0;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    while (r + 1 < nreq && b[r + 1].off <= w) ++r;
    traces2_elem<P, NF, R>(b[r].U, b[r].T, b[r].ne, w - b[r].off);
  }
}

// odd: the stage reads the traces of T2 and writes the next ones into T
// (else T and T2); next: write the next stage's traces at all.
template <int P, bool SHALLOW, typename R = double>
__global__ void __launch_bounds__(64)
k_batch_step2(const Batch2 *__restrict__ b, int nreq, int nwork, double A, double B, int odd, int next) {
  constexpr int NF = SHALLOW ? 3 : 4;
  int r = 0;
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    while (r + 1 < nreq && b[r + 1].off <= w) ++r;
    const double *Tc = odd ? b[r].T2 : b[r].T;
    double *Tn = odd ? b[r].T : b[r].T2;
    double *U = b[r].U, *Ut = b[r].Ut;
    const double mdt = -b[r].dt;
    auto update = [=](size_t o, double rr) {
      const double ut = fma(mdt, rr, A * Ut[o]);   // as the direct path's axpby
      Ut[o] = ut;
      U[o] = fma(B, ut, U[o]);
    };
    const Geo geo = b[r].geo;
    const Phys ph = b[r].phys;
    if constexpr (SHALLOW)
      shallow2_elem<P, R>(U, Tc, b[r].Bal, b[r].bal, geo, ph, b[r].ne, w - b[r].off, b[r].flag, update);
    else
      rhs2_elem<P, false, R>(U, Tc, nullptr, nullptr, geo, ph, b[r].ne, w - b[r].off, b[r].flag, update);
    if (next) traces2_core<P, NF, R>(U, Tn, b[r].ne, w - b[r].off, rhs2_smem<P, NF, R>());
  }
}

struct LsrkTab {
  double A[14], B[14];
};

// A persistent slice of a batched group: nsteps LSRK14 steps of every member
// in one cooperative launch, whose workgroups are all resident at once.  The
// face traces of the slice's first state are computed first; then each stage
// is one pass over the members' elements, in which an element takes its
// residual from the current traces, updates its register and state in place
// and writes the traces of its new state into the other trace buffer,
// followed by one grid-wide barrier.  The buffers alternate T, T2, T, ... so
// that no element overwrites traces a neighbour may still be reading; after
// the even number of stages of a step the current traces are in T again.
// Each element's arithmetic is that of k_batch_traces2 and k_batch_step2, so
// the results are bitwise those of the per-stage launches.
template <int P, bool SHALLOW, typename R = double>
__global__ void __launch_bounds__(64)
k_batch_persist2(const Batch2 *__restrict__ b, int nreq, int nwork, int nsteps, LsrkTab c) {
  constexpr int NF = SHALLOW ? 3 : 4;
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  {
    int r = 0;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
      while (r + 1 < nreq && b[r + 1].off <= w) ++r;
      traces2_core<P, NF, R>(b[r].U, b[r].T, b[r].ne, w - b[r].off, rhs2_smem<P, NF, R>());
    }
  }
  grid.sync();
  for (int step = 0; step < nsteps; ++step)
    for (int st = 0; st < 14; ++st) {
      const double A = c.A[st], B = c.B[st];
      const bool odd = st & 1, last = step == nsteps - 1 && st == 13;
      int r = 0;
      for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        while (r + 1 < nreq && b[r + 1].off <= w) ++r;
        const int e = w - b[r].off, ne = b[r].ne;
        double *U = b[r].U, *Ut = b[r].Ut;
        const double *Tc = odd ? b[r].T2 : b[r].T;
        double *Tn = odd ? b[r].T : b[r].T2;
        const double mdt = -b[r].dt;
        auto update = [=](size_t o, double rr) {
          const double ut = fma(mdt, rr, A * Ut[o]);   // as the direct path's axpby
          Ut[o] = ut;
          U[o] = fma(B, ut, U[o]);
        };
        const Geo geo = b[r].geo;
        const Phys ph = b[r].phys;
        if constexpr (SHALLOW)
          shallow2_elem<P, R>(U, Tc, b[r].Bal, b[r].bal, geo, ph, ne, e, b[r].flag, update);
        else
          rhs2_elem<P, false, R>(U, Tc, nullptr, nullptr, geo, ph, ne, e, b[r].flag, update);
        if (!last) traces2_core<P, NF, R>(U, Tn, ne, e, rhs2_smem<P, NF, R>());   // its new state, read back
      }
      grid.sync();
    }
}

template <int P, bool SHALLOW, typename R>
static hipError_t launch_persist2(BatchLaunch &a) {
  const void *kern = reinterpret_cast<const void *>(&k_batch_persist2<P, SHALLOW, R>);
  int dev = 0, ncu = 0, per = 0;
  hipError_t e;
  if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
  if ((e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
  if ((e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, 64, 0)) != hipSuccess) return e;
  // every workgroup must be resident at once for the grid-wide barriers: keep
  // well inside the occupancy the runtime reports, counting a CU's LDS as
  // 64 KiB
  hipFuncAttributes fa;
  if ((e = hipFuncGetAttributes(&fa, kern)) != hipSuccess) return e;
  per = std::min(per, (int)(65536 / std::max<size_t>(fa.sharedSizeBytes, 1)));
  if (per < 1 || ncu < 1) return hipErrorInvalidConfiguration;
  const int grid = std::max(1, std::min(a.nwork, per * ncu));
  LsrkTab c;
  for (int s = 0; s < 14; ++s) { c.A[s] = a.lsrkA[s]; c.B[s] = a.lsrkB[s]; }
  const Batch2 *b = a.b;
  int nreq = a.nreq, nwork = a.nwork, nsteps = a.nsteps;
  void *args[] = {&b, &nreq, &nwork, &nsteps, &c};
  a.grid = grid;
  return hipLaunchCooperativeKernel(kern, dim3(grid), dim3(64), args, 0, a.stream);
}

template <int P, typename R>
static hipError_t launch_batch2_pr(BatchLaunch &a) {
  if (a.op == 2) return a.shallow ? launch_persist2<P, true, R>(a) : launch_persist2<P, false, R>(a);
  const int grid = std::max(1, std::min(a.nwork, kBatchGrid));
  if (a.op == 0) {
    if (a.shallow)
      hipLaunchKernelGGL((k_batch_traces2<P, 3, R>), dim3(grid), dim3(64), 0, a.stream, a.b, a.nreq, a.nwork);
    else
      hipLaunchKernelGGL((k_batch_traces2<P, 4, R>), dim3(grid), dim3(64), 0, a.stream, a.b, a.nreq, a.nwork);
  } else {
    if (a.shallow)
      hipLaunchKernelGGL((k_batch_step2<P, true, R>), dim3(grid), dim3(64), 0, a.stream, a.b, a.nreq, a.nwork,
                         a.A, a.B, a.odd, a.next);
    else
      hipLaunchKernelGGL((k_batch_step2<P, false, R>), dim3(grid), dim3(64), 0, a.stream, a.b, a.nreq, a.nwork,
                         a.A, a.B, a.odd, a.next);
  }
  return hipGetLastError();
}

template <int P>
static hipError_t launch_batch2_p(BatchLaunch &a) {
  return a.mixed ? launch_batch2_pr<P, float>(a) : launch_batch2_pr<P, double>(a);
}

hipError_t launch_batch2(BatchLaunch &a) {
  if (a.nreq < 1 || a.nwork < 1 || !a.b || (a.op == 2 && a.nsteps < 1)) return hipErrorInvalidValue;
  switch (a.p) {
#define DGX_CASE(PP) case PP: return launch_batch2_p<PP>(a);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9) DGX_CASE(10)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

template <int P>
static hipError_t launch_lin2_p(const LinArgs &a) {
  if (a.adjoint && a.mixed)
    hipLaunchKernelGGL((k_lin2<P, true, float>), dim3(a.ne), dim3(64), 0, a.stream, a);
  else if (a.adjoint)
    hipLaunchKernelGGL((k_lin2<P, true>), dim3(a.ne), dim3(64), 0, a.stream, a);
  else if (a.mixed)
    hipLaunchKernelGGL((k_lin2<P, false, float>), dim3(a.ne), dim3(64), 0, a.stream, a);
  else
    hipLaunchKernelGGL((k_lin2<P, false>), dim3(a.ne), dim3(64), 0, a.stream, a);
  return hipGetLastError();
}

hipError_t launch_lin2(const LinArgs &a) {
  switch (a.p) {
#define DGX_CASE(PP) case PP: return launch_lin2_p<PP>(a);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9) DGX_CASE(10)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_observe(const double *U, const int *rec_e, const double *rec_w, const double *cobs, int nrec,
                          int ne, int L2, double *y, hipStream_t s) {
  hipLaunchKernelGGL(k_observe, dim3(nrec), dim3(64), 0, s, U, rec_e, rec_w, cobs, ne, L2, y);
  return hipGetLastError();
}

hipError_t launch_observe_adjoint(const double *U, double *Ubar, const int *rec_e, const double *rec_w,
                                  const double *cobs, int nrec, int ne, int L2, const double *d, double dt,
                                  double *J, hipStream_t s) {
  hipLaunchKernelGGL(k_observe_adjoint, dim3(1), dim3(64), 0, s, U, Ubar, rec_e, rec_w, cobs, nrec, ne, L2, d,
                     dt, J);
  return hipGetLastError();
}

template <int P>
static hipError_t launch2(const LaunchArgs &a) {
  const int ne = a.ne;
  switch (a.op) {
    case OP_TRACES:
      if (a.mixed)
        hipLaunchKernelGGL((k_traces2<P, float>), dim3(ne), dim3(64), 0, a.stream, a.U, a.T, a.tab, ne, a.nfield);
      else
        hipLaunchKernelGGL(k_traces2<P>, dim3(ne), dim3(64), 0, a.stream, a.U, a.T, a.tab, ne, a.nfield);
      break;
    case OP_RHS:
      if (a.mixed && a.eqn == 30)
        hipLaunchKernelGGL((k_shallow2<P, float>), dim3(ne), dim3(64), 0, a.stream, a.U, a.T, a.R, a.Bal, a.bal,
                           a.tab, a.geo, a.phys, ne, a.flag);
      else if (a.mixed)
        hipLaunchKernelGGL((k_rhs2<P, false, float>), dim3(ne), dim3(64), 0, a.stream, a.U, a.T, a.V, a.TV,
                           a.R, a.tab, a.geo, a.phys, ne, a.flag);
      else if (a.eqn == 30)
        hipLaunchKernelGGL(k_shallow2<P>, dim3(ne), dim3(64), 0, a.stream, a.U, a.T, a.R, a.Bal, a.bal,
                           a.tab, a.geo, a.phys, ne, a.flag);
      else if (a.ns)
        hipLaunchKernelGGL((k_rhs2<P, true>), dim3(ne), dim3(64), 0, a.stream, a.U, a.T, a.V, a.TV,
                           a.R, a.tab, a.geo, a.phys, ne, a.flag);
      else
        hipLaunchKernelGGL((k_rhs2<P, false>), dim3(ne), dim3(64), 0, a.stream, a.U, a.T, a.V, a.TV,
                           a.R, a.tab, a.geo, a.phys, ne, a.flag);
      break;
    case OP_STEP:
      if (a.mixed && a.eqn == 30)
        hipLaunchKernelGGL((k_shallow_step2<P, float>), dim3(ne), dim3(64), 0, a.stream, a.Us, a.T, a.Ut, a.Bal,
                           a.bal, a.geo, a.phys, ne, a.flag, a.ca, a.cb, -a.dt, a.Tn);
      else if (a.mixed)
        hipLaunchKernelGGL((k_step2<P, false, float>), dim3(ne), dim3(64), 0, a.stream, a.Us, a.T, a.V, a.TV,
                           a.Ut, a.geo, a.phys, ne, a.flag, a.ca, a.cb, -a.dt, a.Tn);
      else if (a.eqn == 30)
        hipLaunchKernelGGL(k_shallow_step2<P>, dim3(ne), dim3(64), 0, a.stream, a.Us, a.T, a.Ut, a.Bal, a.bal,
                           a.geo, a.phys, ne, a.flag, a.ca, a.cb, -a.dt, a.Tn);
      else if (a.ns)
        hipLaunchKernelGGL((k_step2<P, true>), dim3(ne), dim3(64), 0, a.stream, a.Us, a.T, a.V, a.TV, a.Ut,
                           a.geo, a.phys, ne, a.flag, a.ca, a.cb, -a.dt, a.Tn);
      else
        hipLaunchKernelGGL((k_step2<P, false>), dim3(ne), dim3(64), 0, a.stream, a.Us, a.T, a.V, a.TV, a.Ut,
                           a.geo, a.phys, ne, a.flag, a.ca, a.cb, -a.dt, a.Tn);
      break;
    case OP_GRAD:
      hipLaunchKernelGGL(k_grad2<P>, dim3(ne), dim3(64), 0, a.stream, a.U, a.T, a.G, a.tab, a.geo, ne);
      break;
    case OP_VISC:
      hipLaunchKernelGGL(k_visc2<P>, dim3(ne), dim3(64), 0, a.stream, a.U, a.G, a.V, a.TV, a.tab,
                         a.geo, a.phys, ne);
      break;
    case OP_FORWARD:
      hipLaunchKernelGGL(k_forward2<P>, dim3(ne * a.nfield), dim3(64), 0, a.stream, a.X, a.Y, a.tab, ne);
      break;
    case OP_BACKWARD:
      hipLaunchKernelGGL(k_backward2<P>, dim3(ne * a.nfield), dim3(64), 0, a.stream, a.X, a.Y, a.tab, ne);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int lds_bytes2(int p, int which) { return (p >= 1 && p <= 10) ? 0 : -1; }

hipError_t launch_2d(const LaunchArgs &a) {
  switch (a.p) {
#define DGX_CASE(PP) case PP: return launch2<PP>(a);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9) DGX_CASE(10)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t upload_tables_2d(const double *h) {
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_tab), h, sizeof(double) * kTabTotal);
  if (e != hipSuccess) return e;
  std::vector<float> f(h, h + kTabTotal);
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tabf), f.data(), sizeof(float) * kTabTotal);
}

}  // namespace dgx
