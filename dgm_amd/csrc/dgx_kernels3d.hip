// Hex (3D) kernels of the DG residual.  One workgroup processes one element;
// its tensors live in LDS and every contraction is a register-blocked pencil
// sweep (dgx_device.hpp).  State is modal ([field][elem][L^3]); nodal values
// at the q^3 GLL points exist only inside a kernel.
//
// Reference operators restated here (paths relative to the reference root):
//   face traces      Hex::fill_sides / interpolate_to_side  src/Hex.cpp:933-947, 1333-1356
//   volume term      Euler3d::convection                    src/Euler3d.cpp:468-575
//   numerical flux   Euler3d::convective_flux (LF)          src/Euler3d.cpp:584-817
//   lift + project   Hex::lift, Element::project            src/Hex.cpp:1786-1952, src/Element.hpp:427-437
//   BR1 viscous      Navier_Stokes3d::diffusion             src/Navier_Stokes3d.cpp:182-391
#include "dgx_kernels.hpp"

namespace dgx {

template <int P> struct H3 {
  static constexpr int L = P + 1, Q = P + 2, QF = P + 1;
  static constexpr int L2 = L * L, L3 = L * L * L, Q2 = Q * Q, Q3 = Q * Q * Q, QF2 = QF * QF;
  static constexpr int cmax(int a, int b) { return a > b ? a : b; }
  static constexpr int NT = (Q2 <= 64) ? 64 : (Q2 <= 128 ? 128 : 256);
  // scratch blocks
  static constexpr int SX = cmax(Q2 * L, 6 * L * QF);     // X / face W
  static constexpr int SY = cmax(2 * Q * L2, 6 * L2);     // Y1,Y2 / face H
};

// ---------------------------------------------------------------------------
// backward transform of one modal block (LDS `m`, L^3) into nodal `out` (Q^3)
template <int P>
__device__ __forceinline__ void backward3(const Tab &tb, const double *m, double *s1, double *s2,
                                          double *out, int tid) {
  using S = H3<P>;
  constexpr int L = S::L, Q = S::Q, NT = S::NT;
  contract<L, L, L, 0, Q, NT>(CT<P>::BT(), m, s1, tid);
  __syncthreads();
  contract<Q, L, L, 1, Q, NT>(CT<P>::BT(), s1, s2, tid);
  __syncthreads();
  contract<Q, Q, L, 2, Q, NT>(CT<P>::BT(), s2, out, tid);
  __syncthreads();
}

// modal projection of a divergence: out = [QD(x)BW(x)BW] Fr + [BW(x)QD(x)BW] Fs + [BW(x)BW(x)QD] Ft
// Fr, Fs, Ft: nodal Q^3 blocks (may be the same buffer refilled by `fill`).
template <int P, typename Fill>
__device__ __forceinline__ void divergence3(const Tab &tb, double *Fb, double *X, double *Y1,
                                            double *Y2, double *out, Fill fill, int tid) {
  using S = H3<P>;
  constexpr int L = S::L, Q = S::Q, NT = S::NT;
  fill(0, Fb);
  __syncthreads();
  contract<Q, Q, Q, 2, L, NT>(CT<P>::BW(), Fb, X, tid);
  __syncthreads();
  contract<Q, Q, L, 1, L, NT>(CT<P>::BW(), X, Y1, tid);
  fill(1, Fb);
  __syncthreads();
  contract<Q, Q, Q, 2, L, NT>(CT<P>::BW(), Fb, X, tid);
  __syncthreads();
  contract<Q, Q, L, 1, L, NT>(CT<P>::QD(), X, Y2, tid);
  fill(2, Fb);
  __syncthreads();
  contract<Q, Q, Q, 2, L, NT>(CT<P>::QD(), Fb, X, tid);
  __syncthreads();
  contract<Q, Q, L, 1, L, NT, true>(CT<P>::BW(), X, Y2, tid);
  __syncthreads();
  contract<Q, L, L, 0, L, NT>(CT<P>::QD(), Y1, out, tid);
  __syncthreads();
  contract<Q, L, L, 0, L, NT, true>(CT<P>::BW(), Y2, out, tid);
  __syncthreads();
}

// Face -> modal: H_s = (PW (x) PW) g_s for the six faces (g: [6][QF][QF]),
// result [6][L][L] in H.  W: scratch [6][L][QF].
template <int P>
__device__ __forceinline__ void face_project3(const Tab &tb, const double *g, double *W, double *H,
                                              int tid) {
  using S = H3<P>;
  constexpr int L = S::L, QF = S::QF, NT = S::NT;
  contract<6, QF, QF, 1, L, NT>(CT<P>::PW(), g, W, tid);
  __syncthreads();
  contract<6, L, QF, 2, L, NT>(CT<P>::PW(), W, H, tid);
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Face traces: modal U -> conservative values at the Gauss points of all six
// faces, T[e][s][field][a] (face-local order of src/Hex.cpp:817-866).
template <int P>
__global__ void __launch_bounds__(H3<P>::NT)
k_traces3(const double *__restrict__ U, double *__restrict__ T, Tab tb, int ne, int nf) {
  using S = H3<P>;
  constexpr int L = S::L, QF = S::QF, L2 = S::L2, L3 = S::L3, QF2 = S::QF2, NT = S::NT;
  extern __shared__ double sm[];
  double *A = sm, *K0 = A + L3, *K1 = K0 + 2 * L2, *K2 = K1 + 2 * L2;
  double *W = K2 + 2 * L2, *O = W + 2 * QF * L;
  const int e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;
  for (int v = 0; v < nf; ++v) {
    copy_block<NT>(U + ((size_t)v * ne + e) * L3, A, L3, tid);
    __syncthreads();
    contract<L, L, L, 0, 2, NT>(CT<P>::Bend(), A, K0, tid);
    contract<L, L, L, 1, 2, NT>(CT<P>::Bend(), A, K1, tid);
    contract<L, L, L, 2, 2, NT>(CT<P>::Bend(), A, K2, tid);
    __syncthreads();
    double *Te = T + (size_t)e * 6 * nf * QF2;
    // faces 4 (xi=-1) and 2 (xi=+1): [2][QF(b)][QF(c)]
    contract<2, L, L, 1, QF, NT>(CT<P>::Bg(), K0, W, tid);
    __syncthreads();
    contract<2, QF, L, 2, QF, NT>(CT<P>::Bg(), W, O, tid);
    __syncthreads();
    for (int i = tid; i < 2 * QF2; i += NT) {
      const int end = i / QF2, a = i - end * QF2;
      Te[((end ? 2 : 4) * nf + v) * QF2 + a] = O[i];
    }
    __syncthreads();
    // faces 1 (eta=-1) and 3 (eta=+1): [QF(a)][2][QF(c)]
    contract<L, 2, L, 0, QF, NT>(CT<P>::Bg(), K1, W, tid);
    __syncthreads();
    contract<QF, 2, L, 2, QF, NT>(CT<P>::Bg(), W, O, tid);
    __syncthreads();
    for (int i = tid; i < 2 * QF2; i += NT) {
      const int ia = i / (2 * QF), r = i - ia * 2 * QF, end = r / QF, ic = r - end * QF;
      Te[((end ? 3 : 1) * nf + v) * QF2 + ia * QF + ic] = O[i];
    }
    __syncthreads();
    // faces 0 (zeta=-1) and 5 (zeta=+1): [QF(a)][QF(b)][2]
    contract<L, L, 2, 0, QF, NT>(CT<P>::Bg(), K2, W, tid);
    __syncthreads();
    contract<QF, L, 2, 1, QF, NT>(CT<P>::Bg(), W, O, tid);
    __syncthreads();
    for (int i = tid; i < 2 * QF2; i += NT) {
      const int ab = i >> 1, end = i & 1;
      Te[((end ? 5 : 0) * nf + v) * QF2 + ab] = O[i];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Residual F = -dU/dt (modal).  NS adds the BR1 viscous volume term V and the
// viscous face traces TV computed by k_grad3/k_visc3.
template <int P, bool NS>
__global__ void __launch_bounds__(H3<P>::NT)
k_rhs3(const double *__restrict__ U, const double *__restrict__ T, const double *__restrict__ V,
       const double *__restrict__ TV, double *__restrict__ R, Tab tb, Geo geo, Phys ph, int ne,
       int *__restrict__ flag) {
  using S = H3<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, L3 = S::L3, Q3 = S::Q3,
                QF2 = S::QF2, NT = S::NT;
  extern __shared__ double sm[];
  double *Un = sm;                 // [5][Q3]
  double *gF = Un + 5 * Q3;        // [5][6][QF2]
  double *Fb = gF + 30 * QF2;      // [Q3]
  double *X = Fb + Q3;             // [SX]
  double *Y = X + S::SX;           // [SY]
  double *Y1 = Y, *Y2 = Y + Q * L2;
  const int e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;
  const double gm1 = ph.gamma - 1.0;

  for (int v = 0; v < 5; ++v) {
    copy_block<NT>(U + ((size_t)v * ne + e) * L3, Fb, L3, tid);
    __syncthreads();
    backward3<P>(tb, Fb, Y, X, Un + v * Q3, tid);
  }

  // face corrections g = (J_side/J_vol) (Fhat - F(U-).n) [+ viscous central flux]
  bool bad = false;
  for (int i = tid; i < 6 * QF2; i += NT) {
    const int s = i / QF2, a = i - s * QF2;
    const int es = e * 6 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int an = geo.perm[geo.permid[es] * QF2 + a];
    const double nx = geo.fnorm[3 * es], ny = geo.fnorm[3 * es + 1], nz = geo.fnorm[3 * es + 2];
    const double jr = geo.fjr[es];
    double UL[5], UR[5], FL[5];
    const double *tl = T + ((size_t)e * 6 + s) * 5 * QF2;
#pragma unroll
    for (int v = 0; v < 5; ++v) UL[v] = tl[v * QF2 + a];
    if (nb >= 0) {
      const double *tr = T + ((size_t)nb * 6 + ns) * 5 * QF2;
#pragma unroll
      for (int v = 0; v < 5; ++v) UR[v] = tr[v * QF2 + an];
    } else {
      ghost3(-1 - nb, UL, nx, ny, nz, ph.bcU, UR);
    }
    double Fh[5];
    riemann3(UL, UR, nx, ny, nz, ph.gamma, ph.flux, FL, Fh, bad);
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      double g = jr * (Fh[v] - FL[v]);
      if (NS && v > 0) {
        const double *vl = TV + ((size_t)e * 6 + s) * 4 * QF2;
        const double *vr = TV + ((size_t)nb * 6 + ns) * 4 * QF2;
        g += 0.5 * jr * (vl[(v - 1) * QF2 + a] + vr[(v - 1) * QF2 + an]);
      }
      gF[(v * 6 + s) * QF2 + a] = g;
    }
  }
  if (bad) atomicOr(flag, 1);
  __syncthreads();

  const double *mt = geo.metric + 9 * (size_t)e;
  const double mtr[9] = {mt[0], mt[1], mt[2], mt[3], mt[4], mt[5], mt[6], mt[7], mt[8]};
  double cw[6];
  for (int v = 0; v < 5; ++v) {
    // nodal contravariant Euler flux of field v in direction d -> buf
    auto fill = [&](int d, double *buf) {
      const double m0 = mtr[3 * d], m1 = mtr[3 * d + 1], m2 = mtr[3 * d + 2];
      for (int k = tid; k < Q3; k += NT) {
        const double r = Un[k], ru = Un[Q3 + k], rv = Un[2 * Q3 + k], rw = Un[3 * Q3 + k],
                     rE = Un[4 * Q3 + k];
        const double ir = 1.0 / r;
        const double u = ru * ir, vv = rv * ir, w = rw * ir;
        const double p = gm1 * (rE - 0.5 * (ru * u + rv * vv + rw * w));
        double fx, fy, fz;
        if (v == 0) { fx = ru; fy = rv; fz = rw; }
        else if (v == 4) { const double h = rE + p; fx = u * h; fy = vv * h; fz = w * h; }
        else {
          const double c = (v == 1) ? ru : (v == 2 ? rv : rw);
          fx = c * u; fy = c * vv; fz = c * w;
          if (v == 1) fx += p; else if (v == 2) fy += p; else fz += p;
        }
        buf[k] = m0 * fx + m1 * fy + m2 * fz;
      }
    };
    double *Rv = Fb;   // L3 <= Q3; written only after the last fill
    divergence3<P>(tb, Fb, X, Y1, Y2, Rv, fill, tid);
    face_project3<P>(tb, gF + v * 6 * QF2, X, Y, tid);
    for (int s = 0; s < 6; ++s) cw[s] = 1.0;
    for (int k = tid; k < L3; k += NT) {
      const int l = k / L2, m = (k / L) % L, n = k % L;
      double r = Rv[k] + face_add3<P>(CT<P>::Bend(), Y, l, m, n, cw);
      if (NS && v > 0) r -= V[((size_t)e * 4 + (v - 1)) * L3 + k];
      R[((size_t)v * ne + e) * L3 + k] = r;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// BR1 lifted, projected gradients of the primitive variables (u, v, w, e),
// modal, G[e][3*prim + dir][L3]  (Euler3d::gradient src/Euler3d.cpp:1144-1160,
// primitive_sides :1162-1199, Domain::average_flux src/Domain.cpp:1883-1937).
template <int P>
__global__ void __launch_bounds__(H3<P>::NT)
k_grad3(const double *__restrict__ U, const double *__restrict__ T, double *__restrict__ G,
        Tab tb, Geo geo, int ne) {
  using S = H3<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, L3 = S::L3, Q3 = S::Q3,
                QF2 = S::QF2, NT = S::NT;
  extern __shared__ double sm[];
  double *Un = sm;                 // [5][Q3]
  double *Jq = Un + 5 * Q3;        // [4][6][QF2]  J_side/J_vol ({q} - q-)
  double *Fb = Jq + 24 * QF2;      // [Q3]
  double *Xa = Fb + Q3;            // [SX]
  double *Xc = Xa + S::SX;         // [SX]
  double *Ya = Xc + S::SX;         // [Q L2] x 3
  double *Yb = Ya + Q * L2, *Yc = Yb + Q * L2;
  double *D3 = Yc + Q * L2;        // [3][L3]
  const int e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;

  for (int v = 0; v < 5; ++v) {
    copy_block<NT>(U + ((size_t)v * ne + e) * L3, Fb, L3, tid);
    __syncthreads();
    backward3<P>(tb, Fb, Ya, Xa, Un + v * Q3, tid);
  }
  for (int i = tid; i < 6 * QF2; i += NT) {
    const int s = i / QF2, a = i - s * QF2;
    const int es = e * 6 + s;
    const int nb = geo.nbr[es], ns = geo.nside[es];
    const int an = geo.perm[geo.permid[es] * QF2 + a];
    const double jr = geo.fjr[es];
    const double *tl = T + ((size_t)e * 6 + s) * 5 * QF2;
    const double *tr = T + ((size_t)nb * 6 + ns) * 5 * QF2;
    double pl[4], pr[4];
    {
      const double r = tl[a], u = tl[QF2 + a] / r, v = tl[2 * QF2 + a] / r, w = tl[3 * QF2 + a] / r;
      double en = tl[4 * QF2 + a] / r;
      double wk = u * u; wk = v * v + wk; wk = w * w + wk;
      en = -0.5 * wk + en;
      pl[0] = u; pl[1] = v; pl[2] = w; pl[3] = en;
    }
    {
      const double r = tr[an], u = tr[QF2 + an] / r, v = tr[2 * QF2 + an] / r, w = tr[3 * QF2 + an] / r;
      double en = tr[4 * QF2 + an] / r;
      double wk = u * u; wk = v * v + wk; wk = w * w + wk;
      en = -0.5 * wk + en;
      pr[0] = u; pr[1] = v; pr[2] = w; pr[3] = en;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) Jq[(q * 6 + s) * QF2 + a] = jr * (0.5 * (pl[q] + pr[q]) - pl[q]);
  }
  __syncthreads();

  const double *mt = geo.metric + 9 * (size_t)e;
  for (int q = 0; q < 4; ++q) {
    for (int k = tid; k < Q3; k += NT) {
      const double r = Un[k], u = Un[Q3 + k] / r, v = Un[2 * Q3 + k] / r, w = Un[3 * Q3 + k] / r;
      double val;
      if (q == 0) val = u;
      else if (q == 1) val = v;
      else if (q == 2) val = w;
      else {
        double en = Un[4 * Q3 + k] / r;
        double wk = u * u; wk = v * v + wk; wk = w * w + wk;
        val = -0.5 * wk + en;
      }
      Fb[k] = val;
    }
    __syncthreads();
    contract<Q, Q, Q, 2, L, NT>(CT<P>::BW(), Fb, Xa, tid);
    contract<Q, Q, Q, 2, L, NT>(CT<P>::QD(), Fb, Xc, tid);
    __syncthreads();
    contract<Q, Q, L, 1, L, NT>(CT<P>::BW(), Xa, Ya, tid);
    contract<Q, Q, L, 1, L, NT>(CT<P>::QD(), Xa, Yb, tid);
    contract<Q, Q, L, 1, L, NT>(CT<P>::BW(), Xc, Yc, tid);
    __syncthreads();
    contract<Q, L, L, 0, L, NT>(CT<P>::QD(), Ya, D3, tid);
    contract<Q, L, L, 0, L, NT>(CT<P>::BW(), Yb, D3 + L3, tid);
    contract<Q, L, L, 0, L, NT>(CT<P>::BW(), Yc, D3 + 2 * L3, tid);
    face_project3<P>(tb, Jq + q * 6 * QF2, Xa, Ya, tid);   // H -> Ya ([6][L][L])
    for (int k = tid; k < L3; k += NT) {
      const int l = k / L2, m = (k / L) % L, n = k % L;
      const double dr = D3[k], ds = D3[L3 + k], dt = D3[2 * L3 + k];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        double c[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) c[s] = geo.fnorm[3 * (e * 6 + s) + d];
        const double g = mt[d] * dr + mt[3 + d] * ds + mt[6 + d] * dt +
                         face_add3<P>(CT<P>::Bend(), Ya, l, m, n, c);
        G[((size_t)e * 12 + 3 * q + d) * L3 + k] = g;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// BR1 viscous fluxes from the lifted gradients: volume divergence (modal,
// V[e][eq][L3], eq = x,y,z momentum and energy) and the normal viscous flux at
// the Gauss points of every face TV[e][s][eq][a] (Hex::fill_sides(fx,fy,fz)
// src/Hex.cpp:950-973 followed by Navier_Stokes3d::diffusion's flux steps).
template <int P>
__global__ void __launch_bounds__(256)
k_visc3(const double *__restrict__ U, const double *__restrict__ G, double *__restrict__ V,
        double *__restrict__ TV, Tab tb, Geo geo, Phys ph, int ne) {
  using S = H3<P>;
  constexpr int L = S::L, Q = S::Q, QF = S::QF, L2 = S::L2, L3 = S::L3, Q2 = S::Q2,
                Q3 = S::Q3, QF2 = S::QF2;
  constexpr int NT = 256;
  extern __shared__ double sm[];
  double *Un = sm;               // [5][Q3]
  double *Gn = Un + 5 * Q3;      // [12][Q3]
  double *Sc = Gn + 12 * Q3;     // scratch
  const int e = xcd_element(blockIdx.x, ne), tid = threadIdx.x;
  // backward transforms (scratch: Sc[0:L3], Sc+L3 [Q L2], then [Q2 L])
  double *m0 = Sc, *s1 = Sc + L3, *s2 = s1 + Q * L2;
  for (int v = 0; v < 17; ++v) {
    const double *src = v < 5 ? U + ((size_t)v * ne + e) * L3 : G + ((size_t)e * 12 + (v - 5)) * L3;
    double *dst = v < 5 ? Un + v * Q3 : Gn + (v - 5) * Q3;
    copy_block<NT>(src, m0, L3, tid);
    __syncthreads();
    contract<L, L, L, 0, Q, NT>(CT<P>::BT(), m0, s1, tid);
    __syncthreads();
    contract<Q, L, L, 1, Q, NT>(CT<P>::BT(), s1, s2, tid);
    __syncthreads();
    contract<Q, Q, L, 2, Q, NT>(CT<P>::BT(), s2, dst, tid);
    __syncthreads();
  }
  const double mu = ph.vis, kap = ph.vis * ph.gamma / ph.Pr;
  // physical viscous flux of equation `eq` (0..3) at node k, direction x/y/z
  auto vflux = [&](int k, double out[4][3]) {
    const double r = Un[k];
    const double u = Un[Q3 + k] / r, v = Un[2 * Q3 + k] / r, w = Un[3 * Q3 + k] / r;
    const double ux = Gn[k], uy = Gn[Q3 + k], uz = Gn[2 * Q3 + k];
    const double vx = Gn[3 * Q3 + k], vy = Gn[4 * Q3 + k], vz = Gn[5 * Q3 + k];
    const double wx = Gn[6 * Q3 + k], wy = Gn[7 * Q3 + k], wz = Gn[8 * Q3 + k];
    const double ex = Gn[9 * Q3 + k], ey = Gn[10 * Q3 + k], ez = Gn[11 * Q3 + k];
    const double t12 = mu * (uy + vx), t13 = mu * (uz + wx), t23 = mu * (vz + wy);
    const double t11 = mu * ((2.0 / 3.0) * ((2.0 * ux - vy) - wz));
    const double t22 = mu * ((2.0 / 3.0) * ((2.0 * vy - ux) - wz));
    const double t33 = mu * ((2.0 / 3.0) * ((2.0 * wz - ux) - vy));
    out[0][0] = t11; out[0][1] = t12; out[0][2] = t13;
    out[1][0] = t12; out[1][1] = t22; out[1][2] = t23;
    out[2][0] = t13; out[2][1] = t23; out[2][2] = t33;
    out[3][0] = t11 * u + t12 * v + t13 * w + kap * ex;
    out[3][1] = t12 * u + t22 * v + t23 * w + kap * ey;
    out[3][2] = t13 * u + t23 * v + t33 * w + kap * ez;
  };
  // normal viscous flux on the GLL face layers: Lay[s][eq][Q2]
  double *Lay = Sc, *Wf = Sc + 24 * Q2;
  for (int i = tid; i < 6 * Q2; i += NT) {
    const int s = i / Q2, ij = i - s * Q2, j1 = ij / Q, j2 = ij - j1 * Q;
    int k;
    switch (s) {
      case 0: k = (j1 * Q + j2) * Q; break;
      case 1: k = (j1 * Q) * Q + j2; break;
      case 2: k = ((Q - 1) * Q + j1) * Q + j2; break;
      case 3: k = (j1 * Q + Q - 1) * Q + j2; break;
      case 4: k = (j1)*Q + j2; break;
      default: k = (j1 * Q + j2) * Q + Q - 1; break;
    }
    const double *nn = geo.fnorm + 3 * (e * 6 + s);
    double f[4][3];
    vflux(k, f);
#pragma unroll
    for (int q = 0; q < 4; ++q)
      Lay[(s * 4 + q) * Q2 + ij] = f[q][0] * nn[0] + f[q][1] * nn[1] + f[q][2] * nn[2];
  }
  __syncthreads();
  contract<24, Q, Q, 1, QF, NT>(CT<P>::Ig(), Lay, Wf, tid);
  __syncthreads();
  contract<24, QF, Q, 2, QF, NT>(CT<P>::Ig(), Wf, Lay, tid);   // [24][QF][QF] == [6][4][QF2]
  __syncthreads();
  copy_block<NT>(Lay, TV + (size_t)e * 24 * QF2, 24 * QF2, tid);
  __syncthreads();
  // contravariant viscous flux in place: Gn[3*eq + d] = sum_x m_dx sigma_eq,x
  const double *mt = geo.metric + 9 * (size_t)e;
  for (int k = tid; k < Q3; k += NT) {
    double f[4][3];
    vflux(k, f);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int d = 0; d < 3; ++d)
        Gn[(3 * q + d) * Q3 + k] = mt[3 * d] * f[q][0] + mt[3 * d + 1] * f[q][1] + mt[3 * d + 2] * f[q][2];
  }
  __syncthreads();
  double *X = Sc, *Y1 = Sc + Q2 * L, *Y2 = Y1 + Q * L2, *Vo = Y2 + Q * L2;
  for (int q = 0; q < 4; ++q) {
    const double *Fr = Gn + 3 * q * Q3, *Fs = Fr + Q3, *Ft = Fs + Q3;
    contract<Q, Q, Q, 2, L, NT>(CT<P>::BW(), Fr, X, tid);
    __syncthreads();
    contract<Q, Q, L, 1, L, NT>(CT<P>::BW(), X, Y1, tid);
    __syncthreads();
    contract<Q, Q, Q, 2, L, NT>(CT<P>::BW(), Fs, X, tid);
    __syncthreads();
    contract<Q, Q, L, 1, L, NT>(CT<P>::QD(), X, Y2, tid);
    __syncthreads();
    contract<Q, Q, Q, 2, L, NT>(CT<P>::QD(), Ft, X, tid);
    __syncthreads();
    contract<Q, Q, L, 1, L, NT, true>(CT<P>::BW(), X, Y2, tid);
    __syncthreads();
    contract<Q, L, L, 0, L, NT>(CT<P>::QD(), Y1, Vo, tid);
    __syncthreads();
    contract<Q, L, L, 0, L, NT, true>(CT<P>::BW(), Y2, Vo, tid);
    __syncthreads();
    copy_block<NT>(Vo, V + ((size_t)e * 4 + q) * L3, L3, tid);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// nodal <-> modal transforms ([field][elem][node] <-> [field][elem][mode]).
// forward: fh = M^-1 B^T W J f = (BW (x) BW (x) BW) f on affine elements.
template <int P>
__global__ void __launch_bounds__(H3<P>::NT)
k_forward3(const double *__restrict__ Fn, double *__restrict__ Fm, Tab tb, int ne) {
  using S = H3<P>;
  constexpr int L = S::L, Q = S::Q, L2 = S::L2, L3 = S::L3, Q3 = S::Q3, NT = S::NT;
  extern __shared__ double sm[];
  double *A = sm, *X = A + Q3, *Y = X + Q * Q * L, *O = Y + Q * L2;
  const int e = blockIdx.x % ne, v = blockIdx.x / ne, tid = threadIdx.x;
  copy_block<NT>(Fn + ((size_t)v * ne + e) * Q3, A, Q3, tid);
  __syncthreads();
  contract<Q, Q, Q, 2, L, NT>(CT<P>::BW(), A, X, tid);
  __syncthreads();
  contract<Q, Q, L, 1, L, NT>(CT<P>::BW(), X, Y, tid);
  __syncthreads();
  contract<Q, L, L, 0, L, NT>(CT<P>::BW(), Y, O, tid);
  __syncthreads();
  copy_block<NT>(O, Fm + ((size_t)v * ne + e) * L3, L3, tid);
}

template <int P>
__global__ void __launch_bounds__(H3<P>::NT)
k_backward3(const double *__restrict__ Fm, double *__restrict__ Fn, Tab tb, int ne) {
  using S = H3<P>;
  constexpr int L = S::L, Q = S::Q, L2 = S::L2, L3 = S::L3, Q3 = S::Q3, NT = S::NT;
  extern __shared__ double sm[];
  double *A = sm, *s1 = A + L3, *s2 = s1 + Q * L2, *O = s2 + Q * Q * L;
  const int e = blockIdx.x % ne, v = blockIdx.x / ne, tid = threadIdx.x;
  copy_block<NT>(Fm + ((size_t)v * ne + e) * L3, A, L3, tid);
  __syncthreads();
  backward3<P>(tb, A, s1, s2, O, tid);
  copy_block<NT>(O, Fn + ((size_t)v * ne + e) * Q3, Q3, tid);
}

// ---------------------------------------------------------------------------
// launchers
template <int P>
static int lds3(int which) {
  using S = H3<P>;
  constexpr int L3 = S::L3, Q3 = S::Q3, QF2 = S::QF2, L2 = S::L2, Q = S::Q, L = S::L, QF = S::QF,
                Q2 = S::Q2;
  switch (which) {
    case 0: return (L3 + 6 * L2 + 2 * QF * L + 2 * QF2) * 8;                       // traces
    case 1: return (5 * Q3 + 30 * QF2 + Q3 + S::SX + S::SY) * 8;                  // rhs
    case 2: return (5 * Q3 + 24 * QF2 + Q3 + 2 * S::SX + 3 * Q * L2 + 3 * L3) * 8;  // grad
    case 3: {                                                                     // visc
      int a = L3 + Q * L2 + Q2 * L, b = 24 * Q2 + 24 * QF * Q, c = Q2 * L + 2 * Q * L2 + L3;
      int m = a > b ? a : b; m = m > c ? m : c;
      return (17 * Q3 + m) * 8;
    }
    default: return (Q3 + Q * Q * L + Q * L2 + L3) * 8;                           // transforms
  }
}

template <int P>
static hipError_t launch3(const LaunchArgs &a) {
  using S = H3<P>;
  const int ne = a.ne;
  switch (a.op) {
    case OP_TRACES: {
      const int sh = lds3<P>(0);
      hipLaunchKernelGGL(k_traces3<P>, dim3(ne), dim3(S::NT), sh, a.stream, a.U, a.T, a.tab, ne, 5);
      break;
    }
    case OP_RHS: {
      const int sh = lds3<P>(1);
      if (a.ns) {
        (void)hipFuncSetAttribute((const void *)k_rhs3<P, true>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
        hipLaunchKernelGGL((k_rhs3<P, true>), dim3(ne), dim3(S::NT), sh, a.stream, a.U, a.T, a.V, a.TV,
                           a.R, a.tab, a.geo, a.phys, ne, a.flag);
      } else {
        (void)hipFuncSetAttribute((const void *)k_rhs3<P, false>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
        hipLaunchKernelGGL((k_rhs3<P, false>), dim3(ne), dim3(S::NT), sh, a.stream, a.U, a.T, a.V, a.TV,
                           a.R, a.tab, a.geo, a.phys, ne, a.flag);
      }
      break;
    }
    case OP_GRAD: {
      const int sh = lds3<P>(2);
      (void)hipFuncSetAttribute((const void *)k_grad3<P>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
      hipLaunchKernelGGL(k_grad3<P>, dim3(ne), dim3(S::NT), sh, a.stream, a.U, a.T, a.G, a.tab, a.geo, ne);
      break;
    }
    case OP_VISC: {
      const int sh = lds3<P>(3);
      if (sh > 160 * 1024) return hipErrorInvalidValue;
      (void)hipFuncSetAttribute((const void *)k_visc3<P>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
      hipLaunchKernelGGL(k_visc3<P>, dim3(ne), dim3(256), sh, a.stream, a.U, a.G, a.V, a.TV, a.tab,
                         a.geo, a.phys, ne);
      break;
    }
    case OP_FORWARD: {
      const int sh = lds3<P>(4);
      (void)hipFuncSetAttribute((const void *)k_forward3<P>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
      hipLaunchKernelGGL(k_forward3<P>, dim3(ne * a.nfield), dim3(S::NT), sh, a.stream, a.X, a.Y,
                         a.tab, ne);
      break;
    }
    case OP_BACKWARD: {
      const int sh = lds3<P>(4);
      (void)hipFuncSetAttribute((const void *)k_backward3<P>, hipFuncAttributeMaxDynamicSharedMemorySize, sh);
      hipLaunchKernelGGL(k_backward3<P>, dim3(ne * a.nfield), dim3(S::NT), sh, a.stream, a.X, a.Y,
                         a.tab, ne);
      break;
    }
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int lds_bytes3(int p, int which) {
  switch (p) {
#define DGX_CASE(PP) case PP: return lds3<PP>(which);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9) DGX_CASE(10)
#undef DGX_CASE
    default: return -1;
  }
}

hipError_t launch_3d(const LaunchArgs &a) {
  switch (a.p) {
#define DGX_CASE(PP) case PP: return launch3<PP>(a);
    DGX_CASE(1) DGX_CASE(2) DGX_CASE(3) DGX_CASE(4) DGX_CASE(5)
    DGX_CASE(6) DGX_CASE(7) DGX_CASE(8) DGX_CASE(9) DGX_CASE(10)
#undef DGX_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t upload_tables_3d(const double *h) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), h, sizeof(double) * kTabTotal);
}

}  // namespace dgx
