// Host pieces shared by the discrete-adjoint drivers (dgx_adjoint.cpp,
// dgx_euler_adjoint.cpp): the binomial checkpointing schedule and receiver
// placement on affine quads.  Internal: not part of the C ABI.
#pragma once
#include "dgx_domain.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <string>
#include <vector>

namespace dgx {

// normalized Legendre phi_l(x) = sqrt((2l+1)/2) P_l(x) (three-term recurrence)
inline double normleg(int l, double x) {
  long double p0 = 1.0L, p1 = x;
  if (l == 0) return std::sqrt(0.5);
  for (int k = 2; k <= l; ++k) {
    const long double p2 = ((2 * k - 1) * (long double)x * p1 - (k - 1) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  return (double)(p1 * sqrtl((2 * l + 1) / 2.0L));
}

// number of steps binomial checkpointing reverses with s free slots and r
// advances per step: C(s + r, s), capped at `cap`
inline long long revolve_beta(int s, int r, long long cap) {
  long long b = 1;
  for (int i = 1; i <= r; ++i) {
    b = b * (s + i) / i;
    if (b >= cap) return cap;
  }
  return b;
}

// Binomial checkpointing schedule.  rev(a, b, slot, free) reverses steps
// [a, b) given state_a in `slot` and `free` unused slots: it advances to a
// split point m, checkpoints there, reverses [m, b) with one slot fewer and
// then [a, m) with the same slots.  m is chosen so that both parts fit the
// smallest repetition number r with C(free + r, free) >= b - a (Griewank and
// Walther's binomial bound).  reverse(k) is always called right after
// advance(k), with state_{k+1} current, in the order k = n-1, n-2, .., 0.
struct Revolve {
  std::function<int(int)> restore, store, advance, reverse;
  std::vector<int> pool;
  long long advances = 0, reversals = 0;
  int max_used = 0, used = 0;

  // head length: the head [a, m) is reversed later with the same slots and one
  // advance fewer (at most C(s + r - 1, s) steps), the tail with one slot fewer
  // (at most C(s - 1 + r, s - 1) steps); the middle of the feasible range
  int split(int n, int s) const {
    int r = 1;
    while (revolve_beta(s, r, n) < n) ++r;
    const long long lo = std::max<long long>(1, n - revolve_beta(s - 1, r, n));
    const long long hi = std::min<long long>(n - 1, revolve_beta(s, r - 1, n));
    return (int)((lo + std::max(lo, hi)) / 2);
  }
  int adv(int a, int b) {
    for (int j = a; j < b; ++j) {
      int rc = advance(j);
      if (rc) return rc;
      ++advances;
    }
    return 0;
  }
  int rev(int a, int b, int slot, int free) {
    int rc;
    if (b - a == 1 || free == 0) {
      for (int k = b - 1; k >= a; --k) {   // free == 0: recompute from `slot` each time
        if ((rc = restore(slot)) || (rc = adv(a, k + 1)) || (rc = reverse(k))) return rc;
        ++reversals;
      }
      return 0;
    }
    const int m = a + split(b - a, free);
    const int s2 = pool.back();
    pool.pop_back();
    max_used = std::max(max_used, ++used);
    if ((rc = restore(slot)) || (rc = adv(a, m)) || (rc = store(s2)) || (rc = rev(m, b, s2, free - 1))) return rc;
    pool.push_back(s2);
    --used;
    return rev(a, m, slot, free);
  }
  // slots: number of state buffers, slot 0 holding state_0
  int run(int n, int slots) {
    pool.clear();
    for (int s = slots - 1; s >= 1; --s) pool.push_back(s);
    used = max_used = 1;
    return rev(0, n, 0, slots - 1);
  }
};

// Receiver points on affine quads: the element holding (x, y) (the first whose
// reference square contains it) and the modal weights phi_l(xi) phi_m(eta),
// [nrec][L^2].  Returns the index of a point outside the mesh, or -1.
inline int place_receivers(const dgx_domain_desc *mesh, const dgx_domain *d, const double *xrec, int nrec,
                           std::vector<int> &rec_e, std::vector<double> &rec_w) {
  const int L = d->p + 1, L2 = L * L;
  rec_e.assign(nrec, -1);
  rec_w.assign((size_t)nrec * L2, 0.0);
  for (int r = 0; r < nrec; ++r) {
    const double x = xrec[2 * r], y = xrec[2 * r + 1];
    int found = -1;
    double xi = 0, eta = 0;
    for (int e = 0; e < d->ne && found < 0; ++e) {
      const double *v = mesh->vert + (size_t)e * 2 * 4;
      const double cx = 0.25 * (v[0] + v[1] + v[2] + v[3]), cy = 0.25 * (v[4] + v[5] + v[6] + v[7]);
      const double *mt = d->metric.data() + 9 * (size_t)e;
      xi = mt[0] * (x - cx) + mt[1] * (y - cy);
      eta = mt[3] * (x - cx) + mt[4] * (y - cy);
      if (std::fabs(xi) <= 1 + 1e-12 && std::fabs(eta) <= 1 + 1e-12) found = e;
    }
    if (found < 0) return r;
    rec_e[r] = found;
    for (int l = 0; l < L; ++l)
      for (int m = 0; m < L; ++m) rec_w[(size_t)r * L2 + l * L + m] = normleg(l, xi) * normleg(m, eta);
  }
  return -1;
}

}  // namespace dgx
