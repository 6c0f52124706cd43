// Batched neighbour-list traversal shared by the pair kernels.
#pragma once

#include "common.hpp"

namespace sophia {

// Fields of one neighbour, gathered before any of them is used.
template <typename R>
struct Nb {
  double x, y, z, m, rho, pres, h;
  R ux, uy, uz, temp, grx, gry, grz;
  double rref, por;
  int pt;
};
template <typename R>
__device__ inline Nb<R> load_nb(const SphView<R>& v, int j) {
  Nb<R> q;
  q.x = v.x[j]; q.y = v.y[j]; q.z = v.z[j];
  q.m = v.m[j]; q.rho = v.rho[j]; q.pres = v.pres[j]; q.h = v.h[j];
  q.ux = v.ux[j]; q.uy = v.uy[j]; q.uz = v.uz[j]; q.temp = v.temp[j]; q.rref = v.rho_ref[j];
  q.grx = v.grx[j]; q.gry = v.gry[j]; q.grz = v.grz[j]; q.por = v.por[j];
  q.pt = v.ptype[j];
  return q;
}

// The list in batches of four, in list order, in two phases: the fields of
// all four neighbours are gathered (straight-line code, so every load of the
// batch is in flight together), then f(q, on) runs per neighbour; the next
// batch's indices are loaded before this batch is consumed.  Loads of fields
// a body never reads are dropped by the compiler.  Slots past the count
// repeat the last entry with on = false; bodies add `on ? term : 0` (a
// select: exact, and NaN-safe for the unused slots).
template <typename R, class F>
__device__ inline void for_nbrs4(const SphView<R>& v, int i, bool eps_rule, F&& f) {
  const int cnt = min(v.nbr_cnt[i], v.nbr_cap);
  if (cnt <= 0) return;
  const size_t n = (size_t)v.n;
  uint32_t en[4];
#pragma unroll
  for (int u = 0; u < 4; u++) en[u] = v.nbr[(size_t)min(u, cnt - 1) * n + i];
  for (int k = 0; k < cnt; k += 4) {
    uint32_t e[4];
#pragma unroll
    for (int u = 0; u < 4; u++) e[u] = en[u];
    Nb<R> q[4];
#pragma unroll
    for (int u = 0; u < 4; u++) q[u] = load_nb(v, (int)(e[u] & ~kNbrFlag));
#pragma unroll
    for (int u = 0; u < 4; u++) en[u] = v.nbr[(size_t)min(k + 4 + u, cnt - 1) * n + i];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const bool on = (k + u < cnt) && !(eps_rule && (e[u] & kNbrFlag));
      f(q[u], on);
    }
  }
}

}  // namespace sophia
