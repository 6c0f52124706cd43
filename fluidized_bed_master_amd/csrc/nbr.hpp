// Batched neighbour-list traversal shared by the pair kernels.
#pragma once

#include <type_traits>

#include "common.hpp"

namespace sophia {

// Fields of one neighbour j of particle i, gathered before any of them is
// used; dx, dy, dz = x_i - x_j.
template <typename R>
struct Nb {
  double dx, dy, dz, m, rho, pres, h;
  R ux, uy, uz, temp, grx, gry, grz;
  double rref, por;
  int pt;
  int j;  // the neighbour's index, for fields a few kernels read on condition
};

// Particle i's side of the differences: FP64 position, or its pos4 record.
struct NbOrigin {
  double x, y, z;
  float4 p4;
};

// REL (FP32 fields only): the differences come from the pos4 records (one
// 16 B gather, error below 2^-21 of the grid spacing) instead of three FP64
// gathers, for stages that use them in FP32.  Otherwise FP64 differences.
template <bool REL, typename R>
__device__ inline Nb<R> load_nb(const SphView<R>& v, int j, const NbOrigin& o) {
  Nb<R> q;
  if constexpr (REL && std::is_same<R, float>::value) {
    const float3 d = rel_diff(o.p4, v.pos4[j], v.pgrid, v.pmask);
    q.dx = d.x; q.dy = d.y; q.dz = d.z;
  } else {
    pdiff(v, o.x, o.y, o.z, j, q.dx, q.dy, q.dz);
  }
  q.m = v.m[j]; q.rho = v.rho[j]; q.pres = v.pres[j]; q.h = v.h[j];
  q.ux = v.ux[j]; q.uy = v.uy[j]; q.uz = v.uz[j]; q.temp = v.temp[j]; q.rref = v.rho_ref[j];
  q.grx = v.grx[j]; q.gry = v.gry[j]; q.grz = v.grz[j]; q.por = v.por[j];
  q.pt = v.ptype[j];
  q.j = j;
  return q;
}

// The list in batches of four, in list order, in two phases: the fields of
// all four neighbours are gathered (straight-line code, so every load of the
// batch is in flight together), then f(q, on) runs per neighbour; the next
// batch's indices are loaded before this batch is consumed.  Loads of fields
// a body never reads are dropped by the compiler.  Slots past the count
// repeat the last entry with on = false; bodies add `on ? term : 0` (a
// select: exact, and NaN-safe for the unused slots).
template <bool REL = false, typename R, class F>
__device__ inline void for_nbrs4(const SphView<R>& v, int i, bool eps_rule, F&& f) {
  const int cnt = min(v.nbr_cnt[i], v.nbr_cap);
  if (cnt <= 0) return;
  const size_t n = (size_t)v.n;
  NbOrigin o;
  if constexpr (REL && std::is_same<R, float>::value) {
    o.p4 = v.pos4[i];
  } else {
    o.x = v.x[i]; o.y = v.y[i]; o.z = v.z[i];
  }
  const int last4 = (cnt - 1) & ~3;
  uint4 en4 = *(const uint4*)(v.nbr + nbr_slot(0, i, (int)n));
  for (int k = 0; k < cnt; k += 4) {
    const uint32_t en[4] = {en4.x, en4.y, en4.z, en4.w};
    uint32_t e[4];
#pragma unroll
    for (int u = 0; u < 4; u++) e[u] = en[u];
    Nb<R> q[4];
#pragma unroll
    for (int u = 0; u < 4; u++) q[u] = load_nb<REL>(v, (int)(e[u] & ~kNbrFlag), o);
    en4 = *(const uint4*)(v.nbr + nbr_slot(min(k + 4, last4), i, (int)n));
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const bool on = (k + u < cnt) && !(eps_rule && (e[u] & kNbrFlag));
      f(q[u], on);
    }
  }
}

// One neighbour from the packed records: the members of Nb that the
// FP64-geometry pair sums read (h and the velocities only with VEL).
template <typename R>
struct NbP {
  double dx, dy, dz, m, rho, pres;
  R ux, uy, uz, temp, h;
  int pt;
  int j;
  double S;  // flux-wall sum (mixed path with VEL; otherwise flux_s[j])
};

// The list in batches of four like for_nbrs4, reading the packed records
// pk = {x, y, z, m} and pk2 = {rho, p}, then with VEL pkv = {ux, uy, uz, T}
// and pkh = {h, p_type}, else p_type: four or three gathers per neighbour in
// place of twelve or seven.  With VEL on the mixed path pk, pkb = {rho, p,
// u, T} and pkc = {S, h, p_type}: three.  dx = x_i - x_j (minimum image on periodic axes),
// the same bits as pdiff.
template <bool VEL, typename R, class F>
__device__ inline void for_nbrs4_pk(const SphView<R>& v, int i, bool eps_rule, F&& f) {
  const int cnt = min(v.nbr_cnt[i], v.nbr_cap);
  if (cnt <= 0) return;
  const int n = v.n;
  const double4 pi = v.pk[i];
  const int last4 = (cnt - 1) & ~3;
  uint4 en4 = *(const uint4*)(v.nbr + nbr_slot(0, i, n));
  for (int k = 0; k < cnt; k += 4) {
    const uint32_t en[4] = {en4.x, en4.y, en4.z, en4.w};
    NbP<R> q[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int j = (int)(en[u] & ~kNbrFlag);
      const double4 a = v.pk[j];
      q[u].dx = pi.x - a.x;
      q[u].dy = pi.y - a.y;
      q[u].dz = pi.z - a.z;
      q[u].m = a.w;
      if constexpr (VEL && std::is_same<R, float>::value) {
        // mixed path: two more records, {rho, p, u, T} and {S, h, p_type}
        const PkB b = v.pkb[j];
        const PkC c = v.pkc[j];
        q[u].rho = b.rho; q[u].pres = b.p;
        q[u].ux = b.ux; q[u].uy = b.uy; q[u].uz = b.uz; q[u].temp = b.temp;
        q[u].S = c.S; q[u].h = c.h; q[u].pt = c.pt;
        q[u].j = j;
        continue;
      }
      const double2 b = v.pk2[j];
      q[u].rho = b.x;
      q[u].pres = b.y;
      if constexpr (VEL) {
        const auto w = v.pkv[j];
        const HRec<R> hr = v.pkh[j];
        q[u].ux = w.x; q[u].uy = w.y; q[u].uz = w.z; q[u].temp = w.w;
        q[u].h = hr.h;
        q[u].pt = hr.pt;
      } else {
        q[u].pt = v.ptype[j];
      }
      q[u].j = j;
    }
    en4 = *(const uint4*)(v.nbr + nbr_slot(min(k + 4, last4), i, n));
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const bool on = (k + u < cnt) && !(eps_rule && (en[u] & kNbrFlag));
      if (v.pmask) {
        if (v.pmask & 1) q[u].dx = per_wrap(q[u].dx, v.plx);
        if (v.pmask & 2) q[u].dy = per_wrap(q[u].dy, v.ply);
        if (v.pmask & 4) q[u].dz = per_wrap(q[u].dz, v.plz);
      }
      f(q[u], on);
    }
  }
}

// One neighbour from the 32 B records (FP32 pair kernels): the members of Nb
// those kernels read, the differences from the position records.
struct NbR {
  float dx, dy, dz, m, rho, aux;
  float ux, uy, uz, temp;
  int pt;
  int j;
  bool flag;  // the entry's eps-rule flag (kNbrFlag), for kernels that apply it themselves
};

// The list in batches of four like for_nbrs4<true>, reading r32 (and with
// VEL pkv = {ux, uy, uz, T}): one or two gathers per neighbour in place of
// four to eight, the same values.
template <bool VEL, class F>
__device__ inline void for_nbrs4_r32(const SphView<float>& v, int i, bool eps_rule, F&& f) {
  const int cnt = min(v.nbr_cnt[i], v.nbr_cap);
  if (cnt <= 0) return;
  const int n = v.n;
  const float4 pi4 = v.pos4[i];
  const int last4 = (cnt - 1) & ~3;
  uint4 en4 = *(const uint4*)(v.nbr + nbr_slot(0, i, n));
  for (int k = 0; k < cnt; k += 4) {
    const uint32_t en[4] = {en4.x, en4.y, en4.z, en4.w};
    NbR q[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int j = (int)(en[u] & ~kNbrFlag);
      const Rec32 a = v.r32[j];
      const float3 d = rel_diff(pi4, a.p, v.pgrid, v.pmask);
      q[u].dx = d.x; q[u].dy = d.y; q[u].dz = d.z;
      q[u].m = a.m; q[u].rho = a.rho; q[u].aux = a.aux; q[u].pt = a.pt;
      if constexpr (VEL) {
        const float4 w = v.pkv[j];
        q[u].ux = w.x; q[u].uy = w.y; q[u].uz = w.z; q[u].temp = w.w;
      }
      q[u].j = j;
      q[u].flag = (en[u] & kNbrFlag) != 0;
    }
    en4 = *(const uint4*)(v.nbr + nbr_slot(min(k + 4, last4), i, n));
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const bool on = (k + u < cnt) && !(eps_rule && (en[u] & kNbrFlag));
      f(q[u], on);
    }
  }
}

}  // namespace sophia
