// ISPH pressure-Poisson equation on gfx950.
//
// Assembly follows KERNEL_PPE3D_sph (functions_PPE2.cuh:542-698) for the
// rows it computes (fluid, 1 <= p_type < 1000, not an open-boundary buffer,
// i_type <= 2): per neighbour slot k of row i it stores
//   a_ik = m_j (eps_i rho_i + eps_j rho_j) / (rho_i rho_j) * c_ij,
//   c_ij = -(x_ij . grad~W_ij) / r^2
// in FP64 arithmetic, together with d_i = sum_k a_ik (PPE2), b_i (PPE3) and
// sum_k a_ik p_j (PPE4).  The reference then takes one Jacobi sweep
// p_i = (b_i + sum_k a_ik p_j) / d_i (functions_PPE2.cuh:868-878); this
// library also solves the system to a relative residual with BiCGStab:
//   (A p)_i = d_i p_i - sum_{j row} a_ij p_j = b_i + sum_{j not row} a_ij p_j
// whose fixed point is the Jacobi update (non-row columns -- walls and
// buffers -- keep their lagged pressures, as in the reference).
//
// Storage of the operator: ELL over the neighbour list, slot major.
//   FP64 path : a_ik as double                        (8 B per entry)
//   mixed path: a_ik split as hi = fl32(a), lo = fl32(a - hi)
//               inner FP32 solves read hi only        (4 B per entry)
//               the FP64 refinement residual reads hi + lo (48-bit a)
// Reductions are FP64, two-stage and deterministic (per-block partials,
// then one block sums them in a fixed order).
#include "ppe.hpp"

namespace sophia {

namespace {

constexpr int kRedThreads = 256;

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// block sum of up to 3 values into partials[slot * nb + blockIdx.x]
template <int NV>
__device__ inline void block_partials(const double* v, double* partials, int nb, const int* slots) {
  __shared__ double red[NV][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = wave_sum(v[q]);
    if (lane == 0) red[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double s = 0;
      for (int w = 0; w < kRedThreads / 64; w++) s += red[q][w];
      partials[slots[q] * nb + blockIdx.x] = s;
    }
  }
}

__global__ void k_rowmask(DevParams P, int n, const int* ptype, const int* btype, const int* itype, int* row) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int t = ptype[i];
  bool on = itype[i] <= 2 && !(P.open_boundary > 0 && btype[i] > 0) && t < 1000 && t > 0;
  row[i] = on ? 1 : 0;
}

template <typename R>
__global__ void k_assemble(DevParams P, SphView<R> v, PpeView pv, int mixed) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  const int n = v.n;
  const int cnt = min(v.nbr_cnt[i], v.nbr_cap);
  if (!pv.row[i]) {
    // non-row: zero coefficients so every matvec can run over all slots
    for (int k = 0; k < cnt; k++) {
      if (mixed) {
        pv.coef_hi[(size_t)k * n + i] = 0.f;
        pv.coef_lo[(size_t)k * n + i] = 0.f;
      } else {
        pv.coef64[(size_t)k * n + i] = 0.0;
      }
    }
    pv.diag[i] = 0;
    pv.dinv[i] = 0;
    pv.rhs[i] = 0;
    return;
  }
  const double hd = v.h[i];
  const double A = knl_tmpA<double>(P.kernel_type, P.dim, hd);
  const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
  const double uxi = v.ux[i], uyi = v.uy[i], uzi = v.uz[i];
  const double rhoi = v.rho[i], pori = v.por[i];
  double bx = 0, by = 0, bz = 0, Aij = 0, AP = 0, lag = 0;
  for (int k = 0; k < cnt; k++) {
    uint32_t e = v.nbr[(size_t)k * n + i];
    int j = (int)(e & ~kNbrFlag);
    double a = 0;
    bool use = !(e & kNbrFlag) && v.ptype[j] < 1000 && v.itype[j] != 4;
    if (use) {
      double dx = xi - v.x[j], dy = yi - v.y[j], dz = zi - v.z[j];
      double r = sqrt(dx * dx + dy * dy + dz * dz) + 1e-20;
      double dw = knl_dw<double>(P.kernel_type, A, hd, r);
      double gx = dw * dx / r, gy = dw * dy / r, gz = dw * dz / r;
      kgc_apply(P, v.cm, n, i, gx, gy, gz);
      double mj = v.m[j], rhoj = v.rho[j], porj = v.por[j];
      bx += mj / rhoj * (pori * uxi - porj * (double)v.ux[j]) * gx;
      by += mj / rhoj * (pori * uyi - porj * (double)v.uy[j]) * gy;
      bz += mj / rhoj * (pori * uzi - porj * (double)v.uz[j]) * gz;
      double c = -(dx * gx + dy * gy + dz * gz) / r / r;
      a = mj * ((pori * rhoi + porj * rhoj) / (rhoj * rhoi)) * c;
      Aij += a;
      AP += a * v.pres[j];
      if (!pv.row[j]) {
        lag += a * v.pres[j];
        a = 0;  // non-row column: lives in the right-hand side
      }
    }
    if (mixed) {
      float hi = (float)a;
      pv.coef_hi[(size_t)k * n + i] = hi;
      pv.coef_lo[(size_t)k * n + i] = (float)(a - (double)hi);
    } else {
      pv.coef64[(size_t)k * n + i] = a;
    }
  }
  const double tdt = P.dt_sph;
  const double bi = P.rho_fluid * (bx + by + bz + 2 * (double)v.dpor[i] - (double)v.dpor_prev[i]) / tdt;
  v.ppe3[i] = bi;
  v.ppe4[i] = AP;
  v.ppe1[i] = rhoi * (bx + by + bz) / tdt;
  v.ppe2[i] = Aij;
  pv.diag[i] = Aij;
  pv.dinv[i] = Aij != 0 ? 1.0 / Aij : 0.0;
  pv.rhs[i] = bi + lag;
}

template <typename R>
__global__ void k_jacobi(SphView<R> v, PpeView pv) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  if (!pv.row[i]) return;
  v.pres[i] = (v.ppe3[i] + v.ppe4[i]) / v.ppe2[i];
}

// ----- operator application ------------------------------------------------
// y = A x (FP64 coefficients or hi+lo), optional fused dot partials with w
template <int MODE>  // 0: coef64, 1: hi+lo
__global__ __launch_bounds__(kRedThreads) void k_apply64(PpeView pv, const double* __restrict__ x,
                                                         double* __restrict__ y, const double* __restrict__ rhs_or_null,
                                                         double* partials, int nb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double sq[1] = {0.0};
  if (i < pv.n) {
    double acc = 0;
    if (pv.row[i]) {
      const int cnt = min(pv.cnt[i], pv.cap);
      const size_t n = (size_t)pv.n;
      double s = 0;
      for (int k = 0; k < cnt; k++) {
        int j = (int)(pv.nbr[(size_t)k * n + i] & ~kNbrFlag);
        double a = MODE == 0 ? pv.coef64[(size_t)k * n + i]
                             : (double)pv.coef_hi[(size_t)k * n + i] + (double)pv.coef_lo[(size_t)k * n + i];
        s += a * x[j];
      }
      acc = pv.diag[i] * x[i] - s;
      if (rhs_or_null) acc = rhs_or_null[i] - acc;  // residual form
    }
    y[i] = acc;
    sq[0] = acc * acc;
  }
  if (partials) {
    const int slots[1] = {0};
    block_partials<1>(sq, partials, nb, slots);
  }
}

// FP32 inner operator: y = A32 x with hi coefficients, and dot partials
// (w, y) into slot s0 and optionally (y, y) / (y, z) into s1, s2
__global__ __launch_bounds__(kRedThreads) void k_apply32(PpeView pv, const float* __restrict__ x,
                                                         float* __restrict__ y, const float* __restrict__ w,
                                                         const float* __restrict__ z, double* partials, int nb,
                                                         int s0, int s1, int s2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d[3] = {0, 0, 0};
  if (i < pv.n) {
    float acc = 0;
    if (pv.row[i]) {
      const int cnt = min(pv.cnt[i], pv.cap);
      const size_t n = (size_t)pv.n;
      float s = 0;
      for (int k = 0; k < cnt; k++) {
        int j = (int)(pv.nbr[(size_t)k * n + i] & ~kNbrFlag);
        s += pv.coef_hi[(size_t)k * n + i] * x[j];
      }
      acc = (float)pv.diag[i] * x[i] - s;
    }
    y[i] = acc;
    d[0] = (double)w[i] * (double)acc;
    if (z) {
      d[1] = (double)acc * (double)acc;
      d[2] = (double)acc * (double)z[i];
    }
  }
  const int slots[3] = {s0, s1, s2};
  if (z) block_partials<3>(d, partials, nb, slots);
  else block_partials<1>(d, partials, nb, slots);
}

template <typename V>
__global__ __launch_bounds__(kRedThreads) void k_apply_v(PpeView pv, const V* __restrict__ x, V* __restrict__ y,
                                                         const V* __restrict__ w, const V* __restrict__ z,
                                                         double* partials, int nb, int s0, int s1, int s2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d[3] = {0, 0, 0};
  if (i < pv.n) {
    double acc = 0;
    if (pv.row[i]) {
      const int cnt = min(pv.cnt[i], pv.cap);
      const size_t n = (size_t)pv.n;
      double s = 0;
      for (int k = 0; k < cnt; k++) {
        int j = (int)(pv.nbr[(size_t)k * n + i] & ~kNbrFlag);
        s += pv.coef64[(size_t)k * n + i] * (double)x[j];
      }
      acc = pv.diag[i] * (double)x[i] - s;
    }
    y[i] = (V)acc;
    d[0] = (double)w[i] * acc;
    if (z) {
      d[1] = acc * acc;
      d[2] = acc * (double)z[i];
    }
  }
  const int slots[3] = {s0, s1, s2};
  if (z) block_partials<3>(d, partials, nb, slots);
  else block_partials<1>(d, partials, nb, slots);
}

// ----- BiCGStab vector updates (scalars live on the device) -----------------
// scalar slots
enum { S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_BETA, S_RR, S_SS, S_RV, S_TS, S_TT, S_COUNT };
// partial slots
enum { P_A = 0, P_B, P_C, P_COUNT };

__global__ void k_finalize(double* partials, int nb, double* sc, int op) {
  __shared__ double red[P_COUNT][1024 / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  double v[P_COUNT] = {0, 0, 0};
  for (int q = 0; q < P_COUNT; q++)
    for (int b = t; b < nb; b += 1024) v[q] += partials[q * nb + b];
  for (int q = 0; q < P_COUNT; q++) {
    double s = wave_sum(v[q]);
    if (lane == 0) red[q][wave] = s;
  }
  __syncthreads();
  if (t != 0) return;
  double tot[P_COUNT];
  for (int q = 0; q < P_COUNT; q++) {
    double s = 0;
    for (int w = 0; w < 1024 / 64; w++) s += red[q][w];
    tot[q] = s;
  }
  switch (op) {
    case 0:  // start: tot[A] = (r, r) = (rh, r)
      sc[S_RR] = tot[P_A];
      sc[S_RHO] = tot[P_A];
      sc[S_RHO_OLD] = 1;
      sc[S_ALPHA] = 1;
      sc[S_OMEGA] = 1;
      sc[S_BETA] = 0;
      break;
    case 1: {  // after v = A ph: tot[A] = (rh, v)
      double rv = tot[P_A];
      sc[S_RV] = rv;
      sc[S_ALPHA] = rv != 0 ? sc[S_RHO] / rv : 0;
      break;
    }
    case 2:  // after s: tot[A] = (s, s)
      sc[S_SS] = tot[P_A];
      break;
    case 3: {  // after t = A sh: tot[A] = (s, t) [w = s], tot[B] = (t, t)
      sc[S_TS] = tot[P_A];
      sc[S_TT] = tot[P_B];
      sc[S_OMEGA] = tot[P_B] > 0 ? tot[P_A] / tot[P_B] : 0;
      break;
    }
    case 4: {  // after x, r update: tot[A] = (r, r), tot[B] = (rh, r)
      sc[S_RR] = tot[P_A];
      double rho_new = tot[P_B];
      double rho_old = sc[S_RHO];
      double omega = sc[S_OMEGA];
      sc[S_RHO_OLD] = rho_old;
      sc[S_RHO] = rho_new;
      sc[S_BETA] = (rho_old != 0 && omega != 0) ? (rho_new / rho_old) * (sc[S_ALPHA] / omega) : 0;
      break;
    }
    case 5:  // plain norm
      sc[S_RR] = tot[P_A];
      break;
  }
}

template <typename V>
__global__ void k_upd_p(int n, const int* row, const double* dinv, const V* r, V* p, const V* v, V* ph,
                        const double* sc, int first) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double pi;
  if (first) pi = (double)r[i];
  else pi = (double)r[i] + sc[S_BETA] * ((double)p[i] - sc[S_OMEGA] * (double)v[i]);
  p[i] = (V)pi;
  ph[i] = (V)(pi * dinv[i]);
}

template <typename V>
__global__ __launch_bounds__(kRedThreads) void k_upd_s(int n, const double* dinv, const V* r, const V* v, V* s, V* sh,
                                                       const double* sc, double* partials, int nb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d[1] = {0};
  if (i < n) {
    double si = (double)r[i] - sc[S_ALPHA] * (double)v[i];
    s[i] = (V)si;
    sh[i] = (V)(si * dinv[i]);
    d[0] = si * si;
  }
  const int slots[1] = {P_A};
  block_partials<1>(d, partials, nb, slots);
}

template <typename V, typename X>
__global__ __launch_bounds__(kRedThreads) void k_upd_xr(int n, X* x, const V* ph, const V* sh, V* r, const V* s,
                                                        const V* t, const V* rh, const double* sc, double scale,
                                                        double* partials, int nb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d[2] = {0, 0};
  if (i < n) {
    const double a = sc[S_ALPHA], w = sc[S_OMEGA];
    x[i] = (X)((double)x[i] + scale * (a * (double)ph[i] + w * (double)sh[i]));
    double ri = (double)s[i] - w * (double)t[i];
    r[i] = (V)ri;
    d[0] = ri * ri;
    d[1] = (double)rh[i] * ri;
  }
  const int slots[2] = {P_A, P_B};
  block_partials<2>(d, partials, nb, slots);
}

template <typename V>
__global__ __launch_bounds__(kRedThreads) void k_init_r(int n, const int* row, const double* src, double scale, V* r,
                                                        V* rh, V* p, V* v, double* partials, int nb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d[1] = {0};
  if (i < n) {
    double ri = row[i] ? src[i] * scale : 0.0;
    r[i] = (V)ri;
    rh[i] = (V)ri;
    p[i] = 0;
    v[i] = 0;
    d[0] = (double)(V)ri * (double)(V)ri;
  }
  const int slots[1] = {P_A};
  block_partials<1>(d, partials, nb, slots);
}

__global__ void k_gather_rows_guess(int n, const int* row, const double* pres, double* x, int zero) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = (row[i] && !zero) ? pres[i] : 0.0;
}

__global__ void k_scatter_rows(int n, const int* row, const double* x, double* pres) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (row[i]) pres[i] = x[i];
}

template <typename V>
__global__ void k_zero(int n, V* a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ __launch_bounds__(kRedThreads) void k_norm2_rows(int n, const int* row, const double* a, double* partials,
                                                             int nb) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double d[1] = {0};
  if (i < n && row[i]) d[0] = a[i] * a[i];
  const int slots[1] = {P_A};
  block_partials<1>(d, partials, nb, slots);
}

__global__ void k_axpy_rows(int n, const int* row, double* x, const float* d, double scale) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (row[i]) x[i] += scale * (double)d[i];
}

}  // namespace

// ---------------------------------------------------------------------------
int ppe_workspace_alloc(PpeWork& w, int n, int cap, bool mixed) {
  w.n = n;
  w.nb = grid_for(n > 0 ? n : 1, kRedThreads);
  auto A = [&](void** p, size_t bytes) -> hipError_t { return hipMalloc(p, bytes ? bytes : 8); };
  size_t N = (size_t)(n > 0 ? n : 1);
  HIP_TRY(A((void**)&w.row, N * 4));
  HIP_TRY(A((void**)&w.diag, N * 8));
  HIP_TRY(A((void**)&w.dinv, N * 8));
  HIP_TRY(A((void**)&w.rhs, N * 8));
  HIP_TRY(A((void**)&w.x64, N * 8));
  HIP_TRY(A((void**)&w.r64, N * 8));
  HIP_TRY(A((void**)&w.partials, (size_t)P_COUNT * w.nb * 8));
  HIP_TRY(A((void**)&w.sc, S_COUNT * 8));
  HIP_TRY(hipHostMalloc((void**)&w.host_sc, S_COUNT * 8, hipHostMallocDefault));
  // Krylov vectors: 8 of them (r, rh, p, v, ph, s, sh, t), typed per path
  size_t vb = mixed ? 4 : 8;
  for (int k = 0; k < 8; k++) HIP_TRY(A(&w.vec[k], N * vb));
  w.mixed = mixed;
  w.cap = cap;
  HIP_TRY(A(&w.d32, N * 4));
  if (mixed) {
    HIP_TRY(A((void**)&w.coef_hi, N * cap * 4));
    HIP_TRY(A((void**)&w.coef_lo, N * cap * 4));
    w.coef64 = nullptr;
  } else {
    HIP_TRY(A((void**)&w.coef64, N * cap * 8));
    w.coef_hi = w.coef_lo = nullptr;
  }
  return 0;
}

void ppe_workspace_free(PpeWork& w) {
  auto F = [](void* p) {
    if (p) (void)hipFree(p);
  };
  F(w.row); F(w.diag); F(w.dinv); F(w.rhs); F(w.x64); F(w.r64); F(w.partials); F(w.sc);
  if (w.host_sc) (void)hipHostFree(w.host_sc);
  for (int k = 0; k < 8; k++) F(w.vec[k]);
  F(w.coef64); F(w.coef_hi); F(w.coef_lo); F(w.d32);
  w = PpeWork{};
}

PpeView ppe_view(const PpeWork& w, const uint32_t* nbr, const int* cnt, int cap) {
  PpeView pv;
  pv.n = w.n;
  pv.cap = cap;
  pv.nbr = nbr;
  pv.cnt = cnt;
  pv.row = w.row;
  pv.diag = w.diag;
  pv.dinv = w.dinv;
  pv.rhs = w.rhs;
  pv.coef64 = w.coef64;
  pv.coef_hi = w.coef_hi;
  pv.coef_lo = w.coef_lo;
  return pv;
}

template <typename R>
int ppe_assemble(hipStream_t s, const DevParams& P, const SphView<R>& v, PpeWork& w) {
  if (v.n <= 0) return 0;
  PpeView pv = ppe_view(w, v.nbr, v.nbr_cnt, v.nbr_cap);
  k_rowmask<<<grid_for(v.n), kBlock, 0, s>>>(P, v.n, v.ptype, v.btype, v.itype, w.row);
  k_assemble<R><<<grid_for(v.n), kBlock, 0, s>>>(P, v, pv, w.mixed ? 1 : 0);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <typename R>
int ppe_jacobi(hipStream_t s, const SphView<R>& v, PpeWork& w) {
  if (v.n <= 0) return 0;
  PpeView pv = ppe_view(w, v.nbr, v.nbr_cnt, v.nbr_cap);
  k_jacobi<R><<<grid_for(v.n), kBlock, 0, s>>>(v, pv);
  HIP_TRY(hipGetLastError());
  return 0;
}

// y = A x (FP64 operator of either path); residual form y = rhs - A x when
// residual != 0; returns ||y||^2 when norm2 != nullptr (synchronises)
int ppe_apply64(hipStream_t s, PpeWork& w, const uint32_t* nbr, const int* cnt, int cap, const double* x, double* y,
                bool residual, double* norm2) {
  PpeView pv = ppe_view(w, nbr, cnt, cap);
  int n = w.n;
  if (n <= 0) {
    if (norm2) *norm2 = 0;
    return 0;
  }
  double* part = norm2 ? w.partials : nullptr;
  if (w.mixed)
    k_apply64<1><<<w.nb, kRedThreads, 0, s>>>(pv, x, y, residual ? w.rhs : nullptr, part, w.nb);
  else
    k_apply64<0><<<w.nb, kRedThreads, 0, s>>>(pv, x, y, residual ? w.rhs : nullptr, part, w.nb);
  HIP_TRY(hipGetLastError());
  if (norm2) {
    k_finalize<<<1, 1024, 0, s>>>(w.partials, w.nb, w.sc, 5);
    HIP_TRY(hipMemcpyAsync(w.host_sc, w.sc, S_COUNT * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    *norm2 = w.host_sc[S_RR];
  }
  return 0;
}

// BiCGStab on A z = src*scale (rows only), V = vector type.  z is
// accumulated into x (type X) with factor xscale.  Stops when
// ||r|| <= tol_abs or after max_iter iterations.  Returns iterations.
template <typename V, typename X>
static int bicgstab(hipStream_t s, PpeWork& w, const uint32_t* nbr, const int* cnt, int cap, const double* src,
                    double scale, X* x, double xscale, double tol_abs, int max_iter, int* iters, double* rr_out,
                    Timers* tm) {
  const char* mv_name = sizeof(V) == 4 ? "ppe_matvec_fp32" : "ppe_matvec_fp64";
  const int n = w.n, nb = w.nb, g = grid_for(n);
  PpeView pv = ppe_view(w, nbr, cnt, cap);
  V* r = (V*)w.vec[0];
  V* rh = (V*)w.vec[1];
  V* p = (V*)w.vec[2];
  V* v = (V*)w.vec[3];
  V* ph = (V*)w.vec[4];
  V* sv = (V*)w.vec[5];
  V* sh = (V*)w.vec[6];
  V* t = (V*)w.vec[7];
  k_init_r<V><<<nb, kRedThreads, 0, s>>>(n, w.row, src, scale, r, rh, p, v, w.partials, nb);
  k_finalize<<<1, 1024, 0, s>>>(w.partials, nb, w.sc, 0);
  HIP_TRY(hipMemcpyAsync(w.host_sc, w.sc, S_COUNT * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  double rr = w.host_sc[S_RR];
  int it = 0;
  const bool f32 = sizeof(V) == 4;
  while (rr > tol_abs * tol_abs && it < max_iter) {
    k_upd_p<V><<<g, kBlock, 0, s>>>(n, w.row, w.dinv, r, p, v, ph, w.sc, it == 0);
    {
      TimeScope ts(tm, mv_name);
      if (f32)
        k_apply32<<<nb, kRedThreads, 0, s>>>(pv, (const float*)ph, (float*)v, (const float*)rh, nullptr,
                                             w.partials, nb, P_A, P_B, P_C);
      else
        k_apply_v<V><<<nb, kRedThreads, 0, s>>>(pv, ph, v, rh, (const V*)nullptr, w.partials, nb, P_A, P_B, P_C);
    }
    k_finalize<<<1, 1024, 0, s>>>(w.partials, nb, w.sc, 1);
    k_upd_s<V><<<nb, kRedThreads, 0, s>>>(n, w.dinv, r, v, sv, sh, w.sc, w.partials, nb);
    k_finalize<<<1, 1024, 0, s>>>(w.partials, nb, w.sc, 2);
    // t = A sh; partial (s, t) with w = s, (t, t), unused third
    {
      TimeScope ts(tm, mv_name);
      if (f32)
        k_apply32<<<nb, kRedThreads, 0, s>>>(pv, (const float*)sh, (float*)t, (const float*)sv, (const float*)sv,
                                             w.partials, nb, P_A, P_B, P_C);
      else
        k_apply_v<V><<<nb, kRedThreads, 0, s>>>(pv, sh, t, sv, sv, w.partials, nb, P_A, P_B, P_C);
    }
    k_finalize<<<1, 1024, 0, s>>>(w.partials, nb, w.sc, 3);
    k_upd_xr<V, X><<<nb, kRedThreads, 0, s>>>(n, x, ph, sh, r, sv, t, rh, w.sc, xscale, w.partials, nb);
    k_finalize<<<1, 1024, 0, s>>>(w.partials, nb, w.sc, 4);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(w.host_sc, w.sc, S_COUNT * 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    rr = w.host_sc[S_RR];
    it++;
    if (!(rr == rr)) {
      set_error("BiCGStab breakdown (NaN residual) at iteration %d", it);
      *iters = it;
      return SOPHIA_E_NUMERIC;
    }
    if (w.host_sc[S_RHO] == 0 || w.host_sc[S_OMEGA] == 0) break;
  }
  *iters = it;
  if (rr_out) *rr_out = rr;
  return 0;
}

int ppe_solve(hipStream_t s, PpeWork& w, const uint32_t* nbr, const int* cnt, int cap, double* pres,
              const sophia_ppe_opts& o, sophia_ppe_stats& st, const double* guess_or_null, Timers* tm) {
  const int n = w.n;
  if (n <= 0) return 0;
  const int g = grid_for(n);
  // guess
  if (guess_or_null) {
    HIP_TRY(hipMemcpyAsync(w.x64, guess_or_null, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  } else {
    k_gather_rows_guess<<<g, kBlock, 0, s>>>(n, w.row, pres, w.x64, o.guess == SOPHIA_GUESS_ZERO);
  }
  double bn = 0;
  {
    int rc0 = ppe_rhs_norm(s, w, &bn);
    if (rc0) return rc0;
  }
  if (bn == 0) bn = 1;
  double rr = 0;
  int rc = ppe_apply64(s, w, nbr, cnt, cap, w.x64, w.r64, true, &rr);
  if (rc) return rc;
  st.r0_rel = sqrt(rr) / bn;
  int total = 0, outer = 0;
  const double tol = o.rtol * bn;
  if (!w.mixed) {
    int it = 0;
    double rr_end = 0;
    if (sqrt(rr) > tol) {
      rc = bicgstab<double, double>(s, w, nbr, cnt, cap, w.r64, 1.0, w.x64, 1.0, tol, o.max_iter, &it, &rr_end,
                                    tm);
      if (rc) return rc;
    }
    total = it;
    outer = 1;
  } else {
    const int max_outer = o.max_outer > 0 ? o.max_outer : 20;
    const double inner = o.inner_rtol > 0 ? o.inner_rtol : 1e-4;
    while (sqrt(rr) > tol && outer < max_outer && total < o.max_iter) {
      double rn = sqrt(rr);
      int it = 0;
      // solve A d = r / ||r|| in FP32 to the inner tolerance; x += ||r|| d
      // (the update is applied inside the solver loop through xscale)
      float* dvec = (float*)w.d32;
      k_zero<float><<<g, kBlock, 0, s>>>(n, dvec);
      double rr_in = 0;
      rc = bicgstab<float, float>(s, w, nbr, cnt, cap, w.r64, 1.0 / rn, dvec, 1.0, inner,
                                  o.max_iter - total, &it, &rr_in, tm);
      if (rc) return rc;
      k_axpy_rows<<<g, kBlock, 0, s>>>(n, w.row, w.x64, dvec, rn);
      total += it;
      outer++;
      rc = ppe_apply64(s, w, nbr, cnt, cap, w.x64, w.r64, true, &rr);
      if (rc) return rc;
      if (it == 0) break;
    }
  }
  // final FP64 residual
  rc = ppe_apply64(s, w, nbr, cnt, cap, w.x64, w.r64, true, &rr);
  if (rc) return rc;
  st.r_rel = sqrt(rr) / bn;
  st.iterations = total;
  st.outer = outer;
  st.converged = st.r_rel <= o.rtol * (1 + 1e-12);
  k_scatter_rows<<<g, kBlock, 0, s>>>(n, w.row, w.x64, pres);
  HIP_TRY(hipGetLastError());
  return 0;
}

int ppe_rhs_norm(hipStream_t s, PpeWork& w, double* bn) {
  if (w.n <= 0) {
    *bn = 0;
    return 0;
  }
  k_norm2_rows<<<w.nb, kRedThreads, 0, s>>>(w.n, w.row, w.rhs, w.partials, w.nb);
  k_finalize<<<1, 1024, 0, s>>>(w.partials, w.nb, w.sc, 5);
  HIP_TRY(hipMemcpyAsync(w.host_sc, w.sc, S_COUNT * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  *bn = sqrt(w.host_sc[S_RR]);
  return 0;
}

int ppe_residual_of(hipStream_t s, PpeWork& w, const uint32_t* nbr, const int* cnt, int cap, const double* x,
                    double* rel) {
  double bn = 0, rr = 0;
  int rc = ppe_rhs_norm(s, w, &bn);
  if (rc) return rc;
  rc = ppe_apply64(s, w, nbr, cnt, cap, x, w.r64, true, &rr);
  if (rc) return rc;
  *rel = sqrt(rr) / (bn == 0 ? 1 : bn);
  return 0;
}

template int ppe_assemble<float>(hipStream_t, const DevParams&, const SphView<float>&, PpeWork&);
template int ppe_assemble<double>(hipStream_t, const DevParams&, const SphView<double>&, PpeWork&);
template int ppe_jacobi<float>(hipStream_t, const SphView<float>&, PpeWork&);
template int ppe_jacobi<double>(hipStream_t, const SphView<double>&, PpeWork&);

}  // namespace sophia
