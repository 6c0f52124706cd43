// Two-level additive preconditioner of the PPE Krylov solve.
//
// Coarse space: aggregates of kAgg^3 cells of the row-major cell grid.  In the
// cell-sorted particle order an aggregate is kAgg^2 contiguous particle
// ranges (one per (y, z) cell row), so restriction is a sum over contiguous
// ranges (one wave per aggregate, fixed order) and the coarse operator
// E = Z^T A Z is accumulated per aggregate from the stored rows of its
// particles (one workgroup per aggregate, per-thread LDS bins for the 27
// neighbouring aggregates, summed in thread order -- deterministic).
//
// Coarse solve: dense X ~ E^-1 (row-major FP32), applied as one GEMV per
// preconditioner application (one wave per coarse row).  X is refreshed with
// Newton-Schulz steps X <- X (2 I - E X) on rocBLAS SGEMM (plain dense GEMMs;
// atomics off, so results are reproducible); from one time step to the next
// E moves little and one GEMM that measures ||I - E X|| usually accepts the
// previous X.  The preconditioner only steers the Krylov iterations: the
// solve still stops on the true FP64 residual.
#include "coarse.hpp"

#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

namespace sophia {

namespace {

constexpr int kResBlocks = 512;

struct AggGrid {
  int NI, NJ, NK, AX, AY, AZ;
};

__device__ inline void cell_xyz(uint32_t key, int NI, int NJ, int& cx, int& cy, int& cz) {
  cx = (int)(key % (uint32_t)NI);
  const int t = (int)(key / (uint32_t)NI);
  cy = t % NJ;
  cz = t / NJ;
}

__device__ inline int agg_of(uint32_t key, const AggGrid& g) {
  int cx, cy, cz;
  cell_xyz(key, g.NI, g.NJ, cx, cy, cz);
  return cx / kAgg + g.AX * (cy / kAgg + g.AY * (cz / kAgg));
}

// Column indices of the four entries of quad q of row i (kFmtRun8: run bases
// from the row's cell key and run lengths, as in the matvec decoder; entries
// past cnt get the row's own index and a zero coefficient from the caller).
struct RunDecode {
  int B[9], E[8];
};

__device__ inline void run_decode(const PpeOp& o, int i, int NI, int NJ, RunDecode& d) {
  int cx, cy, cz;
  cell_xyz(o.keys[i], NI, NJ, cx, cy, cz);
  const int xl = cx > 0 ? cx - 1 : 0;
  const uint64_t rc = o.runc[i];
  int acc = 0;
#pragma unroll
  for (int r = 0; r < 9; r++) {
    const int b = cy + (r % 3) - 1, c = cz + (r / 3) - 1;
    const int len = (int)((rc >> (kRunBits * r)) & kRunMax);
    d.B[r] = len ? (int)o.cell_off[xl + NI * (b + NJ * c)] : 0;
    acc += len;
    if (r < 8) d.E[r] = acc;
  }
}

// E = Z^T A Z in two passes.
//
// Pass 1, one thread per particle: a row i in cell (cx, cy, cz) only couples
// to aggregates that differ from its own by at most one step per axis, and
// per axis only in one direction (-1 if cx % kAgg == 0, +1 if it is
// kAgg - 1, none in between).  So its row of A restricted to aggregates has
// at most 8 non-zero sums, slot s = sx + 2 sy + 4 sz with sx = "column is in
// the neighbouring aggregate along x".  Slot 0 holds d_i minus the row's own
// aggregate columns.  Non-rows write zeros.
template <int FMT, int CM>
__global__ __launch_bounds__(256) void k_coarse_rows(PpeOp o, AggGrid g, double* __restrict__ slots) {
  const int i = particle_index();
  const int n = o.n;
  if (i >= n) return;
  double acc[8];
#pragma unroll
  for (int t = 0; t < 8; t++) acc[t] = 0.0;
  if (o.row[i]) {
    int cx, cy, cz;
    cell_xyz(o.keys[i], g.NI, g.NJ, cx, cy, cz);
    const int ax = cx / kAgg, ay = cy / kAgg, az = cz / kAgg;
    acc[0] = o.diag[i];
    const int cnt = o.ccnt[i];
    RunDecode d;
    if (FMT == kFmtRun8 && cnt > 0) run_decode(o, i, g.NI, g.NJ, d);
    const int nq = (cnt + 3) >> 2;
    for (int q = 0; q < nq; q++) {
      const size_t qi = (size_t)q * n + (size_t)i;
      double c[4];
      uint32_t hb[4] = {0u, 0u, 0u, 0u};  // mixed: hi words, kFmtRun8 offset in the low byte
      if (CM == 2) {
        const double2 a = reinterpret_cast<const double2*>(o.c64)[2 * qi];
        const double2 b = reinterpret_cast<const double2*>(o.c64)[2 * qi + 1];
        c[0] = a.x; c[1] = a.y; c[2] = b.x; c[3] = b.y;
      } else {
        const uint4 a = reinterpret_cast<const uint4*>(o.hi)[qi];
        const float4 b = reinterpret_cast<const float4*>(o.lo)[qi];
        hb[0] = a.x; hb[1] = a.y; hb[2] = a.z; hb[3] = a.w;
        c[0] = (double)hi_val(hb[0]) + (double)b.x; c[1] = (double)hi_val(hb[1]) + (double)b.y;
        c[2] = (double)hi_val(hb[2]) + (double)b.z; c[3] = (double)hi_val(hb[3]) + (double)b.w;
      }
      int j[4];
      if (FMT == kFmtRun8) {
        const uint32_t w = CM == 2 ? o.idx[qi] : 0u;
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int k = 4 * q + u;
          int base = d.B[0];
#pragma unroll
          for (int r = 0; r < 8; r++) base = k >= d.E[r] ? d.B[r + 1] : base;
          const uint32_t off = CM == 2 ? (w >> (8 * u)) & 255u : hb[u] & 255u;
          j[u] = k < cnt ? base + (int)off : i;
        }
      } else {
        const uint4 w = reinterpret_cast<const uint4*>(o.idx)[qi];
        j[0] = (int)w.x; j[1] = 4 * q + 1 < cnt ? (int)w.y : i;
        j[2] = 4 * q + 2 < cnt ? (int)w.z : i; j[3] = 4 * q + 3 < cnt ? (int)w.w : i;
      }
      uint32_t kj[4];
#pragma unroll
      for (int u = 0; u < 4; u++) kj[u] = o.keys[j[u]];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        int cxj, cyj, czj;
        cell_xyz(kj[u], g.NI, g.NJ, cxj, cyj, czj);
        const int sl = (cxj / kAgg != ax) + 2 * (cyj / kAgg != ay) + 4 * (czj / kAgg != az);
        const double a = 4 * q + u < cnt ? c[u] : 0.0;
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] -= sl == t ? a : 0.0;
      }
    }
  }
#pragma unroll
  for (int t = 0; t < 8; t++) slots[(size_t)t * n + i] = acc[t];
}

// Pass 2, one wave per aggregate I: for each of its cells (fixed order) the
// lanes sum the 8 slots over the cell's particles, and each slot sum goes to
// the aggregate that the slot names for that cell's position.  Each of the
// 27 E entries of row I is then a fixed-order sum -- deterministic.  The 27
// entries around the diagonal are all rewritten; the rest of E stays zero
// from allocation.  An aggregate without rows is decoupled (E_II = 1; its
// restriction is always 0).
__global__ __launch_bounds__(64) void k_coarse_reduce(PpeOp o, AggGrid g, const double* __restrict__ slots,
                                                      float* __restrict__ E, int ld) {
  __shared__ double bins[27];
  __shared__ int nrow;
  const int I = blockIdx.x;
  const int lane = threadIdx.x;
  const int ax = I % g.AX, ay = (I / g.AX) % g.AY, az = I / (g.AX * g.AY);
  if (lane < 27) bins[lane] = 0.0;
  if (lane == 0) nrow = 0;
  __syncthreads();
  const size_t n = (size_t)o.n;
  int rows = 0;
  for (int cz = az * kAgg; cz < min(az * kAgg + kAgg, g.NK); cz++) {
    const int dz = cz % kAgg == 0 ? -1 : (cz % kAgg == kAgg - 1 ? 1 : 0);
    for (int cy = ay * kAgg; cy < min(ay * kAgg + kAgg, g.NJ); cy++) {
      const int dy = cy % kAgg == 0 ? -1 : (cy % kAgg == kAgg - 1 ? 1 : 0);
      for (int cx = ax * kAgg; cx < min(ax * kAgg + kAgg, g.NI); cx++) {
        const int dx = cx % kAgg == 0 ? -1 : (cx % kAgg == kAgg - 1 ? 1 : 0);
        const int key = cx + g.NI * (cy + g.NJ * cz);
        const int p0 = (int)o.cell_off[key], p1 = (int)o.cell_off[key + 1];
        double sv[8];
#pragma unroll
        for (int t = 0; t < 8; t++) sv[t] = 0.0;
        for (int i = p0 + lane; i < p1; i += 64) {
#pragma unroll
          for (int t = 0; t < 8; t++) sv[t] += slots[(size_t)t * n + i];
          rows += o.row[i] ? 1 : 0;
        }
#pragma unroll
        for (int t = 0; t < 8; t++) {
          double v = sv[t];
          for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
          sv[t] = v;
        }
        if (lane == 0) {
#pragma unroll
          for (int t = 0; t < 8; t++) {
            const int ox = (t & 1) ? dx : 0, oy = (t & 2) ? dy : 0, oz = (t & 4) ? dz : 0;
            bins[(ox + 1) + 3 * (oy + 1) + 9 * (oz + 1)] += sv[t];
          }
        }
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) rows += __shfl_xor(rows, off);
  if (lane == 0) nrow = rows;
  __syncthreads();
  if (lane < 27) {
    const int jx = ax + lane % 3 - 1, jy = ay + (lane / 3) % 3 - 1, jz = az + lane / 9 - 1;
    if (jx >= 0 && jx < g.AX && jy >= 0 && jy < g.AY && jz >= 0 && jz < g.AZ) {
      const int J = jx + g.AX * (jy + g.AY * jz);
      E[(size_t)I * ld + J] = nrow == 0 ? (J == I ? 1.0f : 0.0f) : (float)bins[lane];
    }
  }
}

// Dense matrices are row-major na x na inside an ld x ld buffer (ld a
// multiple of 64, so every row starts 256 B aligned).

// X = diag(E)^-1 (cold start)
__global__ void k_diag_inverse(const float* __restrict__ E, float* __restrict__ X, int na, int ld) {
  const size_t m = (size_t)na * ld;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (size_t)gridDim.x * blockDim.x) {
    const size_t I = e / ld, J = e % ld;
    const float d = E[I * ld + I];
    X[e] = (I == J && d != 0.0f) ? 1.0f / d : 0.0f;
  }
}

__global__ void k_identity(float* __restrict__ U, int na, int ld) {
  const size_t m = (size_t)na * ld;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (size_t)gridDim.x * blockDim.x)
    U[e] = (e / ld == e % ld) ? 1.0f : 0.0f;
}

// partials of ||U - I||_F^2 over the na x na block (U = 2 I - E X, so
// U - I = I - E X)
__global__ __launch_bounds__(256) void k_res_partials(const float* __restrict__ U, int na, int ld,
                                                      double* partials) {
  __shared__ double red[4];
  const size_t m = (size_t)na * ld;
  double s = 0;
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (size_t)gridDim.x * blockDim.x) {
    const size_t I = e / ld, J = e % ld;
    if (J >= (size_t)na) continue;
    const double r = (double)U[e] - (I == J ? 1.0 : 0.0);
    s += r * r;
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partials[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// The BiCGStab vector update that precedes each coarse application, done in
// the restriction's pass over the particles (the update is elementwise, the
// restriction visits every particle of the grid once):
//   KIND 1, direction: p = r + beta (p - omega v)  (p = r on the first
//           iteration), out = p / D;
//   KIND 2, s:         r = r - alpha v,                     out = r / D;
// each in FP64 and stored as V, and c[I] = the sum over aggregate I of the
// updated vector as stored (the vectors are zero off the rows).
// One 256-thread block per aggregate: every wave fetches the particle ranges
// of the aggregate's kAgg^2 cell rows (lanes 0..8), the block walks the
// concatenated ranges with independent loads and sums in a fixed order (lane
// butterflies, then waves 0..3).  Block na updates the particles past the last
// cell, which belong to no aggregate.
template <typename V, int KIND>
__device__ inline double update_one(const CoarseUpd<V>& u, int i, bool first, double alpha, double beta,
                                    double omega) {
  double val;
  V* dst;
  if constexpr (KIND == 1) {
    val = first ? (double)u.r[i] : (double)u.r[i] + beta * ((double)u.p[i] - omega * (double)u.v[i]);
    dst = u.p;
  } else {
    val = (double)u.r[i] - alpha * (double)u.v[i];
    dst = u.r;
  }
  const V st = (V)val;
  dst[i] = st;
  u.out[i] = (V)(val * u.dinv[i]);
  return (double)st;
}

template <typename V, int KIND>
__global__ __launch_bounds__(256) void k_update_restrict(PpeOp o, AggGrid g, int na, CoarseUpd<V> u,
                                                         double* __restrict__ c, const int* done, const int* brk) {
  if (*done || *brk) return;
  const bool first = KIND == 1 && *u.first != 0;
  const double alpha = KIND == 2 ? *u.alpha : 0.0;
  const double beta = KIND == 1 ? *u.beta : 0.0, omega = KIND == 1 ? *u.omega : 0.0;
  const int I = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63;
  if (I == na) {
    for (int i = (int)o.cell_off[g.NI * g.NJ * g.NK] + t; i < o.n; i += 256)
      update_one<V, KIND>(u, i, first, alpha, beta, omega);
    return;
  }
  const int ax = I % g.AX, ay = (I / g.AX) % g.AY, az = I / (g.AX * g.AY);
  const int x0 = ax * kAgg, x1 = min(x0 + kAgg, g.NI);
  int p0 = 0, len = 0;
  if (lane < kAgg * kAgg) {
    const int cy = ay * kAgg + lane % kAgg, cz = az * kAgg + lane / kAgg;
    if (cy < g.NJ && cz < g.NK) {
      const int kr = g.NI * (cy + g.NJ * cz);
      p0 = (int)o.cell_off[kr + x0];
      len = (int)o.cell_off[kr + x1] - p0;
    }
  }
  int start[kAgg * kAgg], first_i[kAgg * kAgg];
  int tot = 0;
#pragma unroll
  for (int r = 0; r < kAgg * kAgg; r++) {
    start[r] = tot;
    first_i[r] = __shfl(p0, r);
    tot += __shfl(len, r);
  }
  double s = 0;
#pragma unroll 2
  for (int e = t; e < tot; e += 256) {
    int i = first_i[0] + e;
#pragma unroll
    for (int r = 1; r < kAgg * kAgg; r++) i = e >= start[r] ? first_i[r] + (e - start[r]) : i;
    s += update_one<V, KIND>(u, i, first, alpha, beta, omega);
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  __shared__ double red[4];
  if (lane == 0) red[t >> 6] = s;
  __syncthreads();
  if (t == 0) c[I] = ((red[0] + red[1]) + red[2]) + red[3];
}

// y = X c: one wave per coarse row, 16 B loads, four in flight per lane
__global__ __launch_bounds__(256) void k_coarse_gemv(const float* __restrict__ X, int na, int ld,
                                                     const double* __restrict__ c, double* __restrict__ y,
                                                     const int* done, const int* brk) {
  if (*done || *brk) return;
  const int I = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (I >= na) return;
  const float4* xr = reinterpret_cast<const float4*>(X + (size_t)I * ld);
  const int nv = ld >> 2;  // float4 per row (padding columns of X are zero)
  double s = 0;
#pragma unroll 4
  for (int q = lane; q < nv; q += 64) {
    const float4 a = xr[q];
    const int J = 4 * q;
    s += (double)a.x * c[J] + (double)a.y * c[J + 1] + (double)a.z * c[J + 2] + (double)a.w * c[J + 3];
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) y[I] = s;
}

// out_i += y[agg(i)] on rows
template <typename V>
__global__ __launch_bounds__(256) void k_prolong_add(PpeOp o, AggGrid g, const double* __restrict__ y,
                                                     V* __restrict__ out, const int* done, const int* brk) {
  if (*done || *brk) return;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= o.n || !o.row[i]) return;
  out[i] = (V)((double)out[i] + y[agg_of(o.keys[i], g)]);
}

AggGrid agg_grid(const CoarseWork& c) { return AggGrid{c.NI, c.NJ, c.NK, c.AX, c.AY, c.AZ}; }

#define BLAS_TRY(x)                                                  \
  do {                                                               \
    rocblas_status st_ = (x);                                        \
    if (st_ != rocblas_status_success) {                             \
      set_error("rocBLAS %s failed: status %d", #x, (int)st_);       \
      return SOPHIA_E_HIP;                                           \
    }                                                                \
  } while (0)

int alloc_coarse(CoarseWork& c, const PpeWork& w) {
  const int NI = w.NI, NJ = w.NJ, NK = w.ncell / (w.NI * w.NJ);
  const int AX = (NI + kAgg - 1) / kAgg, AY = (NJ + kAgg - 1) / kAgg, AZ = (NK + kAgg - 1) / kAgg;
  const int na = AX * AY * AZ;
  const int ld = (na + 63) / 64 * 64;
  if (c.na == na && c.NI == NI && c.NJ == NJ && c.NK == NK && c.slots_n >= w.n) return 0;
  if (na > kCoarseMaxNa) {
    set_error("two-level preconditioner: %d aggregates exceed %d", na, kCoarseMaxNa);
    return SOPHIA_E_CAPACITY;
  }
  void* blas = c.blas;
  hipStream_t side = c.side;
  hipEvent_t ev_in = c.ev_in, ev_rows = c.ev_rows;
  c.blas = nullptr;
  c.side = nullptr;
  c.ev_in = c.ev_rows = nullptr;
  coarse_free(c);
  c.blas = blas;
  c.side = side;
  c.ev_in = ev_in;
  c.ev_rows = ev_rows;
  c.NI = NI; c.NJ = NJ; c.NK = NK;
  c.AX = AX; c.AY = AY; c.AZ = AZ; c.na = na; c.ld = ld;
  const size_t m = (size_t)ld * ld * sizeof(float);
  // padding rows and columns stay zero: X's padding columns multiply the
  // zero padding of c in the GEMV
  for (float** p : {&c.E, &c.X, &c.X2, &c.U}) {
    HIP_TRY(hipMalloc((void**)p, m));
    HIP_TRY(hipMemset(*p, 0, m));
  }
  HIP_TRY(hipMalloc((void**)&c.c, (size_t)ld * 8));
  HIP_TRY(hipMemset(c.c, 0, (size_t)ld * 8));
  HIP_TRY(hipMalloc((void**)&c.y, (size_t)na * 8));
  HIP_TRY(hipMalloc((void**)&c.partials, kResBlocks * 8));
  HIP_TRY(hipMalloc((void**)&c.slots, (size_t)8 * w.n * 8));
  c.slots_n = w.n;
  HIP_TRY(hipHostMalloc((void**)&c.host, kResBlocks * 8, hipHostMallocDefault));
  c.valid = false;
  return 0;
}

// U = 2 I - E X; returns ||U - I||_F / sqrt(na) (synchronises)
int ns_residual(hipStream_t s, CoarseWork& c, double* res) {
  const int na = c.na, ld = c.ld;
  const int g = (int)std::min<size_t>(((size_t)na * ld + 255) / 256, 4096);
  k_identity<<<g, 256, 0, s>>>(c.U, na, ld);
  const float alpha = -1.0f, beta = 2.0f;
  // row-major buffers read as column-major matrices are transposes:
  // X^T E^T = (E X)^T, so the column-major product (X)(E) is E X row-major
  BLAS_TRY(rocblas_sgemm((rocblas_handle)c.blas, rocblas_operation_none, rocblas_operation_none, na, na, na, &alpha,
                         c.X, ld, c.E, ld, &beta, c.U, ld));
  k_res_partials<<<kResBlocks, 256, 0, s>>>(c.U, na, ld, c.partials);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(c.host, c.partials, kResBlocks * 8, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  double t = 0;
  for (int b = 0; b < kResBlocks; b++) t += c.host[b];
  *res = std::sqrt(t / na);
  return 0;
}

// X <- X U (U from ns_residual)
int ns_update(hipStream_t s, CoarseWork& c) {
  const int na = c.na, ld = c.ld;
  const float alpha = 1.0f, beta = 0.0f;
  // column-major (U)(X) = U^T X^T = (X U)^T: X U row-major
  BLAS_TRY(rocblas_sgemm((rocblas_handle)c.blas, rocblas_operation_none, rocblas_operation_none, na, na, na, &alpha,
                         c.U, ld, c.X, ld, &beta, c.X2, ld));
  std::swap(c.X, c.X2);
  return 0;
}

}  // namespace

// E = Z^T A Z: per-particle aggregate-slot sums, then one row of E per aggregate
static int coarse_rows(hipStream_t s, CoarseWork& c, const PpeWork& w) {
  const PpeOp o = ppe_op(w);
  const AggGrid g = agg_grid(c);
  const int gr = grid_xcd(w.n);
  if (w.fmt == kFmtRun8) {
    if (w.mixed) k_coarse_rows<kFmtRun8, 1><<<gr, 256, 0, s>>>(o, g, c.slots);
    else k_coarse_rows<kFmtRun8, 2><<<gr, 256, 0, s>>>(o, g, c.slots);
  } else {
    if (w.mixed) k_coarse_rows<kFmtIdx32, 1><<<gr, 256, 0, s>>>(o, g, c.slots);
    else k_coarse_rows<kFmtIdx32, 2><<<gr, 256, 0, s>>>(o, g, c.slots);
  }
  k_coarse_reduce<<<c.na, 64, 0, s>>>(o, g, c.slots, c.E, c.ld);
  HIP_TRY(hipGetLastError());
  return 0;
}

int coarse_rows_early(hipStream_t s, CoarseWork& c, const PpeWork& w, Timers* tm) {
  if (!w.rowmajor || w.n <= 0) return 0;
  TRY(alloc_coarse(c, w));
  if (!c.side) {
    HIP_TRY(hipStreamCreateWithFlags(&c.side, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c.ev_in, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c.ev_rows, hipEventDisableTiming));
  }
  HIP_TRY(hipEventRecord(c.ev_in, s));
  HIP_TRY(hipStreamWaitEvent(c.side, c.ev_in, 0));
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (tm) {
    t0 = tm->get();
    t1 = tm->get();
    HIP_TRY(hipEventRecord(t0, c.side));
  }
  TRY(coarse_rows(c.side, c, w));
  if (tm) {
    HIP_TRY(hipEventRecord(t1, c.side));
    tm->add_span("ppe_coarse_rows_side", t0, t1);
  }
  HIP_TRY(hipEventRecord(c.ev_rows, c.side));
  c.rows_pending = true;
  c.rows_gen = w.gen;
  c.e_gen = w.gen;
  return 0;
}

int coarse_settle(hipStream_t s, CoarseWork& c) {
  if (!c.rows_pending) return 0;
  HIP_TRY(hipStreamWaitEvent(s, c.ev_rows, 0));
  c.rows_pending = false;
  return 0;
}

int coarse_setup(hipStream_t s, CoarseWork& c, const PpeWork& w, bool* usable, Timers* tm) {
  *usable = false;
  if (!w.rowmajor || w.n <= 0) return 0;
  TRY(alloc_coarse(c, w));
  if (!c.blas) {
    rocblas_handle h = nullptr;
    BLAS_TRY(rocblas_create_handle(&h));
    c.blas = h;
    BLAS_TRY(rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed));
    BLAS_TRY(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
  }
  BLAS_TRY(rocblas_set_stream((rocblas_handle)c.blas, s));
  TimeScope ts(tm, "ppe_coarse_setup");
  const bool early = c.rows_pending && c.rows_gen == w.gen;
  TRY(coarse_settle(s, c));
  if (!early) {
    // no side-stream build of this operator: build E here
    TRY(coarse_rows(s, c, w));
    c.e_gen = w.gen;
  }
  const int gsz = (int)std::min<size_t>(((size_t)c.na * c.ld + 255) / 256, 4096);
  c.ns_steps = 0;
  for (int attempt = 0; attempt < 2; attempt++) {
    if (!c.valid || attempt > 0) {
      k_diag_inverse<<<gsz, 256, 0, s>>>(c.E, c.X, c.na, c.ld);
      HIP_TRY(hipGetLastError());
    }
    // warm: a few steps from the previous inverse; cold: quadratic
    // convergence sets in once ||I - E X|| < 1, after ~log2(cond E) steps
    const int max_steps = (c.valid && attempt == 0) ? 6 : 60;
    double res = 0, prev = INFINITY;
    bool ok = false;
    for (int k = 0; k <= max_steps; k++) {
      TRY(ns_residual(s, c, &res));
      if (!(res == res) || (res > 1.0 && res > prev * 1.0001 && k > 2)) break;  // diverging
      if (res <= kCoarseTol) {
        ok = true;
        break;
      }
      if (k == max_steps) break;
      prev = res;
      TRY(ns_update(s, c));
      c.ns_steps++;
    }
    c.res = res;
    c.x_gen = c.e_gen;
    if (ok) {
      c.valid = true;
      *usable = true;
      return 0;
    }
    c.valid = false;
  }
  return 0;
}

template <typename V>
int coarse_apply_update(hipStream_t s, const CoarseWork& c, const PpeWork& w, const CoarseUpd<V>& u, int kind,
                        const int* done, const int* brk) {
  const PpeOp o = ppe_op(w);
  const AggGrid g = agg_grid(c);
  if (kind == 1) k_update_restrict<V, 1><<<c.na + 1, 256, 0, s>>>(o, g, c.na, u, c.c, done, brk);
  else k_update_restrict<V, 2><<<c.na + 1, 256, 0, s>>>(o, g, c.na, u, c.c, done, brk);
  const int gb = (c.na + 3) / 4;
  k_coarse_gemv<<<gb, 256, 0, s>>>(c.X, c.na, c.ld, c.c, c.y, done, brk);
  k_prolong_add<V><<<grid_for(w.n), kBlock, 0, s>>>(o, g, c.y, u.out, done, brk);
  HIP_TRY(hipGetLastError());
  return 0;
}

int coarse_export(hipStream_t s, const CoarseWork& c, float* E, float* X) {
  const size_t row = (size_t)c.na * sizeof(float), pitch = (size_t)c.ld * sizeof(float);
  if (E) HIP_TRY(hipMemcpy2DAsync(E, row, c.E, pitch, row, (size_t)c.na, hipMemcpyDeviceToHost, s));
  if (X) HIP_TRY(hipMemcpy2DAsync(X, row, c.X, pitch, row, (size_t)c.na, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

void coarse_free(CoarseWork& c) {
  auto F = [](void* p) {
    if (p) (void)hipFree(p);
  };
  F(c.E); F(c.X); F(c.X2); F(c.U); F(c.c); F(c.y); F(c.partials); F(c.slots);
  if (c.host) (void)hipHostFree(c.host);
  if (c.blas) (void)rocblas_destroy_handle((rocblas_handle)c.blas);
  if (c.side) {
    (void)hipStreamSynchronize(c.side);
    (void)hipStreamDestroy(c.side);
  }
  if (c.ev_in) (void)hipEventDestroy(c.ev_in);
  if (c.ev_rows) (void)hipEventDestroy(c.ev_rows);
  c = CoarseWork{};
}

template int coarse_apply_update<float>(hipStream_t, const CoarseWork&, const PpeWork&, const CoarseUpd<float>&,
                                        int, const int*, const int*);
template int coarse_apply_update<double>(hipStream_t, const CoarseWork&, const PpeWork&, const CoarseUpd<double>&,
                                         int, const int*, const int*);

}  // namespace sophia
