// NNPS on gfx950: cell keys, stable LSD radix sort, cell ranges, SoA
// permutation, and the SPH-SPH neighbour list.
//
// Reference: KERNEL_index_particle_to_cell_sph/_dem (functions_NNPS.cuh:
// 762-821), cub::DeviceRadixSort::SortPairs (ISPH_Calc_seperate.cuh:82-85),
// KERNEL_reorder_sph/_dem (functions_NNPS.cuh:670-730).
//
// The sort is a stable least-significant-digit radix sort over 8-bit digits
// and only the bits the largest key needs (2-3 passes for 10^5-10^6 cells).
// Each pass is three launches: per-tile digit histograms, one scan of the
// digit-major histogram table, and a stable scatter in which every wave64
// ranks its lanes by digit with eight ballots (lanes with equal digits get
// consecutive slots in lane order) and the four waves of a workgroup are
// ordered through an LDS prefix.  Stability makes the result equal to the
// reference's SortPairs of (key, identity index).
#include "common.hpp"

#include <mutex>

namespace sophia {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr int kTile = kSortThreads * kSortItems;  // 4096 keys per workgroup

__global__ void k_keys(DevParams P, int n, const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, const int* __restrict__ itype, uint32_t* __restrict__ keys,
                       uint32_t* __restrict__ vals) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t key;
  if (itype[i] == 3) {
    key = (uint32_t)(P.num_cells - 1);  // dummy particles go to the spare slot
  } else {
    int a, b, c;
    cell_of(P, x[i], y[i], z[i], a, b, c);
    key = (uint32_t)cell_key(P, a, b, c);
  }
  keys[i] = key;
  vals[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kSortThreads) void k_hist(int n, int shift, const uint32_t* __restrict__ keys,
                                                       uint32_t* __restrict__ hist, int ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  int base = blockIdx.x * kTile;
#pragma unroll
  for (int r = 0; r < kSortItems; r++) {
    int i = base + r * kSortThreads + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Exclusive scan of the digit-major histogram hist[d * ntiles + tile] in two
// launches of 256 workgroups: k_scan_rows scans each digit's row in place and
// writes the row total to tot[d]; k_scan_add adds the totals of the digits
// below d to row d.
__global__ __launch_bounds__(256) void k_scan_rows(uint32_t* __restrict__ a, int m, uint32_t* __restrict__ tot) {
  __shared__ uint32_t ws[4];
  uint32_t* row = a + (size_t)blockIdx.x * m;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t carry = 0;
  for (int base = 0; base < m; base += 256) {
    const int i = base + t;
    const uint32_t x = i < m ? row[i] : 0u;
    uint32_t s = x;  // inclusive scan over the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t y = __shfl_up(s, off);
      if (lane >= off) s += y;
    }
    if (lane == 63) ws[w] = s;
    __syncthreads();
    uint32_t wp = 0;
    for (int q = 0; q < w; q++) wp += ws[q];
    const uint32_t all = ws[0] + ws[1] + ws[2] + ws[3];
    if (i < m) row[i] = carry + wp + s - x;
    carry += all;
    __syncthreads();
  }
  if (t == 0) tot[blockIdx.x] = carry;
}

__global__ __launch_bounds__(256) void k_scan_add(uint32_t* __restrict__ a, int m, const uint32_t* __restrict__ tot) {
  __shared__ uint32_t ws[4];
  const int t = threadIdx.x, d = blockIdx.x;
  uint32_t v = t < d ? tot[t] : 0u;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((t & 63) == 0) ws[t >> 6] = v;
  __syncthreads();
  const uint32_t off = ws[0] + ws[1] + ws[2] + ws[3];
  if (off == 0) return;
  uint32_t* row = a + (size_t)d * m;
  for (int i = t; i < m; i += 256) row[i] += off;
}

__global__ __launch_bounds__(kSortThreads) void k_scatter(int n, int shift, const uint32_t* __restrict__ kin,
                                                          const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
                                                          uint32_t* __restrict__ vout, const uint32_t* __restrict__ hist,
                                                          int ntiles) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[kSortThreads / 64][256];
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  base[t] = hist[t * ntiles + blockIdx.x];
  for (int w = 0; w < kSortThreads / 64; w++) wcnt[w][t] = 0;
  __syncthreads();
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const int tile0 = blockIdx.x * kTile;
  for (int r = 0; r < kSortItems; r++) {
    int i = tile0 + r * kSortThreads + t;
    bool valid = i < n;
    uint32_t key = valid ? kin[i] : 0u;
    uint32_t val = valid ? vin[i] : 0u;
    uint32_t d = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      bool bit = (d >> b) & 1u;
      uint64_t m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    uint32_t rank = (uint32_t)__popcll(peers & lt);
    uint32_t cnt = (uint32_t)__popcll(peers);
    bool leader = valid && rank == 0;
    if (leader) wcnt[wave][d] = cnt;
    __syncthreads();
    // exclusive prefix over waves for every digit, then advance the tile base
    {
      uint32_t run = base[t];
      for (int w = 0; w < kSortThreads / 64; w++) {
        uint32_t c = wcnt[w][t];
        wcnt[w][t] = run;
        run += c;
      }
      base[t] = run;
    }
    __syncthreads();
    if (valid) {
      uint32_t pos = wcnt[wave][d] + rank;
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
    for (int w = 0; w < kSortThreads / 64; w++) wcnt[w][t] = 0;
    __syncthreads();
  }
}

__global__ void k_ranges(int n, const uint32_t* __restrict__ keys, uint32_t* __restrict__ start,
                         uint32_t* __restrict__ end) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) start[k] = (uint32_t)i;
  if (i == n - 1 || keys[i + 1] != k) end[k] = (uint32_t)(i + 1);
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

constexpr int kMaxFill = 24;
struct FillArgs {
  uint32_t* p[kMaxFill];
  size_t n[kMaxFill];
  uint32_t v[kMaxFill];
};

__global__ void k_fill_many(FillArgs a) {
  uint32_t* p = a.p[blockIdx.y];
  const size_t n = a.n[blockIdx.y];
  const uint32_t v = a.v[blockIdx.y];
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

constexpr int kMaxPermute = 48;
struct PermuteArgs {
  PermuteField f[kMaxPermute];
  int nf;
};

__global__ void k_permute(int n, const uint32_t* __restrict__ order, PermuteArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t src = order[i];
  for (int k = 0; k < a.nf; k++) {
    if (a.f[k].bytes == 8)
      ((uint64_t*)a.f[k].dst)[i] = ((const uint64_t*)a.f[k].src)[src];
    else
      ((uint32_t*)a.f[k].dst)[i] = ((const uint32_t*)a.f[k].src)[src];
  }
}

// SPH-SPH neighbour list within kappa*h_i over the 27 cells, in the
// reference's visiting order.  Entries that pass sqrt(r2) < R but fail
// sqrt(r2)+1e-20 < R carry kNbrFlag (those stages that use the +1e-20
// test skip them).
// pos4 record: grid cell c = floor((x - x0) / g) per axis (clamped to
// 0..1023) and the FP32 offset x - (x0 + c g), computed in FP64
__device__ inline uint32_t pos_cell(double x, double x0, double g, float& off) {
  const int c = min(max((int)floor((x - x0) / g), 0), 1023);
  off = (float)(x - (x0 + (double)c * g));
  return (uint32_t)c;
}
template <typename R>
__global__ void k_pos4(SphView<R> v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  const double g = v.pgrid;
  float ox, oy, oz;
  const uint32_t cx = pos_cell(v.x[i], v.pg0x, g, ox);
  const uint32_t cy = pos_cell(v.y[i], v.pg0y, g, oy);
  const uint32_t cz = pos_cell(v.z[i], v.pg0z, g, oz);
  v.pos4[i] = make_float4(ox, oy, oz, __uint_as_float(cx | cy << 10 | cz << 20));
}

// Candidates of the 27 cells are tested on the pos4 records first: farther
// than Rs + tol or nearer than Rs - tol is decided there (one 16 B load);
// only the band |r - Rs| <= tol takes the exact FP64 test, so the list, its
// order and the r + 1e-20 < Rs flag are exactly those of the FP64 test.  tol
// bounds the FP32 distance error: 2^-21 max|coordinate| + 2^-20 Rs, well
// above the rounding of the record differences (2^-21 of the grid spacing,
// itself below the domain extent) and of r^2.
template <typename R>
__global__ void k_neighbors(DevParams P, SphView<R> v, CellView cv, int* maxcnt, double tol_abs) {
  int i = particle_index();
  int c = 0;
  if (i < v.n) {
    const double xi = v.x[i], yi = v.y[i], zi = v.z[i];
    const double Rs = P.kappa * v.h[i];
    const double tol = tol_abs + Rs * (1.0 / 1048576.0);
    const float lo2 = (float)((Rs - tol) * (Rs - tol)), hi2 = (float)((Rs + tol) * (Rs + tol));
    const float4 pi = v.pos4[i];
    const float g = v.pgrid;
    const int n = v.n, cap = v.nbr_cap;
    // the last 8 entries, oldest first; a full chunk goes out as two 16-byte stores
    uint32_t r[kNbrChunk] = {};
    auto put = [&](int k0) {
      uint4* p = (uint4*)(v.nbr + nbr_slot(k0, i, n));
      p[0] = make_uint4(r[0], r[1], r[2], r[3]);
      p[1] = make_uint4(r[4], r[5], r[6], r[7]);
    };
    auto shift_in = [&](uint32_t e) {
#pragma unroll
      for (int t = 0; t < kNbrChunk - 1; t++) r[t] = r[t + 1];
      r[kNbrChunk - 1] = e;
    };
    for_cells(P, xi, yi, zi, cv, [&](int j) {
      const float3 dd = rel_diff(pi, v.pos4[j], g, v.pmask);
      const float d2 = dd.x * dd.x + dd.y * dd.y + dd.z * dd.z;
      if (d2 > hi2) return;
      uint32_t e = (uint32_t)j;
      if (!(d2 < lo2)) {
        double ex, ey, ez;
        pdiff(v, xi, yi, zi, j, ex, ey, ez);
        const double r = dist_exact(ex, ey, ez);
        if (!(r < Rs)) return;
        if (!(r + 1e-20 < Rs)) e |= kNbrFlag;
      }
      shift_in(e);
      c++;
      if ((c & (kNbrChunk - 1)) == 0 && c <= cap) put(c - kNbrChunk);
    });
    const int m = c & (kNbrChunk - 1);
    if (m != 0 && c < cap) {  // pad the last chunk with the last entry
      const uint32_t last = r[kNbrChunk - 1];
      for (int t = m; t < kNbrChunk; t++) shift_in(last);
      put(c - m);
    }
    v.nbr_cnt[i] = c;
  }
  // one atomic per wave for the largest count (capacity check on the host)
  for (int off = 32; off > 0; off >>= 1) c = max(c, __shfl_xor(c, off));
  if ((threadIdx.x & 63) == 0) atomicMax(maxcnt, c);
}

__global__ void k_knl(int kt, int dim, double h, const double* r, int n, double* w, double* dw) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double A = knl_tmpA<double>(kt, dim, h);
  w[i] = knl_w<double>(kt, A, h, r[i]);
  dw[i] = knl_dw<double>(kt, A, h, r[i]);
}

}  // namespace

int nnps_keys(hipStream_t s, const DevParams& P, int n, const double* x, const double* y, const double* z,
              const int* itype, uint32_t* keys, uint32_t* vals) {
  if (n <= 0) return 0;
  k_keys<<<grid_for(n), kBlock, 0, s>>>(P, n, x, y, z, itype, keys, vals);
  HIP_TRY(hipGetLastError());
  return 0;
}

int radix_sort_pairs(hipStream_t s, int n, int nbits, SortScratch& sc, uint32_t** keys_out, uint32_t** vals_out) {
  uint32_t *ka = sc.keys_a, *kb = sc.keys_b, *va = sc.vals_a, *vb = sc.vals_b;
  if (n > 0) {
    int ntiles = (n + kTile - 1) / kTile;
    if (ntiles > sc.max_tiles) {
      set_error("radix sort: %d tiles exceed scratch capacity %d", ntiles, sc.max_tiles);
      return SOPHIA_E_CAPACITY;
    }
    for (int shift = 0; shift < nbits; shift += 8) {
      k_hist<<<ntiles, kSortThreads, 0, s>>>(n, shift, ka, sc.hist, ntiles);
      k_scan_rows<<<256, 256, 0, s>>>(sc.hist, ntiles, sc.hist + (size_t)256 * sc.max_tiles);
      k_scan_add<<<256, 256, 0, s>>>(sc.hist, ntiles, sc.hist + (size_t)256 * sc.max_tiles);
      k_scatter<<<ntiles, kSortThreads, 0, s>>>(n, shift, ka, va, kb, vb, sc.hist, ntiles);
      HIP_TRY(hipGetLastError());
      uint32_t* t = ka; ka = kb; kb = t;
      t = va; va = vb; vb = t;
    }
  }
  *keys_out = ka;
  *vals_out = va;
  return 0;
}

int cell_ranges(hipStream_t s, int n, int num_cells, const uint32_t* keys, uint32_t* start, uint32_t* end) {
  const FillSpan f[2] = {{start, (size_t)num_cells, kEmpty}, {end, (size_t)num_cells, 0u}};
  int rc = fill_u32_many(s, f, 2);
  if (rc) return rc;
  if (n <= 0) return 0;
  k_ranges<<<grid_for(n), kBlock, 0, s>>>(n, keys, start, end);
  HIP_TRY(hipGetLastError());
  return 0;
}

int permute_fields(hipStream_t s, int n, const uint32_t* order, const PermuteField* fields, int nf) {
  if (n <= 0) return 0;
  for (int off = 0; off < nf; off += kMaxPermute) {
    PermuteArgs a;
    a.nf = nf - off < kMaxPermute ? nf - off : kMaxPermute;
    for (int k = 0; k < a.nf; k++) a.f[k] = fields[off + k];
    k_permute<<<grid_for(n), kBlock, 0, s>>>(n, order, a);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}

int fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n) {
  if (n == 0) return 0;
  size_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  k_fill_u32<<<(unsigned)blocks, kBlock, 0, s>>>(p, v, n);
  HIP_TRY(hipGetLastError());
  return 0;
}

int fill_u32_many(hipStream_t s, const FillSpan* spans, int count) {
  for (int off = 0; off < count; off += kMaxFill) {
    const int k = count - off < kMaxFill ? count - off : kMaxFill;
    FillArgs a;
    size_t nmax = 0;
    for (int q = 0; q < k; q++) {
      a.p[q] = spans[off + q].p;
      a.n[q] = spans[off + q].n;
      a.v[q] = spans[off + q].v;
      if (a.n[q] > nmax) nmax = a.n[q];
    }
    if (nmax == 0) continue;
    size_t blocks = (nmax + kBlock - 1) / kBlock;
    if (blocks > 2048) blocks = 2048;
    k_fill_many<<<dim3((unsigned)blocks, (unsigned)k), kBlock, 0, s>>>(a);
    HIP_TRY(hipGetLastError());
  }
  return 0;
}

int fill_zero_bytes(hipStream_t s, void* p, size_t bytes) {
  if (bytes % 4) {
    set_error("fill_zero_bytes: size %zu not a multiple of 4", bytes);
    return SOPHIA_E_ARG;
  }
  return fill_u32(s, (uint32_t*)p, 0u, bytes / 4);
}

template <typename R>
int build_sph_neighbors(hipStream_t s, const DevParams& P, const SphView<R>& v, const CellView& cv, int* d_maxcnt) {
  if (v.n <= 0) return 0;
  k_pos4<R><<<grid_for(v.n), kBlock, 0, s>>>(v);
  const double m = fmax(fmax(fmax(fabs(P.x_min), fabs(P.x_max)), fmax(fabs(P.y_min), fabs(P.y_max))),
                        fmax(fabs(P.z_min), fabs(P.z_max)));
  k_neighbors<R><<<grid_xcd(v.n), kBlock, 0, s>>>(P, v, cv, d_maxcnt, m * (1.0 / 2097152.0));
  HIP_TRY(hipGetLastError());
  return 0;
}
template <typename R>
int refresh_pos4(hipStream_t s, const SphView<R>& v) {
  if (v.n <= 0) return 0;
  k_pos4<R><<<grid_for(v.n), kBlock, 0, s>>>(v);
  HIP_TRY(hipGetLastError());
  return 0;
}
template int refresh_pos4<float>(hipStream_t, const SphView<float>&);
template int refresh_pos4<double>(hipStream_t, const SphView<double>&);
template int build_sph_neighbors<float>(hipStream_t, const DevParams&, const SphView<float>&, const CellView&, int*);
template int build_sph_neighbors<double>(hipStream_t, const DevParams&, const SphView<double>&, const CellView&, int*);

int knl_eval_device(hipStream_t s, int kt, int dim, double h, const double* r, int n, double* w, double* dw) {
  if (n <= 0) return 0;
  k_knl<<<grid_for(n), kBlock, 0, s>>>(kt, dim, h, r, n, w, dw);
  HIP_TRY(hipGetLastError());
  return 0;
}

// Allocation-free kernel evaluation: the three staging arrays (r, W, dW/dr)
// live in one file-scope device array, kKnlScratchMax values each.  Host code
// reaches it only through the address hipGetSymbolAddress returns for the
// current device (the array's name is a host-side stand-in, not a device
// address); every device holds its own copy of the array.
__device__ __align__(256) double scratch[3 * kKnlScratchMax];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int knl_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device %d outside the %d devices the scratch table covers", device, kMaxDevices);
    return SOPHIA_E_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<double*>(p);
  return 0;
}

int knl_eval_noalloc(int kt, int dim, double h, const double* r, int n, double* w, double* dw) {
  if (n > kKnlScratchMax) {
    set_error("sophia_knl_eval_noalloc: n = %d exceeds the scratch capacity of %d values", n, kKnlScratchMax);
    return SOPHIA_E_CAPACITY;
  }
  if (n <= 0) return 0;
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  TRY(knl_scratch_init(dev));
  double* base = g_scratch_ptr[dev];
  double *dr = base, *dwv = base + kKnlScratchMax, *ddw = base + 2 * kKnlScratchMax;
  const size_t bytes = (size_t)n * sizeof(double);
  HIP_TRY(hipMemcpy(dr, r, bytes, hipMemcpyHostToDevice));
  // every output value is written by the kernel, so nothing needs clearing
  k_knl<<<grid_for(n), kBlock, 0, nullptr>>>(kt, dim, h, dr, n, dwv, ddw);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(w, dwv, bytes, hipMemcpyDeviceToHost));
  HIP_TRY(hipMemcpy(dw, ddw, bytes, hipMemcpyDeviceToHost));
  return 0;
}

}  // namespace sophia
