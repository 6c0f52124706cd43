// ISPH pressure-Poisson equation on gfx950.
//
// Assembly follows KERNEL_PPE3D_sph (functions_PPE2.cuh:542-698) for the
// rows it computes (fluid, 1 <= p_type < 1000, not an open-boundary buffer,
// i_type <= 2): per neighbour j of row i
//   a_ij = m_j (eps_i rho_i + eps_j rho_j) / (rho_i rho_j) * c_ij,
//   c_ij = -(x_ij . grad~W_ij) / r^2
// in FP64 arithmetic, together with d_i = sum_j a_ij (PPE2), b_i (PPE3) and
// sum_j a_ij p_j (PPE4).  The reference then takes one Jacobi sweep
// p_i = (b_i + sum_j a_ij p_j) / d_i (functions_PPE2.cuh:868-878); this
// library also solves the system to a relative residual with BiCGStab:
//   (A p)_i = d_i p_i - sum_{j row} a_ij p_j = b_i + sum_{j not row} a_ij p_j
// whose fixed point is the Jacobi update (non-row columns -- walls and
// buffers -- keep their lagged pressures, as in the reference).
//
// Operator storage (PpeOp): only row columns are kept, compacted per row,
// slot-quad major.  Columns are u8 offsets inside the 9 x-runs of the
// 27-cell stencil (kFmtRun8) or u32 indices (kFmtIdx32).
//   FP64 path : a_ij as double                        (8 B + 1 B per entry)
//   mixed path: hi = fl32(a), lo = fl32(a - hi)
//               inner FP32 iterations read hi only    (4 B + 1 B per entry)
//               FP64 residuals read hi + lo (48-bit a)
//
// Solver: BiCGStab (van der Vorst 1992) with Jacobi right preconditioning.
// The mixed path runs the recurrences on FP32 vectors with the hi operator
// and folds the iterate into an FP64 solution with a true FP64 residual
// whenever the recursive residual has dropped by a factor delta since the
// last fold ("reliable updates", Sleijpen & van der Vorst 1996), so the
// Krylov space is never restarted.  All scalars stay on the device; every
// reduction is two-stage and deterministic (per-block partials summed in a
// fixed order by the last block to finish).  The host checks convergence
// once per chunk of iterations.
#include "ppe.hpp"
#include "nbr.hpp"

#include "coarse.hpp"

#include <cstring>
#include <type_traits>

namespace sophia {

namespace {

constexpr int kRedThreads = 256;
constexpr int kMaxGrid = 2048;  // 8 XCDs x 256 blocks
enum { kFlagOffset = 1u, kFlagRunCount = 2u, kFlagOrder = 4u };

int grid_rows(int n) {
  int g = grid_for(n > 0 ? n : 1, kRedThreads);
  g = ((g + 7) / 8) * 8;
  return g < kMaxGrid ? g : kMaxGrid;
}

// Device-resident BiCGStab state.
struct BiState {
  double rho, rho_old, alpha, omega, beta;
  double rr;        // recursive residual norm^2 (scaled units)
  double rr_true;   // last FP64 residual norm^2 (pressure units)
  double max_rr;    // largest recursive rr since the last fold
  double tol2;      // stopping threshold on rr (scaled units)
  double tol2_true; // stopping threshold on rr_true
  double delta2;    // fold when rr < delta2 * max_rr
  double xscale;    // x64 += xscale * xs
  double norm;      // result slot of plain reductions
  int iter, max_iter;
  int done;         // 1 converged, 2 breakdown, 3 iteration limit
  int update, stop, brk, first, folds;
};

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// Rows of this block: the rows are split into 8 contiguous chunks, one per
// XCD (workgroups are dealt to XCDs round robin), so the neighbour gathers
// of a block hit the L2 of the XCD that owns the surrounding rows.
struct RowIter {
  int i, end, stride;
};
__device__ inline RowIter xcd_rows(int n) {
  const int G = gridDim.x;
  const int x = blockIdx.x & 7, b = blockIdx.x >> 3, per = G >> 3;
  const int chunk = ((((n + 7) >> 3) + 63) >> 6) << 6;
  const int lo = x * chunk;
  const int hi = min(n, lo + chunk);
  return RowIter{lo + b * (int)blockDim.x + (int)threadIdx.x, hi, per * (int)blockDim.x};
}

// Block sums of NV values into partials[q * G + block].  A one-block
// finalize kernel adds the G partials in a fixed order (deterministic); a
// "last block" reduction inside the producing kernel would need a
// device-scope release fence per block, which writes back the L2 of every
// XCD and costs more than the extra launch.
template <int NV>
__device__ void block_partials(const double (&v)[NV], double* partials) {
  __shared__ double red[NV][kRedThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = wave_sum(v[q]);
    if (lane == 0) red[q][wave] = s;
  }
  __syncthreads();
  if (threadIdx.x < NV) {
    const int q = threadIdx.x;
    double s = 0;
    for (int w = 0; w < kRedThreads / 64; w++) s += red[q][w];
    partials[q * gridDim.x + blockIdx.x] = s;
  }
}

// fixed-order sum of G partials per value, one block of kFinThreads
constexpr int kFinThreads = 1024;
template <int NV>
__device__ void sum_partials(const double* partials, int G, double (&tot)[NV]) {
  __shared__ double red[NV][kFinThreads / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = 0;
    for (int b = threadIdx.x; b < G; b += kFinThreads) s += partials[q * G + b];
    s = wave_sum(s);
    if (lane == 0) red[q][wave] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = 0;
    for (int w = 0; w < kFinThreads / 64; w++) s += red[q][w];
    tot[q] = s;
  }
}

// ----- row sums over the compact operator ----------------------------------
// CM: 0 = hi (FP32 accumulation), 1 = hi + lo (FP64), 2 = FP64 coefficients.
// For CM 0 and 1, b holds the raw hi words: their low byte is the kFmtRun8
// column offset.
template <int CM>
struct CoefQuad;
template <>
struct CoefQuad<0> {
  float c[4];
  uint32_t b[4];
  __device__ void load(const PpeOp& o, size_t qi) {
    const uint4 a = reinterpret_cast<const uint4*>(o.hi)[qi];
    b[0] = a.x; b[1] = a.y; b[2] = a.z; b[3] = a.w;
#pragma unroll
    for (int u = 0; u < 4; u++) c[u] = hi_val(b[u]);
  }
};
template <>
struct CoefQuad<1> {
  double c[4];
  uint32_t b[4];
  __device__ void load(const PpeOp& o, size_t qi) {
    const uint4 a = reinterpret_cast<const uint4*>(o.hi)[qi];
    const float4 l = reinterpret_cast<const float4*>(o.lo)[qi];
    b[0] = a.x; b[1] = a.y; b[2] = a.z; b[3] = a.w;
    const float lv[4] = {l.x, l.y, l.z, l.w};
#pragma unroll
    for (int u = 0; u < 4; u++) c[u] = (double)hi_val(b[u]) + (double)lv[u];
  }
};
template <>
struct CoefQuad<2> {
  double c[4];
  uint32_t b[4];  // unused: the FP64 path keeps its offsets in idx
  __device__ void load(const PpeOp& o, size_t qi) {
    double2 a = reinterpret_cast<const double2*>(o.c64)[2 * qi];
    double2 e = reinterpret_cast<const double2*>(o.c64)[2 * qi + 1];
    c[0] = a.x; c[1] = a.y; c[2] = e.x; c[3] = e.y;
  }
};

// sum_j a_ij x_j over the stored (row) columns of row i.
// Entries are consumed eight at a time (two slot quads) while the next two
// quads' column words and coefficients are in flight, so a wave keeps about
// 2.5 KB of operator stream outstanding.  The loop body has no branches, so
// that the eight gathers of x go out together and the waits for them do not
// include the prefetch: the run of entry k comes from the prefix ends of the
// run lengths through a select chain, the prefetch index is clamped to the
// last quad (a re-read that hits the cache), and entries k >= cnt are masked
// to a zero coefficient on the row's own column.
//
// STAGED (kFmtRun8 only): x is the block's LDS copy of the particles of its
// nine run rows, and adj[r] maps a particle index of run row r to its LDS
// slot (slot = index + adj[r]); the row's own particle lies in run row 4.
template <int FMT, int CM, typename AT, typename XT, bool STAGED = false>
__device__ __forceinline__ AT row_sum(const PpeOp& o, int i, const XT* __restrict__ x,
                                      const int* adj = nullptr) {
  const int cnt = o.ccnt[i];
  AT s = 0;
  if (cnt == 0) return s;
  const size_t n = (size_t)o.n;
  const int nq = (cnt + 3) >> 2;
  const int last = nq - 1;
  if (FMT == kFmtRun8) {
    const uint32_t key = o.keys[i];
    const int NI = o.NI, NJ = o.NJ;
    const int cx = (int)(key % (uint32_t)NI);
    const int t = (int)(key / (uint32_t)NI);
    const int cy = t % NJ, cz = t / NJ;
    const int xl = cx > 0 ? cx - 1 : 0;
    const uint64_t rc = o.runc[i];
    // first particle of each run, in the (dz, dy) order of the list (a run
    // outside the grid has length 0 and keeps base 0), and the prefix ends
    int B[9], E[8];
    int acc = 0;
#pragma unroll
    for (int r = 0; r < 9; r++) {
      const int b = cy + (r % 3) - 1, c = cz + (r / 3) - 1;
      const int len = (int)((rc >> (kRunBits * r)) & kRunMax);
      B[r] = len ? (int)o.cell_off[xl + NI * (b + NJ * c)] + (STAGED ? adj[r] : 0) : 0;
      acc += len;
      if (r < 8) E[r] = acc;
    }
    const int self = STAGED ? i + adj[4] : i;
    // mixed operator: the offsets ride in the coefficient words; FP64: idx
    constexpr bool PK = CM < 2;
    const size_t q1 = (size_t)min(1, last) * n + i;
    uint32_t wn0 = PK ? 0u : o.idx[i], wn1 = PK ? 0u : o.idx[q1];
    CoefQuad<CM> cn0, cn1;
    cn0.load(o, (size_t)i);
    cn1.load(o, q1);
    for (int q = 0; q < nq; q += 2) {
      const uint32_t w0 = wn0, w1 = wn1;
      const CoefQuad<CM> c0 = cn0, c1 = cn1;
      int j[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int k = 4 * q + u;
        int base = B[0];
#pragma unroll
        for (int r = 0; r < 8; r++) base = k >= E[r] ? B[r + 1] : base;
        uint32_t off;
        if constexpr (PK) off = (u < 4 ? c0.b[u & 3] : c1.b[u & 3]) & 255u;
        else off = ((u < 4 ? w0 : w1) >> (8 * (u & 3))) & 255u;
        j[u] = k < cnt ? base + (int)off : self;
      }
      XT xv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) xv[u] = x[j[u]];
      const size_t qa = (size_t)min(q + 2, last) * n + i;
      const size_t qb = (size_t)min(q + 3, last) * n + i;
      if constexpr (!PK) {
        wn0 = o.idx[qa];
        wn1 = o.idx[qb];
      }
      cn0.load(o, qa);
      cn1.load(o, qb);
#pragma unroll
      for (int u = 0; u < 4; u++) s += (4 * q + u < cnt ? (AT)c0.c[u] : (AT)0) * (AT)xv[u];
#pragma unroll
      for (int u = 0; u < 4; u++) s += (4 * q + 4 + u < cnt ? (AT)c1.c[u] : (AT)0) * (AT)xv[4 + u];
    }
  } else {
    const uint4* ix = reinterpret_cast<const uint4*>(o.idx);
    const size_t q1 = (size_t)min(1, last) * n + i;
    uint4 jn0 = ix[i], jn1 = ix[q1];
    CoefQuad<CM> cn0, cn1;
    cn0.load(o, (size_t)i);
    cn1.load(o, q1);
    for (int q = 0; q < nq; q += 2) {
      const uint4 j0 = jn0, j1 = jn1;
      const CoefQuad<CM> c0 = cn0, c1 = cn1;
      const int jj[8] = {(int)j0.x, (int)j0.y, (int)j0.z, (int)j0.w, (int)j1.x, (int)j1.y, (int)j1.z, (int)j1.w};
      XT xv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) xv[u] = x[4 * q + u < cnt ? jj[u] : i];
      const size_t qa = (size_t)min(q + 2, last) * n + i;
      const size_t qb = (size_t)min(q + 3, last) * n + i;
      jn0 = ix[qa];
      jn1 = ix[qb];
      cn0.load(o, qa);
      cn1.load(o, qb);
#pragma unroll
      for (int u = 0; u < 4; u++) s += (4 * q + u < cnt ? (AT)c0.c[u] : (AT)0) * (AT)xv[u];
#pragma unroll
      for (int u = 0; u < 4; u++) s += (4 * q + 4 + u < cnt ? (AT)c1.c[u] : (AT)0) * (AT)xv[4 + u];
    }
  }
  return s;
}

// ----- assembly ---------------------------------------------------------------
struct AsmOut {
  int n, NI, NJ, NK;
  const uint32_t* keys;
  const uint32_t* cell_off;
  const int* row;
  int* ccnt;
  uint64_t* runc;
  uint32_t* idx;
  float *hi, *lo;
  double* c64;
  double *diag, *dinv, *rhs;
  float* dgi;
  unsigned* flags;
};

__device__ inline int pick9(const int (&a)[9], int r) {
  int v = a[0];
#pragma unroll
  for (int q = 1; q < 9; q++) v = (r == q) ? a[q] : v;
  return v;
}

__global__ void k_rowmask(DevParams P, int n, const int* ptype, const int* btype, const int* itype, const int* surf,
                          double* pres, int* row) {
  int i = particle_index();
  if (i >= n) return;
  int t = ptype[i];
  const bool fluid = itype[i] <= 2 && !(P.open_boundary > 0 && btype[i] > 0) && t < 1000 && t > 0;
  const bool dirichlet = fluid && P.free_surface > 0 && surf[i] != 0;
  // a free-surface particle is a known value p_ref, not an unknown
  if (dirichlet) pres[i] = P.p_ref;
  bool on = fluid && !dirichlet;
  row[i] = on ? 1 : 0;
}

__global__ void k_cell_offsets(int n, const uint32_t* __restrict__ keys, int nc, uint32_t* __restrict__ off) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c > nc) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (keys[mid] < (uint32_t)c) lo = mid + 1;
    else hi = mid;
  }
  off[c] = (uint32_t)lo;
}

// The assembly's 64 B neighbour records (mixed path), after the row mask
// and the Dirichlet pressures are set
__global__ void k_pack_asm(SphView<float> v, const int* __restrict__ row) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  RecAsm r;
  r.p = v.pos4[i];
  const double m = v.m[i], rho = v.rho[i], por = v.por[i];
  r.vol = m / rho;
  r.por = por;
  r.porrho = por * rho;
  r.pres = v.pres[i];
  r.ux = v.ux[i];
  r.uy = v.uy[i];
  r.uz = v.uz[i];
  const int pt = min(max(v.ptype[i], -32768), 32767);
  r.flags = (pt & 0xFFFF) | ((v.itype[i] & 0xFF) << 16) | ((row[i] ? 1 : 0) << 24);
  v.ra[i] = r;
}

template <typename R, int FMT, int MIXED>
__global__ __launch_bounds__(kBlock) void k_assemble(DevParams P, SphView<R> v, AsmOut o) {
  int i = particle_index();
  if (i >= v.n) return;
  const int n = v.n;
  const size_t nn = (size_t)n;
  const int cnt = min(v.nbr_cnt[i], v.nbr_cap);
  if (!o.row[i]) {
    o.ccnt[i] = 0;
    o.runc[i] = 0;
    o.diag[i] = 0;
    o.dinv[i] = 0;
    o.rhs[i] = 0;
    if (MIXED) o.dgi[i] = 0.f;
    return;
  }
  // runs of the 27-cell stencil (kFmtRun8)
  int B[9], E[9];
  if (FMT == kFmtRun8) {
    const uint32_t key = o.keys[i];
    const int cx = (int)(key % (uint32_t)o.NI);
    const int t = (int)(key / (uint32_t)o.NI);
    const int cy = t % o.NJ, cz = t / o.NJ;
    const int xl = cx > 0 ? cx - 1 : 0, xh = cx < o.NI - 1 ? cx + 1 : cx;
#pragma unroll
    for (int r = 0; r < 9; r++) {
      const int b = cy + (r % 3) - 1, c = cz + (r / 3) - 1;
      if (b >= 0 && b < o.NJ && c >= 0 && c < o.NK) {
        B[r] = (int)o.cell_off[xl + o.NI * (b + o.NJ * c)];
        E[r] = (int)o.cell_off[xh + o.NI * (b + o.NJ * c) + 1];
      } else {
        B[r] = 0;
        E[r] = 0;
      }
    }
  }
  const double hd = v.h[i], ihd = 1.0 / hd;
  const double A = knl_tmpA<double>(P.kernel_type, P.dim, hd);
  // geometry: FP64 positions, or with FP32 fields the pos4 records (one 16 B
  // gather per neighbour, difference error below 2^-21 of the grid spacing)
  constexpr bool REL = std::is_same<R, float>::value;
  const double xi = REL ? 0.0 : v.x[i], yi = REL ? 0.0 : v.y[i], zi = REL ? 0.0 : v.z[i];
  const float4 pi4 = REL ? v.pos4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  const double uxi = v.ux[i], uyi = v.uy[i], uzi = v.uz[i];
  const double rhoi = v.rho[i], pori = v.por[i], irhoi = 1.0 / rhoi;
  double bx = 0, by = 0, bz = 0, Aij = 0, AP = 0, lag = 0;
  double dcomp = 0;  // sum over stored columns of a - hi16(a) (mixed)
  int kc = 0, run = 0;
  uint64_t rcw = 0;
  uint32_t word = 0;
  uint32_t qi4[4] = {0u, 0u, 0u, 0u}, qh4[4] = {0u, 0u, 0u, 0u};
  float ql4[4] = {0.f, 0.f, 0.f, 0.f};
  double qc4[4] = {0.0, 0.0, 0.0, 0.0};
  auto push_quad = [&](uint32_t jj, uint32_t h, float l, double c) {
#pragma unroll
    for (int t = 0; t < 3; t++) {
      qi4[t] = qi4[t + 1];
      qh4[t] = qh4[t + 1];
      ql4[t] = ql4[t + 1];
      qc4[t] = qc4[t + 1];
    }
    qi4[3] = jj;
    qh4[3] = h;
    ql4[3] = l;
    qc4[3] = c;
  };
  auto store_quad = [&](int q) {
    const size_t qs = (size_t)q * nn + i;
    if (FMT == kFmtIdx32) reinterpret_cast<uint4*>(o.idx)[qs] = make_uint4(qi4[0], qi4[1], qi4[2], qi4[3]);
    if (MIXED) {
      reinterpret_cast<uint4*>(o.hi)[qs] = make_uint4(qh4[0], qh4[1], qh4[2], qh4[3]);
      reinterpret_cast<float4*>(o.lo)[qs] = make_float4(ql4[0], ql4[1], ql4[2], ql4[3]);
    } else {
      reinterpret_cast<double2*>(o.c64)[2 * qs] = make_double2(qc4[0], qc4[1]);
      reinterpret_cast<double2*>(o.c64)[2 * qs + 1] = make_double2(qc4[2], qc4[3]);
    }
  };
  unsigned bad = 0;
  // the list in batches of four: every field of the batch is gathered
  // before any is used (straight-line loads, all in flight together), the
  // next batch's indices are loaded, then the four entries run in list order
  typedef typename std::conditional<REL, float, double>::type GT;
  struct AsmNb {
    GT x, y, z;  // x_i - x_j
    double vol, por, porrho, pres;  // V_j = m_j / rho_j, eps_j, eps_j rho_j, p_j
    R ux, uy, uz;
    int pt, it, row;
  };
  const int last4 = (cnt - 1) & ~3;
  uint4 en4 = cnt > 0 ? *(const uint4*)(v.nbr + nbr_slot(0, i, (int)n)) : make_uint4(0u, 0u, 0u, 0u);
  for (int k0 = 0; k0 < cnt; k0 += 4) {
    uint32_t eb[4];
    const uint32_t en[4] = {en4.x, en4.y, en4.z, en4.w};
    AsmNb nb[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      eb[u] = en[u];
      const int j = (int)(eb[u] & ~kNbrFlag);
      if constexpr (REL) {
        // one 64 B record (k_pack_asm) in place of eleven field gathers
        const RecAsm ra = v.ra[j];
        const float3 d = rel_diff(pi4, ra.p, v.pgrid, v.pmask);
        nb[u].x = d.x; nb[u].y = d.y; nb[u].z = d.z;
        nb[u].vol = ra.vol; nb[u].por = ra.por; nb[u].porrho = ra.porrho; nb[u].pres = ra.pres;
        nb[u].ux = ra.ux; nb[u].uy = ra.uy; nb[u].uz = ra.uz;
        nb[u].pt = (int)(int16_t)(ra.flags & 0xFFFF); nb[u].it = (ra.flags >> 16) & 0xFF; nb[u].row = ra.flags >> 24;
      } else {
        pdiff(v, xi, yi, zi, j, nb[u].x, nb[u].y, nb[u].z);
        const double mj = v.m[j], rhoj = v.rho[j], porj = v.por[j];
        nb[u].vol = mj / rhoj; nb[u].por = porj; nb[u].porrho = porj * rhoj; nb[u].pres = v.pres[j];
        nb[u].ux = v.ux[j]; nb[u].uy = v.uy[j]; nb[u].uz = v.uz[j];
        nb[u].pt = v.ptype[j]; nb[u].it = v.itype[j]; nb[u].row = o.row[j];
      }
    }
    en4 = *(const uint4*)(v.nbr + nbr_slot(min(k0 + 4, last4), i, (int)n));
#pragma unroll
    for (int ub = 0; ub < 4; ub++) {
    if (k0 + ub >= cnt) break;
    const uint32_t e = eb[ub];
    const int j = (int)(e & ~kNbrFlag);
    const AsmNb& nq = nb[ub];
    if (e & kNbrFlag) continue;
    if (!(nq.pt < 1000 && nq.it != 4)) continue;
    double dx = nq.x, dy = nq.y, dz = nq.z;
    double r = sqrt(dx * dx + dy * dy + dz * dz) + 1e-20;
    // two reciprocals per pair (1/r, V_j = m_j / rho_j) in place of the
    // reference expressions' divisions; the rounding moves by an ulp
    const double ir = 1.0 / r;
    double dw = knl_dw_ih<double>(P.kernel_type, A, hd, ihd, r);
    const double sd = dw * ir;
    double gx = sd * dx, gy = sd * dy, gz = sd * dz;
    kgc_apply(P, v.cm, n, i, gx, gy, gz);
    const double porj = nq.por, Vj = nq.vol;
    bx += Vj * (pori * uxi - porj * (double)nq.ux) * gx;
    by += Vj * (pori * uyi - porj * (double)nq.uy) * gy;
    bz += Vj * (pori * uzi - porj * (double)nq.uz) * gz;
    double c = -(dx * gx + dy * gy + dz * gz) * (ir * ir);
    double a = (pori * rhoi + nq.porrho) * Vj * irhoi * c;
    Aij += a;
    AP += a * nq.pres;
    if (!nq.row) {
      lag += a * nq.pres;  // non-row column: lives in the right-hand side
      continue;
    }
    if (a == 0.0) continue;
    // store column kc: the quad's values collect in registers (oldest first)
    // and leave as whole 16-byte stores
    const int q = kc >> 2, u = kc & 3;
    uint32_t off8 = 0;
    if (FMT == kFmtRun8) {
      while (run < 9 && j >= pick9(E, run)) run++;
      int off = run < 9 ? j - pick9(B, run) : -1;
      if (off < 0) bad |= kFlagOrder;
      if (off > 255) bad |= kFlagOffset;
      if (run < 9) {
        if ((int)((rcw >> (kRunBits * run)) & kRunMax) == kRunMax) bad |= kFlagRunCount;
        else rcw += 1ull << (kRunBits * run);
      }
      off8 = (uint32_t)off & 255u;
      if (!MIXED) {
        word |= off8 << (8 * u);
        if (u == 3) {
          o.idx[(size_t)q * nn + i] = word;
          word = 0;
        }
      }
    }
    if (MIXED) {
      const uint32_t hb = hi16_bits(a);
      const double hv = (double)hi_val(hb);
      dcomp += a - hv;
      push_quad((uint32_t)j, hb | off8, (float)(a - hv), a);
    } else {
      push_quad((uint32_t)j, 0u, 0.f, a);
    }
    kc++;
    if ((kc & 3) == 0) store_quad(q);
    }
  }
  // pad the last quad so whole-quad loads read defined values
  if (kc & 3) {
    const int q = kc >> 2;
    for (int u = kc & 3; u < 4; u++) push_quad((uint32_t)i, 0u, 0.f, 0.0);
    store_quad(q);
    if (FMT == kFmtRun8 && !MIXED) o.idx[(size_t)q * nn + i] = word;
  }
  if (bad) atomicOr(o.flags, bad);
  o.ccnt[i] = kc;
  o.runc[i] = rcw;
  const double tdt = P.dt_sph;
  const double bi = P.rho_fluid * (bx + by + bz + 2 * (double)v.dpor[i] - (double)v.dpor_prev[i]) / tdt;
  v.ppe3[i] = bi;
  v.ppe4[i] = AP;
  v.ppe1[i] = rhoi * (bx + by + bz) / tdt;
  v.ppe2[i] = Aij;
  o.diag[i] = Aij;
  if (MIXED) o.dgi[i] = (float)(Aij - dcomp);
  o.dinv[i] = Aij != 0 ? 1.0 / Aij : 0.0;
  o.rhs[i] = bi + lag;
}

template <typename R>
__global__ void k_jacobi(SphView<R> v, const int* row) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= v.n) return;
  if (!row[i]) return;
  v.pres[i] = (v.ppe3[i] + v.ppe4[i]) / v.ppe2[i];
}

// ----- full-precision operator ----------------------------------------------
// y = A x (FP64 or hi + lo); residual form y = rhs - A x; optional ||y||^2
// partials (finished by k_fin with kFinNorm)
template <int FMT, int CM>
__global__ __launch_bounds__(kRedThreads) void k_apply_full(PpeOp o, const double* __restrict__ x,
                                                            double* __restrict__ y, int residual, int want_norm,
                                                            double* partials) {
  double d[1] = {0.0};
  for (RowIter it = xcd_rows(o.n); it.i < it.end; it.i += it.stride) {
    const int i = it.i;
    double acc = 0;
    if (o.row[i]) {
      const double s = row_sum<FMT, CM, double>(o, i, x);
      acc = o.diag[i] * x[i] - s;
      if (residual) acc = o.rhs[i] - acc;
    }
    y[i] = acc;
    d[0] += acc * acc;
  }
  if (want_norm) block_partials<1>(d, partials);
}

// ||rhs||^2 over rows (partials)
__global__ __launch_bounds__(kRedThreads) void k_rhs_norm(int n, const int* row, const double* rhs,
                                                          double* partials) {
  double d[1] = {0.0};
  for (RowIter it = xcd_rows(n); it.i < it.end; it.i += it.stride)
    if (row[it.i]) d[0] += rhs[it.i] * rhs[it.i];
  block_partials<1>(d, partials);
}

// number of stored entries (partials)
__global__ __launch_bounds__(kRedThreads) void k_count_nnz(int n, const int* ccnt, double* partials) {
  double d[1] = {0.0};
  for (RowIter it = xcd_rows(n); it.i < it.end; it.i += it.stride) d[0] += (double)ccnt[it.i];
  block_partials<1>(d, partials);
}

// ----- BiCGStab kernels ----------------------------------------------------------
// vectors (type V): r (recursive residual, also holds s), rh (shadow), p, v,
// ph = D^-1 p, sh = D^-1 s, t, xs (scaled correction since the last fold).
// Each reducing kernel writes per-block partials; k_fin then updates the
// device state in one block.

enum FinOp { kFinNorm = 0, kFinInit, kFinAlpha, kFinOmega, kFinRho, kFinTrue, kFinCount };

__device__ inline bool iter_live(const BiState* st) { return !st->done && !st->brk; }

template <typename V>
__global__ __launch_bounds__(kRedThreads) void k_bi_init(int n, const int* row, const double* r64, double inv_scale,
                                                         V* r, V* rh, V* p, V* v, V* xs, double* partials) {
  double d[1] = {0.0};
  for (RowIter it = xcd_rows(n); it.i < it.end; it.i += it.stride) {
    const int i = it.i;
    const V ri = row[i] ? (V)(r64[i] * inv_scale) : (V)0;
    r[i] = ri;
    rh[i] = ri;
    p[i] = 0;
    v[i] = 0;
    xs[i] = 0;
    d[0] += (double)ri * (double)ri;
  }
  block_partials<1>(d, partials);
}

// p = r + beta (p - omega v) (p = r on the first iteration); ph = D^-1 p
template <typename V>
__global__ __launch_bounds__(kRedThreads) void k_bi_dir(int n, const double* __restrict__ dinv,
                                                        const V* __restrict__ r, V* __restrict__ p,
                                                        const V* __restrict__ v, V* __restrict__ ph,
                                                        const BiState* st) {
  if (!iter_live(st)) return;
  const bool first = st->first != 0;
  const double beta = st->beta, omega = st->omega;
  for (RowIter it = xcd_rows(n); it.i < it.end; it.i += it.stride) {
    const int i = it.i;
    double pi = first ? (double)r[i] : (double)r[i] + beta * ((double)p[i] - omega * (double)v[i]);
    p[i] = (V)pi;
    ph[i] = (V)(pi * dinv[i]);
  }
}

// y = A x with the inner operator; partials (w, y), (y, y)
template <int FMT, int CM, typename V>
__global__ __launch_bounds__(kRedThreads) void k_bi_mv(PpeOp o, const V* __restrict__ x, V* __restrict__ y,
                                                       const V* __restrict__ w, double* partials,
                                                       const BiState* st) {
  if (!iter_live(st)) return;
  typedef typename std::conditional<CM == 0, float, double>::type AT;
  double d[2] = {0.0, 0.0};
  for (RowIter it = xcd_rows(o.n); it.i < it.end; it.i += it.stride) {
    const int i = it.i;
    AT acc = 0;
    if (o.row[i]) {
      const AT s = row_sum<FMT, CM, AT>(o, i, x);
      acc = (CM == 0 ? (AT)o.dgi[i] : (AT)o.diag[i]) * (AT)x[i] - s;
    }
    const V yi = (V)acc;
    y[i] = yi;
    d[0] += (double)w[i] * (double)yi;
    d[1] += (double)yi * (double)yi;
  }
  block_partials<2>(d, partials);
}

// ----- LDS-staged inner matvec (kFmtRun8) -------------------------------------
// A block takes its rows in tiles of kRedThreads consecutive particles.  The
// columns of a tile lie in nine particle ranges, one per (dy, dz) row of
// neighbour cells: cell keys [kmin - 1 + off, kmax + 1 + off] with
// off = NI (dy + NJ dz), contiguous in the cell-sorted order.  The tile copies
// those ranges of x into LDS with coalesced loads (about 11 values per row on
// a 1 mm column, against ~65 gathered entries per row) and its rows gather
// from LDS.  A tile whose ranges do not fit takes the global gathers.
constexpr int kStageCap = 4864;  // values: 19 KB (FP32) / 38 KB (FP64) per block

template <typename XT>
struct Stage {
  XT x[kStageCap];
  int lo[9];    // first staged particle of run row r
  int sz[9];    // staged particles of run row r
  int adj[9];   // LDS slot = particle index + adj[r]
  int total;
};

// all threads of the block; returns (uniformly) whether the tile is staged
template <typename XT>
__device__ bool stage_tile(const PpeOp& o, const XT* __restrict__ x, int i0, int i1, Stage<XT>& S) {
  __syncthreads();  // the previous tile is done with S
  if (threadIdx.x < 9) {
    const int r = threadIdx.x;
    const int off = o.NI * ((r % 3 - 1) + o.NJ * (r / 3 - 1));
    const uint32_t ka = o.keys[i0], kb = o.keys[i1 - 1];
    int a = 0, len = kStageCap + 1;  // a key outside the grid: no staging
    if (ka < (uint32_t)o.ncell && kb < (uint32_t)o.ncell) {
      const int klo = max(0, (int)ka - 1 + off), khi = min(o.ncell - 1, (int)kb + 1 + off);
      len = 0;
      if (klo <= khi) {
        a = (int)o.cell_off[klo];
        len = (int)o.cell_off[khi + 1] - a;
      }
    }
    S.lo[r] = a;
    S.sz[r] = len;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int acc = 0;
    for (int r = 0; r < 9; r++) {
      S.adj[r] = acc - S.lo[r];
      acc += S.sz[r];
    }
    S.total = acc;
  }
  __syncthreads();
  const int total = S.total;
  if (total > kStageCap) return false;
  int start[9];
#pragma unroll
  for (int r = 0; r < 9; r++) start[r] = S.lo[r] + S.adj[r];  // LDS slot of run row r's first value
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    int src = e - S.adj[0];
#pragma unroll
    for (int r = 1; r < 9; r++) src = e >= start[r] ? e - S.adj[r] : src;
    S.x[e] = x[src];
  }
  __syncthreads();
  return true;
}

template <int CM, typename V>
__global__ __launch_bounds__(kRedThreads, 8) void k_bi_mv_staged(PpeOp o, const V* __restrict__ x,
                                                              V* __restrict__ y, const V* __restrict__ w,
                                                              double* partials, const BiState* st) {
  if (!iter_live(st)) return;
  typedef typename std::conditional<CM == 0, float, double>::type AT;
  __shared__ Stage<V> S;
  double d[2] = {0.0, 0.0};
  const RowIter it = xcd_rows(o.n);
  const int t = (int)threadIdx.x;
  for (int i0 = it.i - t; i0 < it.end; i0 += it.stride) {
    const int i1 = min(i0 + (int)blockDim.x, it.end);
    const bool staged = stage_tile(o, x, i0, i1, S);
    const int i = i0 + t;
    if (i < i1) {
      AT acc = 0;
      if (o.row[i]) {
        const AT s = staged ? row_sum<kFmtRun8, CM, AT, V, true>(o, i, S.x, S.adj)
                            : row_sum<kFmtRun8, CM, AT, V>(o, i, x);
        acc = (CM == 0 ? (AT)o.dgi[i] : (AT)o.diag[i]) * (AT)x[i] - s;
      }
      const V yi = (V)acc;
      y[i] = yi;
      d[0] += (double)w[i] * (double)yi;
      d[1] += (double)yi * (double)yi;
    }
  }
  block_partials<2>(d, partials);
}

// The FP64 operator passes (y = A x or rhs - A x with an optional norm, and
// the true residual after a fold) with x64 staged in LDS per tile like the
// inner matvec (kFmtRun8).  Each thread sums the same rows in the same order
// as the unstaged kernels, so the bits are the same.
template <int CM>
__global__ __launch_bounds__(kRedThreads) void k_apply_staged(PpeOp o, const double* __restrict__ x,
                                                              double* __restrict__ y, int residual, int want_norm,
                                                              double* partials) {
  __shared__ Stage<double> S;
  double d[1] = {0.0};
  const RowIter it = xcd_rows(o.n);
  const int t = (int)threadIdx.x;
  for (int i0 = it.i - t; i0 < it.end; i0 += it.stride) {
    const int i1 = min(i0 + (int)blockDim.x, it.end);
    const bool staged = stage_tile(o, x, i0, i1, S);
    const int i = i0 + t;
    if (i < i1) {
      double acc = 0;
      if (o.row[i]) {
        const double sm = staged ? row_sum<kFmtRun8, CM, double, double, true>(o, i, S.x, S.adj)
                                 : row_sum<kFmtRun8, CM, double>(o, i, x);
        acc = o.diag[i] * x[i] - sm;
        if (residual) acc = o.rhs[i] - acc;
      }
      y[i] = acc;
      d[0] += acc * acc;
    }
  }
  if (want_norm) block_partials<1>(d, partials);
}

template <int CMF, typename V>
__global__ __launch_bounds__(kRedThreads) void k_bi_true_staged(PpeOp o, const double* __restrict__ x64,
                                                                double* __restrict__ r64, V* __restrict__ r,
                                                                const V* __restrict__ rh, double* partials,
                                                                const BiState* st) {
  if (st->done || !st->update) return;
  const double inv = 1.0 / st->xscale;
  __shared__ Stage<double> S;
  double d[3] = {0.0, 0.0, 0.0};
  const RowIter it = xcd_rows(o.n);
  const int t = (int)threadIdx.x;
  for (int i0 = it.i - t; i0 < it.end; i0 += it.stride) {
    const int i1 = min(i0 + (int)blockDim.x, it.end);
    const bool staged = stage_tile(o, x64, i0, i1, S);
    const int i = i0 + t;
    if (i < i1) {
      double acc = 0;
      if (o.row[i]) {
        const double sm = staged ? row_sum<kFmtRun8, CMF, double, double, true>(o, i, S.x, S.adj)
                                 : row_sum<kFmtRun8, CMF, double>(o, i, x64);
        acc = o.rhs[i] - (o.diag[i] * x64[i] - sm);
      }
      r64[i] = acc;
      const V ri = (V)(acc * inv);
      r[i] = ri;
      d[0] += (double)ri * (double)ri;
      d[1] += (double)rh[i] * (double)ri;
      d[2] += acc * acc;
    }
  }
  block_partials<3>(d, partials);
}

// s = r - alpha v (in place in r); sh = D^-1 s
template <typename V>
__global__ __launch_bounds__(kRedThreads) void k_bi_s(int n, const double* __restrict__ dinv, V* __restrict__ r,
                                                      const V* __restrict__ v, V* __restrict__ sh,
                                                      const BiState* st) {
  if (!iter_live(st)) return;
  const double alpha = st->alpha;
  for (RowIter it = xcd_rows(n); it.i < it.end; it.i += it.stride) {
    const int i = it.i;
    const double si = (double)r[i] - alpha * (double)v[i];
    r[i] = (V)si;
    sh[i] = (V)(si * dinv[i]);
  }
}

// xs += alpha ph + omega sh; r = s - omega t; partials (r, r), (rh, r)
template <typename V>
__global__ __launch_bounds__(kRedThreads) void k_bi_xr(int n, V* __restrict__ xs, const V* __restrict__ ph,
                                                       const V* __restrict__ sh, V* __restrict__ r,
                                                       const V* __restrict__ t, const V* __restrict__ rh,
                                                       double* partials, const BiState* st) {
  if (!iter_live(st)) return;
  const double alpha = st->alpha, omega = st->omega;
  double d[2] = {0.0, 0.0};
  for (RowIter it = xcd_rows(n); it.i < it.end; it.i += it.stride) {
    const int i = it.i;
    xs[i] = (V)((double)xs[i] + alpha * (double)ph[i] + omega * (double)sh[i]);
    const V ri = (V)((double)r[i] - omega * (double)t[i]);
    r[i] = ri;
    d[0] += (double)ri * (double)ri;
    d[1] += (double)rh[i] * (double)ri;
  }
  block_partials<2>(d, partials);
}

// fold: x64 += xscale * xs; xs = 0
template <typename V>
__global__ __launch_bounds__(kRedThreads) void k_bi_fold(int n, const int* row, double* __restrict__ x64,
                                                         V* __restrict__ xs, const BiState* st) {
  if (st->done || !st->update) return;
  const double sc = st->xscale;
  for (RowIter it = xcd_rows(n); it.i < it.end; it.i += it.stride) {
    const int i = it.i;
    if (row[i]) x64[i] += sc * (double)xs[i];
    xs[i] = 0;
  }
}

// true residual after a fold: r64 = rhs - A x64 (full precision),
// r = r64 / xscale; partials (r, r), (rh, r), (r64, r64)
template <int FMT, int CMF, typename V>
__global__ __launch_bounds__(kRedThreads) void k_bi_true(PpeOp o, const double* __restrict__ x64,
                                                         double* __restrict__ r64, V* __restrict__ r,
                                                         const V* __restrict__ rh, double* partials,
                                                         const BiState* st) {
  if (st->done || !st->update) return;
  const double inv = 1.0 / st->xscale;
  double d[3] = {0.0, 0.0, 0.0};
  for (RowIter it = xcd_rows(o.n); it.i < it.end; it.i += it.stride) {
    const int i = it.i;
    double acc = 0;
    if (o.row[i]) {
      const double s = row_sum<FMT, CMF, double>(o, i, x64);
      acc = o.rhs[i] - (o.diag[i] * x64[i] - s);
    }
    r64[i] = acc;
    const V ri = (V)(acc * inv);
    r[i] = ri;
    d[0] += (double)ri * (double)ri;
    d[1] += (double)rh[i] * (double)ri;
    d[2] += acc * acc;
  }
  block_partials<3>(d, partials);
}

// One block: sums the G partials of the preceding kernel and updates the
// state.  Skips exactly when the producer skipped.
__global__ __launch_bounds__(kFinThreads) void k_fin(const double* partials, int G, int op, BiState* st,
                                                     unsigned* count_out) {
  if ((op == kFinAlpha || op == kFinOmega || op == kFinRho) && !iter_live(st)) return;
  if (op == kFinTrue && (st->done || !st->update)) return;
  double tot[3];
  if (op == kFinTrue) {
    sum_partials<3>(partials, G, tot);
  } else if (op == kFinAlpha || op == kFinOmega || op == kFinRho) {
    double t2[2];
    sum_partials<2>(partials, G, t2);
    tot[0] = t2[0];
    tot[1] = t2[1];
  } else {
    double t1[1];
    sum_partials<1>(partials, G, t1);
    tot[0] = t1[0];
  }
  if (threadIdx.x != 0) return;
  switch (op) {
    case kFinNorm:
      st->norm = tot[0];
      break;
    case kFinCount:
      if (count_out) count_out[0] = (unsigned)tot[0];
      break;
    case kFinInit:
      st->rho = tot[0];
      st->rho_old = 1;
      st->alpha = 1;
      st->omega = 1;
      st->beta = 0;
      st->rr = tot[0];
      st->max_rr = tot[0];
      st->first = 1;
      st->done = 0;
      st->update = 0;
      st->stop = 0;
      st->brk = 0;
      break;
    case kFinAlpha: {
      const double rv = tot[0];
      if (rv != 0 && rv == rv) {
        st->alpha = st->rho / rv;
      } else {
        st->brk = 1;
        st->update = 1;
      }
      break;
    }
    case kFinOmega: {
      const double ts = tot[0], tt = tot[1];
      if (tt > 0 && ts != 0) {
        st->omega = ts / tt;
      } else {
        st->brk = 1;
        st->update = 1;
      }
      break;
    }
    case kFinRho: {
      const double rr = tot[0], rho = tot[1];
      const double alpha = st->alpha, omega = st->omega;
      st->rho_old = st->rho;
      st->rho = rho;
      st->beta = (st->rho_old != 0 && omega != 0) ? (rho / st->rho_old) * (alpha / omega) : 0;
      st->rr = rr;
      st->first = 0;
      st->iter += 1;
      if (rr > st->max_rr) st->max_rr = rr;
      bool upd = rr <= st->tol2 || rr < st->delta2 * st->max_rr;
      if (st->iter >= st->max_iter) {
        st->stop = 1;
        upd = true;
      }
      if (rho == 0 || !(rr == rr)) {
        st->brk = 1;
        upd = true;
      }
      st->update = upd ? 1 : 0;
      break;
    }
    case kFinTrue: {
      const double rho = tot[1];
      st->rr = tot[0];
      st->rr_true = tot[2];
      st->max_rr = tot[0];
      st->rho = rho;
      st->beta = (st->rho_old != 0 && st->omega != 0) ? (rho / st->rho_old) * (st->alpha / st->omega) : 0;
      st->update = 0;
      st->folds += 1;
      if (tot[2] <= st->tol2_true) st->done = 1;
      else if (st->stop) st->done = 3;
      else if (st->brk || rho == 0) st->done = 2;
      break;
    }
  }
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

// polynomial extrapolation of the pressure history on rows
__global__ void k_extrapolate(int n, const int* row, const double* p, const double* h1, const double* h2, int depth,
                              double* out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = p[i];
  if (row[i]) {
    if (depth >= 2) v = 3.0 * p[i] - 3.0 * h1[i] + h2[i];
    else if (depth == 1) v = 2.0 * p[i] - h1[i];
  }
  out[i] = v;
}

__global__ void k_zero_u32(unsigned* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

void launch_fin(hipStream_t s, const PpeWork& w, int op, unsigned* count_out = nullptr) {
  k_fin<<<1, kFinThreads, 0, s>>>(w.partials, w.nbg, op, (BiState*)w.state, count_out);
}

// host-side dispatch on (format, coefficient mode)
template <int CM>
void launch_apply_full(hipStream_t s, const PpeWork& w, const PpeOp& o, const double* x, double* y, bool residual,
                       bool norm) {
  if (o.fmt == kFmtRun8)
    k_apply_staged<CM><<<w.nbg, kRedThreads, 0, s>>>(o, x, y, residual, norm, w.partials);
  else
    k_apply_full<kFmtIdx32, CM><<<w.nbg, kRedThreads, 0, s>>>(o, x, y, residual, norm, w.partials);
  if (norm) launch_fin(s, w, kFinNorm);
}

int read_state(hipStream_t s, PpeWork& w, BiState* out) {
  HIP_TRY(hipMemcpyAsync(w.host_state, w.state, sizeof(BiState), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  std::memcpy(out, w.host_state, sizeof(BiState));
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
int ppe_workspace_alloc(PpeWork& w, int n, int cap, bool mixed) {
  w.n = n;
  w.nbg = grid_rows(n);
  w.cap4 = ((cap > 0 ? cap : 1) + 3) / 4 * 4;
  auto A = [&](void** p, size_t bytes) -> hipError_t { return hipMalloc(p, bytes ? bytes : 16); };
  const size_t N = (size_t)(n > 0 ? n : 1);
  const size_t E = N * (size_t)w.cap4;
  HIP_TRY(A((void**)&w.row, N * 4));
  HIP_TRY(A((void**)&w.ccnt, N * 4));
  HIP_TRY(A((void**)&w.runc, N * 8));
  HIP_TRY(A((void**)&w.idx, E * 4));  // sized for kFmtIdx32
  HIP_TRY(A((void**)&w.diag, N * 8));
  HIP_TRY(A((void**)&w.dinv, N * 8));
  HIP_TRY(A((void**)&w.rhs, N * 8));
  HIP_TRY(A((void**)&w.x64, N * 8));
  HIP_TRY(A((void**)&w.r64, N * 8));
  HIP_TRY(A((void**)&w.partials, (size_t)4 * kMaxGrid * 8));
  HIP_TRY(A(&w.state, sizeof(BiState)));
  HIP_TRY(hipMemset(w.state, 0, sizeof(BiState)));
  HIP_TRY(hipHostMalloc(&w.host_state, sizeof(BiState), hipHostMallocDefault));
  HIP_TRY(A((void**)&w.flags, 16));
  HIP_TRY(hipHostMalloc((void**)&w.host_flags, 16, hipHostMallocDefault));
  w.vec_bytes = mixed ? 4 : 8;
  for (int k = 0; k < 8; k++) HIP_TRY(A(&w.vec[k], N * w.vec_bytes));
  w.mixed = mixed;
  if (mixed) {
    HIP_TRY(A((void**)&w.hi, E * 4));
    HIP_TRY(A((void**)&w.lo, E * 4));
    HIP_TRY(A((void**)&w.dgi, N * 4));
    w.c64 = nullptr;
  } else {
    HIP_TRY(A((void**)&w.c64, E * 8));
    w.hi = w.lo = w.dgi = nullptr;
  }
  return 0;
}

void ppe_workspace_free(PpeWork& w) {
  auto F = [](void* p) {
    if (p) (void)hipFree(p);
  };
  F(w.row); F(w.ccnt); F(w.runc); F(w.idx); F(w.diag); F(w.dinv); F(w.rhs); F(w.x64); F(w.r64);
  F(w.partials); F(w.state); F(w.flags);
  if (w.host_state) (void)hipHostFree(w.host_state);
  if (w.host_flags) (void)hipHostFree(w.host_flags);
  for (int k = 0; k < 8; k++) F(w.vec[k]);
  F(w.hi); F(w.lo); F(w.dgi); F(w.c64);
  if (w.coarse) {
    coarse_free(*w.coarse);
    delete w.coarse;
  }
  w = PpeWork{};
}

PpeOp ppe_op(const PpeWork& w) {
  PpeOp o;
  o.n = w.n;
  o.nq = w.cap4 / 4;
  o.fmt = w.fmt;
  o.NI = w.NI;
  o.NJ = w.NJ;
  o.ncell = w.ncell;
  o.keys = w.keys;
  o.cell_off = w.cell_off;
  o.row = w.row;
  o.ccnt = w.ccnt;
  o.runc = w.runc;
  o.idx = w.idx;
  o.hi = w.hi;
  o.lo = w.lo;
  o.c64 = w.c64;
  o.diag = w.diag;
  o.dgi = w.dgi;
  o.dinv = w.dinv;
  o.rhs = w.rhs;
  return o;
}

int cell_offsets(hipStream_t s, int n, const uint32_t* keys, int num_cells, uint32_t* off) {
  k_cell_offsets<<<grid_for(num_cells + 1), kBlock, 0, s>>>(n, keys, num_cells, off);
  HIP_TRY(hipGetLastError());
  return 0;
}

template <typename R>
int ppe_assemble(hipStream_t s, const DevParams& P, const SphView<R>& v, PpeWork& w, const uint32_t* keys,
                 const uint32_t* cell_off) {
  if (v.n <= 0) return 0;
  if (v.nbr_cap > w.cap4) {
    set_error("PPE workspace holds %d slots per row, neighbour list has %d", w.cap4, v.nbr_cap);
    return SOPHIA_E_STATE;
  }
  w.keys = keys;
  w.cell_off = cell_off;
  w.bn = -1;
  w.gen++;
  w.NI = P.NI;
  w.NJ = P.NJ;
  w.ncell = P.NI * P.NJ * P.NK;
  w.rowmajor = P.z_index == 0;
  AsmOut o{v.n, P.NI, P.NJ, P.NK, keys, cell_off, w.row, w.ccnt, w.runc, w.idx, w.hi, w.lo, w.c64,
           w.diag, w.dinv, w.rhs, w.dgi, w.flags};
  const int g = grid_xcd(v.n);
  k_rowmask<<<g, kBlock, 0, s>>>(P, v.n, v.ptype, v.btype, v.itype, v.surf, v.pres, w.row);
  if constexpr (std::is_same<R, float>::value) k_pack_asm<<<grid_for(v.n), kBlock, 0, s>>>(v, w.row);
  // the x-runs of the stencil are contiguous only in row-major keys without
  // periodic wrap
  int fmt = (P.z_index || P.pmask) ? kFmtIdx32 : kFmtRun8;
  for (;;) {
    k_zero_u32<<<1, 64, 0, s>>>(w.flags, 4);
    if (fmt == kFmtRun8) {
      if (w.mixed) k_assemble<R, kFmtRun8, 1><<<g, kBlock, 0, s>>>(P, v, o);
      else k_assemble<R, kFmtRun8, 0><<<g, kBlock, 0, s>>>(P, v, o);
    } else {
      if (w.mixed) k_assemble<R, kFmtIdx32, 1><<<g, kBlock, 0, s>>>(P, v, o);
      else k_assemble<R, kFmtIdx32, 0><<<g, kBlock, 0, s>>>(P, v, o);
    }
    k_count_nnz<<<w.nbg, kRedThreads, 0, s>>>(v.n, w.ccnt, w.partials);
    launch_fin(s, w, kFinCount, w.flags + 1);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(w.host_flags, w.flags, 8, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    const unsigned f = w.host_flags[0];
    if (fmt == kFmtRun8 && (f & (kFlagOffset | kFlagRunCount))) {
      fmt = kFmtIdx32;  // a run longer than 256 particles: plain indices
      continue;
    }
    if (f & kFlagOrder) {
      set_error("PPE assembly: neighbour list not in cell-run order");
      return SOPHIA_E_STATE;
    }
    break;
  }
  w.fmt = fmt;
  w.nnz = (int64_t)w.host_flags[1];
  return 0;
}

template <typename R>
int ppe_jacobi(hipStream_t s, const SphView<R>& v, PpeWork& w) {
  if (v.n <= 0) return 0;
  k_jacobi<R><<<grid_for(v.n), kBlock, 0, s>>>(v, w.row);
  HIP_TRY(hipGetLastError());
  return 0;
}

int ppe_apply64(hipStream_t s, PpeWork& w, const double* x, double* y, bool residual, double* norm2) {
  const int n = w.n;
  if (n <= 0) {
    if (norm2) *norm2 = 0;
    return 0;
  }
  PpeOp o = ppe_op(w);
  if (w.mixed) launch_apply_full<1>(s, w, o, x, y, residual, norm2 != nullptr);
  else launch_apply_full<2>(s, w, o, x, y, residual, norm2 != nullptr);
  HIP_TRY(hipGetLastError());
  if (norm2) {
    BiState h;
    TRY(read_state(s, w, &h));
    *norm2 = h.norm;
  }
  return 0;
}

// v -= (mean of v over the rows): the projection that makes the right-hand
// side of a periodic operator without walls (constant null space, A
// symmetric without KGC) consistent, and fixes the level of its solution.
// Grid-stride block partials of (sum, count), summed by one block in a
// fixed order, then subtracted.
__global__ __launch_bounds__(kBlock) void k_mean_part(int n, const int* row, const double* v, double* part) {
  double a = 0, c = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (row[i]) {
      a += v[i];
      c += 1.0;
    }
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    c += __shfl_xor(c, off);
  }
  __shared__ double sa[kBlock / 64], sc[kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    sa[wv] = a;
    sc[wv] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double ta = 0, tc = 0;
    for (int k = 0; k < kBlock / 64; k++) {
      ta += sa[k];
      tc += sc[k];
    }
    part[2 * blockIdx.x] = ta;
    part[2 * blockIdx.x + 1] = tc;
  }
}
__global__ void k_mean_fin(const double* part, int nb, double* mean) {
  if (threadIdx.x != 0) return;
  double ta = 0, tc = 0;
  for (int b = 0; b < nb; b++) {
    ta += part[2 * b];
    tc += part[2 * b + 1];
  }
  *mean = tc > 0 ? ta / tc : 0.0;
}
__global__ void k_mean_sub(int n, const int* row, double* v, const double* mean) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && row[i]) v[i] -= *mean;
}
int ppe_project_mean(hipStream_t s, PpeWork& w, double* v) {
  if (w.n <= 0) return 0;
  const int nb = std::min(grid_for(w.n), kMaxGrid);
  k_mean_part<<<nb, kBlock, 0, s>>>(w.n, w.row, v, w.partials);
  k_mean_fin<<<1, 64, 0, s>>>(w.partials, nb, w.partials + 2 * kMaxGrid);
  k_mean_sub<<<grid_for(w.n), kBlock, 0, s>>>(w.n, w.row, v, w.partials + 2 * kMaxGrid);
  HIP_TRY(hipGetLastError());
  if (v == w.rhs) w.bn = -1;  // the cached norm belongs to the old rhs
  return 0;
}

int ppe_rhs_norm(hipStream_t s, PpeWork& w, double* bn) {
  if (w.n <= 0) {
    *bn = 0;
    return 0;
  }
  if (w.bn >= 0) {  // the right-hand side is fixed from assembly to assembly
    *bn = w.bn;
    return 0;
  }
  k_rhs_norm<<<w.nbg, kRedThreads, 0, s>>>(w.n, w.row, w.rhs, w.partials);
  launch_fin(s, w, kFinNorm);
  HIP_TRY(hipGetLastError());
  BiState h;
  TRY(read_state(s, w, &h));
  *bn = sqrt(h.norm);
  w.bn = *bn;
  return 0;
}

int ppe_extrapolate(hipStream_t s, const PpeWork& w, const double* p, const double* h1, const double* h2, int depth,
                    double* out) {
  if (w.n <= 0) return 0;
  k_extrapolate<<<grid_for(w.n), kBlock, 0, s>>>(w.n, w.row, p, h1, h2, depth, out);
  HIP_TRY(hipGetLastError());
  return 0;
}

int ppe_residual_of(hipStream_t s, PpeWork& w, const double* x, double* rel, double* rr_out) {
  double bn = 0, rr = 0;
  TRY(ppe_rhs_norm(s, w, &bn));
  TRY(ppe_apply64(s, w, x, w.r64, true, &rr));
  *rel = sqrt(rr) / (bn == 0 ? 1 : bn);
  if (rr_out) *rr_out = rr;
  return 0;
}

namespace {

// One BiCGStab run from the residual in w.r64 (norm rn), folding into w.x64.
// V: vector type; CM: inner operator; CMF: full-precision operator.
template <typename V, int FMT, int CM, int CMF>
int bicgstab_run(hipStream_t s, PpeWork& w, const CoarseWork* cw, double rn, double tol, double delta, int max_iter,
                 int* iters, int* folds, int* done, Timers* tm) {
  const int n = w.n, G = w.nbg;
  PpeOp o = ppe_op(w);
  BiState* st = (BiState*)w.state;
  V* r = (V*)w.vec[0];
  V* rh = (V*)w.vec[1];
  V* p = (V*)w.vec[2];
  V* v = (V*)w.vec[3];
  V* ph = (V*)w.vec[4];
  V* sh = (V*)w.vec[5];
  V* t = (V*)w.vec[6];
  V* xs = (V*)w.vec[7];
  {
    BiState h{};
    h.xscale = rn;
    h.tol2 = (tol / rn) * (tol / rn);
    h.tol2_true = tol * tol;
    h.delta2 = delta * delta;
    h.max_iter = max_iter;
    std::memcpy(w.host_state, &h, sizeof(h));
    HIP_TRY(hipMemcpyAsync(st, w.host_state, sizeof(BiState), hipMemcpyHostToDevice, s));
  }
  k_bi_init<V><<<G, kRedThreads, 0, s>>>(n, w.row, w.r64, 1.0 / rn, r, rh, p, v, xs, w.partials);
  launch_fin(s, w, kFinInit);
  HIP_TRY(hipGetLastError());
  const char* mv_name = sizeof(V) == 4 ? "ppe_matvec_fp32" : "ppe_matvec_fp64";
  const char* mv_all = sizeof(V) == 4 ? "ppe_matvec_fp32.launches" : "ppe_matvec_fp64.launches";
  // the inner matvec is timed with events on one launch in kMvSample (each
  // event costs time on the stream); every launch is counted
  auto mv_timer = [&]() -> Timers* {
    if (!tm || !tm->on) return nullptr;
    tm->count(mv_all);
    return (w.mv_seq++ % kMvSample) == 0 ? tm : nullptr;
  };
  int chunk = 8;
  BiState h{};
  double rr0 = 1.0;
  for (;;) {
    for (int c = 0; c < chunk; c++) {
      if (cw) {
        const CoarseUpd<V> u{w.dinv, r, p, v, ph, &st->alpha, &st->beta, &st->omega, &st->first};
        TRY(coarse_apply_update<V>(s, *cw, w, u, 1, &st->done, &st->brk));
      } else {
        k_bi_dir<V><<<G, kRedThreads, 0, s>>>(n, w.dinv, r, p, v, ph, st);
      }
      {
        TimeScope ts(mv_timer(), mv_name);
        if (FMT == kFmtRun8) k_bi_mv_staged<CM, V><<<G, kRedThreads, 0, s>>>(o, ph, v, rh, w.partials, st);
        else k_bi_mv<FMT, CM, V><<<G, kRedThreads, 0, s>>>(o, ph, v, rh, w.partials, st);
      }
      launch_fin(s, w, kFinAlpha);
      if (cw) {
        const CoarseUpd<V> u{w.dinv, r, p, v, sh, &st->alpha, &st->beta, &st->omega, &st->first};
        TRY(coarse_apply_update<V>(s, *cw, w, u, 2, &st->done, &st->brk));
      } else {
        k_bi_s<V><<<G, kRedThreads, 0, s>>>(n, w.dinv, r, v, sh, st);
      }
      {
        TimeScope ts(mv_timer(), mv_name);
        if (FMT == kFmtRun8) k_bi_mv_staged<CM, V><<<G, kRedThreads, 0, s>>>(o, sh, t, r, w.partials, st);
        else k_bi_mv<FMT, CM, V><<<G, kRedThreads, 0, s>>>(o, sh, t, r, w.partials, st);
      }
      launch_fin(s, w, kFinOmega);
      k_bi_xr<V><<<G, kRedThreads, 0, s>>>(n, xs, ph, sh, r, t, rh, w.partials, st);
      launch_fin(s, w, kFinRho);
      k_bi_fold<V><<<G, kRedThreads, 0, s>>>(n, w.row, w.x64, xs, st);
      if (FMT == kFmtRun8) k_bi_true_staged<CMF, V><<<G, kRedThreads, 0, s>>>(o, w.x64, w.r64, r, rh, w.partials, st);
      else k_bi_true<FMT, CMF, V><<<G, kRedThreads, 0, s>>>(o, w.x64, w.r64, r, rh, w.partials, st);
      launch_fin(s, w, kFinTrue);
    }
    HIP_TRY(hipGetLastError());
    TRY(read_state(s, w, &h));
    if (h.done) break;
    // next chunk: about the iterations still needed at the observed rate
    const double rr = h.rr > 0 ? h.rr : 1e-300;
    if (h.iter > 0 && rr < rr0 && h.tol2 > 0) {
      const double rate = log(rr / rr0) / h.iter;  // < 0
      const double need = log(h.tol2 / rr) / rate;
      int c = (int)(need * 0.9);
      chunk = c < 2 ? 2 : (c > 32 ? 32 : c);
    } else {
      chunk = 8;
    }
  }
  *iters = h.iter;
  *folds = h.folds;
  *done = h.done;
  return 0;
}

template <typename V, int CM, int CMF>
int bicgstab_dispatch(hipStream_t s, PpeWork& w, const CoarseWork* cw, double rn, double tol, double delta,
                      int max_iter, int* iters, int* folds, int* done, Timers* tm) {
  if (w.fmt == kFmtRun8)
    return bicgstab_run<V, kFmtRun8, CM, CMF>(s, w, cw, rn, tol, delta, max_iter, iters, folds, done, tm);
  return bicgstab_run<V, kFmtIdx32, CM, CMF>(s, w, cw, rn, tol, delta, max_iter, iters, folds, done, tm);
}

}  // namespace

int ppe_solve(hipStream_t s, PpeWork& w, double* pres, const sophia_ppe_opts& o, sophia_ppe_stats& st,
              const double* guess_or_null, Timers* tm, double known_rr) {
  const int n = w.n;
  if (n <= 0) return 0;
  const int g = grid_for(n);
  if (guess_or_null) {
    HIP_TRY(hipMemcpyAsync(w.x64, guess_or_null, (size_t)n * 8, hipMemcpyDeviceToDevice, s));
  } else {
    k_gather_rows_guess<<<g, kBlock, 0, s>>>(n, w.row, pres, w.x64, o.guess == SOPHIA_GUESS_ZERO);
  }
  double bn = 0;
  TRY(ppe_rhs_norm(s, w, &bn));
  if (bn == 0) bn = 1;
  double rr = known_rr;
  if (!(guess_or_null && known_rr >= 0)) TRY(ppe_apply64(s, w, w.x64, w.r64, true, &rr));
  st.r0_rel = sqrt(rr) / bn;
  const double tol = o.rtol * bn;
  // reliable-update threshold: FP32 inner iterations fold into the FP64
  // solution each time their residual drops by this factor; the FP64 path
  // folds only when it believes it has converged
  const double delta = w.mixed ? (o.inner_rtol > 0 && o.inner_rtol < 1 ? o.inner_rtol : 0.01) : 0.0;
  const int max_restart = o.max_outer > 0 ? o.max_outer : 20;
  // two-level preconditioner: coarse operator and inverse for this operator
  const CoarseWork* cw = nullptr;
  st.precond_used = SOPHIA_PRECOND_JACOBI;
  st.n_coarse = 0;
  st.coarse_res = 0;
  if (o.precond == SOPHIA_PRECOND_TWO_LEVEL && o.zero_mean == 0 && sqrt(rr) > tol) {
    if (!w.coarse) w.coarse = new CoarseWork();
    bool usable = false;
    TRY(coarse_setup(s, *w.coarse, w, &usable, tm));
    st.n_coarse = w.coarse->na;
    st.coarse_res = w.coarse->res;
    if (usable) {
      cw = w.coarse;
      st.precond_used = SOPHIA_PRECOND_TWO_LEVEL;
    }
  }
  int total = 0, folds = 0, restarts = 0;
  while (sqrt(rr) > tol && total < o.max_iter && restarts < max_restart) {
    int it = 0, fo = 0, done = 0;
    const int rc = w.mixed
        ? bicgstab_dispatch<float, 0, 1>(s, w, cw, sqrt(rr), tol, delta, o.max_iter - total, &it, &fo, &done, tm)
        : bicgstab_dispatch<double, 2, 2>(s, w, cw, sqrt(rr), tol, delta, o.max_iter - total, &it, &fo, &done, tm);
    if (rc) return rc;
    total += it;
    folds += fo;
    restarts++;
    BiState h;
    TRY(read_state(s, w, &h));
    rr = h.rr_true;
    if (!(rr == rr)) {
      set_error("BiCGStab breakdown (NaN residual) after %d iterations", total);
      return SOPHIA_E_NUMERIC;
    }
    if (done != 2) break;  // converged or out of iterations; 2 = breakdown -> restart
    if (it == 0) break;
  }
  // rr is the full-precision residual of x64 as it stands: every BiCGStab
  // run ends with a fold followed by the FP64 residual of the folded x64
  // (k_bi_true), and without a run it is the residual of the guess
  st.r_rel = sqrt(rr) / bn;
  st.iterations = total;
  st.outer = folds;
  st.converged = st.r_rel <= o.rtol * (1 + 1e-12);
  if (o.zero_mean) TRY(ppe_project_mean(s, w, w.x64));  // the level of a singular system
  k_scatter_rows<<<g, kBlock, 0, s>>>(n, w.row, w.x64, pres);
  HIP_TRY(hipGetLastError());
  return 0;
}

// --- matrix-free matvec (timing probe) -----------------------------------
// y_i = d_i x_i - sum_{j row} a_ij x_j with a_ij = c_ij (eps_i V_j + m_j eps_j
// / rho_i), c_ij = -(x_ij . Cm_i grad W_ij) / r_ij^2, recomputed per launch
// in FP32 from the neighbour list, the position records and one float2 per
// column (V_j, m_j eps_j; zero for non-row columns).  Probe only: it reads
// the 4-byte neighbour list where the stored operator reads 5 bytes per
// entry, and adds the kernel gradient per pair.
template <typename R>
__global__ void k_mf_pack(int n, const SphView<R> v, const int* row, float2* pk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  pk[i] = row[i] ? make_float2((float)(v.m[i] / v.rho[i]), (float)(v.m[i] * v.por[i])) : make_float2(0.f, 0.f);
}

template <typename R>
__global__ __launch_bounds__(kBlock) void k_mf_matvec(DevParams P, SphView<R> v, const int* __restrict__ row,
                                                      const double* __restrict__ diag, const float2* __restrict__ pk,
                                                      const float* __restrict__ x, float* __restrict__ y) {
  const int i = particle_index();
  if (i >= v.n) return;
  if (!row[i]) {
    y[i] = 0.f;
    return;
  }
  const int n = v.n;
  const float h = (float)v.h[i], A = knl_tmpA<float>(P.kernel_type, P.dim, h);
  const float epsi = (float)v.por[i], irhoi = (float)(1.0 / v.rho[i]);
  float acc = 0.f;
  for_nbrs4<true>(v, i, true, [&](const Nb<R>& q, bool on) {
    const float dx = (float)q.dx, dy = (float)q.dy, dz = (float)q.dz;
    const float r = sqrtf(dx * dx + dy * dy + dz * dz) + 1e-20f;
    const float dw = knl_dw<float>(P.kernel_type, A, h, r);
    float gx = dw * dx / r, gy = dw * dy / r, gz = dw * dz / r;
    kgc_apply(P, v.cm, n, i, gx, gy, gz);
    const float c = -(dx * gx + dy * gy + dz * gz) / (r * r);
    const float2 pj = pk[q.j];
    const float a = c * (epsi * pj.x + pj.y * irhoi);
    acc += on ? a * x[q.j] : 0.f;
  });
  y[i] = (float)diag[i] * x[i] - acc;
}

__global__ void k_to_f32(int n, const double* __restrict__ a, float* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}
__global__ void k_to_f64(int n, const float* __restrict__ a, double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

template <typename R>
int ppe_matvec_probe(hipStream_t s, const DevParams& P, const SphView<R>& v, PpeWork& w, int variant, int reps,
                     const double* x64, double* y64, float2* pk, float* ms) {
  if (!w.mixed || w.fmt != kFmtRun8) {
    set_error("matvec probe: needs the mixed-precision operator with run offsets");
    return SOPHIA_E_STATE;
  }
  const int n = w.n;
  float* x = (float*)w.vec[4];
  float* y = (float*)w.vec[3];
  const float* rh = (const float*)w.vec[1];
  k_to_f32<<<grid_for(n), kBlock, 0, s>>>(n, x64, x);
  BiState* st = (BiState*)w.state;
  if (variant == 1) {
    k_mf_pack<R><<<grid_for(n), kBlock, 0, s>>>(n, v, w.row, pk);
  } else {
    // a live solver state: no convergence, no breakdown
    BiState h{};
    h.max_iter = 1 << 30;
    std::memcpy(w.host_state, &h, sizeof(h));
    HIP_TRY(hipMemcpyAsync(st, w.host_state, sizeof(BiState), hipMemcpyHostToDevice, s));
  }
  PpeOp o = ppe_op(w);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  for (int k = 0; k <= reps; k++) {
    if (k == 1) HIP_TRY(hipEventRecord(e0, s));  // the first launch warms the caches, untimed
    if (variant == 1) k_mf_matvec<R><<<grid_xcd(n), kBlock, 0, s>>>(P, v, w.row, w.diag, pk, x, y);
    else k_bi_mv_staged<0, float><<<w.nbg, kRedThreads, 0, s>>>(o, x, y, rh, w.partials, st);
  }
  HIP_TRY(hipEventRecord(e1, s));
  HIP_TRY(hipEventSynchronize(e1));
  float t = 0.f;
  HIP_TRY(hipEventElapsedTime(&t, e0, e1));
  *ms = reps > 0 ? t / reps : 0.f;
  HIP_TRY(hipEventDestroy(e0));
  HIP_TRY(hipEventDestroy(e1));
  k_to_f64<<<grid_for(n), kBlock, 0, s>>>(n, y, y64);
  HIP_TRY(hipStreamSynchronize(s));
  return 0;
}

template int ppe_assemble<float>(hipStream_t, const DevParams&, const SphView<float>&, PpeWork&, const uint32_t*,
                                 const uint32_t*);
template int ppe_assemble<double>(hipStream_t, const DevParams&, const SphView<double>&, PpeWork&, const uint32_t*,
                                  const uint32_t*);
template int ppe_jacobi<float>(hipStream_t, const SphView<float>&, PpeWork&);
template int ppe_jacobi<double>(hipStream_t, const SphView<double>&, PpeWork&);

template int ppe_matvec_probe<float>(hipStream_t, const DevParams&, const SphView<float>&, PpeWork&, int, int,
                                     const double*, double*, float2*, float*);
template int ppe_matvec_probe<double>(hipStream_t, const DevParams&, const SphView<double>&, PpeWork&, int, int,
                                      const double*, double*, float2*, float*);

}  // namespace sophia
