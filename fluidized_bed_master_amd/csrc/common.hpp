// Shared declarations of the HIP SPH-DEM library (gfx950).
//
// Layout in HBM: every particle field is its own contiguous stream
// (structure of arrays) in the cell-sorted order produced by NNPS, so a
// wavefront reading field f of 64 consecutive particles issues one
// coalesced 256 B (FP32) or 512 B (FP64) request.  Positions, smoothing
// lengths and pressure are FP64 on every path (cell keys, neighbour
// membership and the pressure solve are decided in FP64), and so are mass,
// density and porosity: pressure-gradient sums sum_j V_j (p_i + p_j) gradW
// cancel a hydrostatic p of order H/h times the result, which amplifies a
// per-particle FP32 rounding of V_j = m_j/rho_j past 1e-5.  The remaining
// SPH fields use the storage type R (float on the mixed path, double on the
// FP64 path).  DEM spheres are few (<= 3% of particles) and are kept in
// FP64 on both paths so contact detection and shear history match the
// oracle exactly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/sophia/abi.h"

namespace sophia {

constexpr uint32_t kEmpty = 0xffffffffu;  // empty cell marker
constexpr int kMaxHist = 16;              // DEM shear-history slots
constexpr int kBlock = 256;               // default workgroup size (4 waves)
constexpr uint32_t kNbrFlag = 0x80000000u;  // set on list entries that fail r+1e-20 < R
// The SPH neighbour list is stored in chunks of 8 consecutive slots of one
// particle (32 bytes): slot k of particle i of n sits at nbr_slot(k, i, n).
// The builder writes whole chunks, so no 32-byte sector is written piecemeal,
// and pads a particle's last chunk with its last entry: a 16-byte load of
// slots 4q..4q+3 with 4q < count reads valid particle indices only.
constexpr int kNbrChunk = 8;
__host__ __device__ inline size_t nbr_slot(int k, int i, int n) {
  return (((size_t)(k >> 3) * (size_t)n + (size_t)i) << 3) + (size_t)(k & 7);
}

// error reporting
void set_error(const char* fmt, ...);
int hip_fail(hipError_t e, const char* what, const char* file, int line);
#define HIP_TRY(call)                                                    \
  do {                                                                   \
    hipError_t e__ = (call);                                             \
    if (e__ != hipSuccess) return hip_fail(e__, #call, __FILE__, __LINE__); \
  } while (0)
// propagate a non-zero status code
#define TRY(x)             \
  do {                     \
    int rc__ = (x);        \
    if (rc__) return rc__; \
  } while (0)

// Parameters passed by value to every kernel.
struct DevParams {
  int dim, kernel_type, kgc, delsph, rho_type;
  int fp_solve, fv_solve, fva_solve, fg_solve, con_solve, xsph_solve;
  int open_boundary, z_index, eulerian, fix_xsph, pgrad_form;
  int NI, NJ, NK, num_cells;
  int dem_slots;
  int filt_step;          // 1 when the Shepard filter is due this step
  int first_step;         // time == 0
  double x_min, x_max, y_min, y_max, z_min, z_max;
  double kappa, dt_sph, dt_dem, u_limit, c_xsph, soundspeed, gravity;
  double rho_fluid, rho_water, h_over_dx, mu_fluid, mu_water, k_fluid, k_dem, cp_fluid, cp_dem;
  double dem_G, dem_nu, dem_friction, dem_damping, dem_hertz, drag_switch;
  double eos_b, gamma, delta_sph, p_ref;
  int compat;             // sophia_params.reference_compat bits
  int free_surface;
  double fs_threshold;
  double pr_fluid, pr_air, dem_k_mat[2], dem_alpha_mat[2], dem_mat_radius;
  double heat_flux;       // W/m^2 through flux walls (0 = off)
  double heat_flux_box[6];
  // property tables (prop_table): per |p_type| < tab_types the block
  // tab[k*50 + tab_index[t] ...] of size tab_size[t], k = T, h, k, mu
  int wcsph;              // solver_type 0: the WCSPH variants of the coupling kernels
  int noslip;             // 2D WCSPH: KERNEL_boundary2D velocity mirroring
  int pmask;              // periodic axes (bit 0 x, 1 y, 2 z)
  double plen[3];         // their periods
  int prop_table, tab_types;
  int tab_index[SOPHIA_TABLE_TYPES], tab_size[SOPHIA_TABLE_TYPES];
  const double* tab;
};

// Piecewise-linear interpolation of ys over the ascending xs (n >= 2),
// extended linearly beyond both ends with the end segments: the reference's
// interp2 (functions_PROP.cuh:143-175), evaluated in its operation order
// ((y1 - y0) / (x1 - x0) * (x - x0) + y0) so that host and device agree.
__host__ __device__ inline double interp_table(const double* xs, const double* ys, int n, double x) {
  int a = 0;
  if (xs[n - 1] < x) {
    a = n - 2;
  } else if (!(x <= xs[0])) {
    a = n - 2;
    for (int i = 0; i < n - 1; i++)
      if (xs[i] < x && x <= xs[i + 1]) {
        a = i;
        break;
      }
  }
  double y = ys[a + 1] - ys[a];
  y /= (xs[a + 1] - xs[a]);
  y *= (x - xs[a]);
  y += ys[a];
  return y;
}

// viscosity and conductivity of an SPH particle: the table of its |p_type|
// when property tables are on and it has one, else the constants
__device__ inline double prop_mu(const DevParams& P, int pt, double T) {
  const int t = pt < 0 ? -pt : pt;
  if (P.prop_table && t < P.tab_types && P.tab_size[t] >= 2)
    return interp_table(P.tab + P.tab_index[t], P.tab + 3 * SOPHIA_TABLE_VALUES + P.tab_index[t], P.tab_size[t], T);
  return P.mu_fluid;
}
__device__ inline double prop_k(const DevParams& P, int pt, double T) {
  const int t = pt < 0 ? -pt : pt;
  if (P.prop_table && t < P.tab_types && P.tab_size[t] >= 2)
    return interp_table(P.tab + P.tab_index[t], P.tab + 2 * SOPHIA_TABLE_VALUES + P.tab_index[t], P.tab_size[t], T);
  return t > 1000 ? P.k_dem : P.k_fluid;
}

// ---------------------------------------------------------------------------
// Device state.  Persistent fields travel with the particles through the
// NNPS permutation; scratch fields are recomputed every step.
// ---------------------------------------------------------------------------
template <typename R> struct Vec4Of;
template <> struct Vec4Of<float> { typedef float4 type; };
template <> struct Vec4Of<double> { typedef double4 type; };
template <typename R> struct alignas(2 * sizeof(R)) HRec {
  R h;
  int pt;
};

// Packed neighbour records of the FP32 pair kernels (mixed path), rebuilt
// before each kernel that reads them: one 32 B record per particle (a single
// cache-line gather) in place of four to five field gathers.  aux: rho/rho_ref
// - 1 for the prep sums, else 0.
struct alignas(32) Rec32 {
  float4 p;  // pos4
  float m, rho;
  int pt;
  float aux;
};
// PPE assembly (mixed path): V = m/rho, eps, eps rho and p in FP64 with the
// FP32 velocity; flags = p_type (low 16 bits) | i_type << 16 | row << 24.
struct alignas(64) RecAsm {
  float4 p;  // pos4
  double vol, por, porrho, pres;
  float ux, uy, uz;
  int flags;
};

// Advection's neighbour records on the mixed path, after pk = {x, y, z, m}:
// {rho, p, u, T} and {S (the flux-wall sum), h, p_type}
struct alignas(32) PkB {
  double rho, p;
  float ux, uy, uz, temp;
};
struct alignas(16) PkC {
  double S;
  float h;
  int pt;
};

// Neumann walls (mixed path): a candidate's pos4, and for those within
// range V = m/rho (FP32), p and rho/eps (FP64), T and p_type, on one line.
struct alignas(64) RecNeu {
  float4 p;
  double pres, rhopor;
  float vol, temp;
  int pt;
};

template <typename R>
struct SphView {
  int n;
  // persistent
  double *x, *y, *z, *h, *pres, *m, *rho;
  R *ux, *uy, *uz, *temp, *flt_s, *dpor;
  int *ptype, *btype, *itype;
  // scratch
  double* por;  // porosity eps (FP64: pressure-gradient sums cancel against it)
  double* flux_s;  // fluid Shepard sum of each flux wall (0 elsewhere)
  int* surf;       // free-surface label (P3.lbl_surf), written before the PPE when free_surface > 0
  R *visp, *condp; // per-particle viscosity and conductivity from the property tables (prop_table)
  R *dpor_prev;
  double *rho_ref, *rho0;  // reference / start-of-step density (P2; the Tait EOS amplifies their rounding)
  double *vol, *vol0;      // WCSPH particle volume and its initial value (P1.vol, P1.vol0)
  double *x0, *y0, *z0;    // WCSPH start-of-step position (P2.x0)
  double *xp, *yp, *rhop;  // 2D WCSPH with XSPH: predicted x, y, rho kept past the corrector
  R *cm;              // 6 * n, symmetric KGC inverse (xx, xy, xz, yy, yz, zz)
  R *grx, *gry, *grz; // delta-SPH density gradient
  R *fx, *fy, *fz, *fmag;   // ftotal (P3)
  R *fpx, *fpy, *fpz;       // pressure acceleration (P3.fp)
  R *pgx, *pgy, *pgz;       // pressure-gradient force density (P1.pgf)
  R *dtemp, *temp0;
  R *ux0, *uy0, *uz0;       // velocity at step start (P2.ux0)
  R *drho;                  // WCSPH continuity rate
  double *ppe1, *ppe2, *ppe3, *ppe4;
  // Packed neighbour records of the pair sums that take FP64 geometry
  // (the symmetric pressure gradient, and every sum of the FP64 path),
  // rebuilt before each use: {x, y, z, m}, {rho, p}, and for advection
  // {ux, uy, uz, T} and {h, p_type}
  double4* pk;
  double2* pk2;
  typename Vec4Of<R>::type* pkv;
  HRec<R>* pkh;
  Rec32* r32;
  RecAsm* ra;
  RecNeu* rn;
  PkB* pkb;
  PkC* pkc;
  R *xsx, *xsy, *xsz, *xst; // XSPH results
  // Positions as FP32 offsets inside a grid of power-of-two spacing pgrid
  // anchored at (pg0x, pg0y, pg0z), the grid cell packed in w (10 bits per
  // axis); differences come from rel_diff.  Rebuilt by NNPS and after every
  // position update (ISPH corrector, WCSPH predictor and corrector).
  float4* pos4;
  // SPH-SPH neighbour list (chunked slots: nbr[nbr_slot(k, i, n)]); nbr_cap is
  // a multiple of 8
  uint32_t* nbr;
  int* nbr_cnt;
  int nbr_cap;
  double pg0x, pg0y, pg0z;
  float pgrid;
  // periodic axes (bit 0 x, 1 y, 2 z) and their periods; on those axes the
  // pos4 grid has exactly 1024 cells per period
  int pmask;
  double plx, ply, plz;
};

// A position moved by one step back into [mn, mn + L) on a periodic axis.
__host__ __device__ inline double per_pos(double x, double mn, double L) {
  if (x >= mn + L) return x - L;
  if (x < mn) return x + L;
  return x;
}

// Periodic minimum image of a coordinate difference (the two positions lie
// within one period of each other): one subtraction, as the oracle does.
__host__ __device__ inline double per_wrap(double d, double L) {
  if (d > 0.5 * L) return d - L;
  if (d < -0.5 * L) return d + L;
  return d;
}

// x_a - x_b from two pos4 records: the integer cell difference times the
// power-of-two spacing (exact) plus the FP32 offset difference, one rounding
// each; for records a few cells apart the error stays below 2^-21 pgrid
__device__ inline float3 rel_diff(float4 a, float4 b, float g) {
  const uint32_t ca = __float_as_uint(a.w), cb = __float_as_uint(b.w);
  const float cx = (float)((int)(ca & 1023u) - (int)(cb & 1023u));
  const float cy = (float)((int)((ca >> 10) & 1023u) - (int)((cb >> 10) & 1023u));
  const float cz = (float)((int)(ca >> 20) - (int)(cb >> 20));
  return make_float3(fmaf(cx, g, a.x - b.x), fmaf(cy, g, a.y - b.y), fmaf(cz, g, a.z - b.z));
}
// the same on periodic axes (pmask bits): the grid has 1024 cells per
// period there, so the minimum image is a wrap of the integer cell
// difference into [-512, 512]
__device__ inline int cell_wrap(int d) { return d > 512 ? d - 1024 : (d < -512 ? d + 1024 : d); }
__device__ inline float3 rel_diff(float4 a, float4 b, float g, int pmask) {
  if (!pmask) return rel_diff(a, b, g);
  const uint32_t ca = __float_as_uint(a.w), cb = __float_as_uint(b.w);
  int dx = (int)(ca & 1023u) - (int)(cb & 1023u);
  int dy = (int)((ca >> 10) & 1023u) - (int)((cb >> 10) & 1023u);
  int dz = (int)(ca >> 20) - (int)(cb >> 20);
  if (pmask & 1) dx = cell_wrap(dx);
  if (pmask & 2) dy = cell_wrap(dy);
  if (pmask & 4) dz = cell_wrap(dz);
  return make_float3(fmaf((float)dx, g, a.x - b.x), fmaf((float)dy, g, a.y - b.y), fmaf((float)dz, g, a.z - b.z));
}
// x_i - x_j in FP64 with the periodic minimum image
template <class V>
__device__ inline void pdiff(const V& v, double xi, double yi, double zi, int j, double& dx, double& dy,
                             double& dz) {
  dx = xi - v.x[j];
  dy = yi - v.y[j];
  dz = zi - v.z[j];
  if (v.pmask) {
    if (v.pmask & 1) dx = per_wrap(dx, v.plx);
    if (v.pmask & 2) dy = per_wrap(dy, v.ply);
    if (v.pmask & 4) dz = per_wrap(dz, v.plz);
  }
}

struct DemView {
  int n;
  double *x, *y, *z, *ux, *uy, *uz, *wx, *wy, *wz, *m, *ri, *rad, *h, *rho, *temp;
  double *fdx, *fdy, *fdz;     // Fd_df (persistent: only written in the fluid regime)
  double *qsd;
  int *ptype, *btype, *itype, *dem_idx;
  int* hist_idx;      // kMaxHist * n
  double* hist_ds;    // kMaxHist * 3 * n
  int* hist_idx_new;
  double* hist_ds_new;
  // scratch
  double *flt_s, *flt_sd, *flt_sd2, *por;
  double *fax, *fay, *faz, *fbx, *fby, *fbz;
  double *fx, *fy, *fz, *tqx, *tqy, *tqz;
  double *ux0, *uy0, *uz0;
  double *test6;
  double *dtemp, *temp0;     // temperature rate (P3.dtemp) and start-of-step temperature (P2.temp0)
  double *hrad, *hcon;       // radiative and contact parts of dtemp (P1.test1, P1.test2)
  // WCSPH: start-of-step position and spin (P2.x0, P2.wx0), saved by the
  // half-step predictor before the spheres are re-sorted
  double *x0, *y0, *z0, *wx0, *wy0, *wz0;
};

struct CellView {
  const uint32_t* start;
  const uint32_t* end;
};

// ---------------------------------------------------------------------------
// cell geometry: functions_NNPS.cuh:306-390, 776-790
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t spread3(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}
__host__ __device__ inline uint64_t spread2(uint64_t v) {
  v = (v | (v << 16)) & 0x0000FFFF0000FFFFULL;
  v = (v | (v << 8)) & 0x00FF00FF00FF00FFULL;
  v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0FULL;
  v = (v | (v << 2)) & 0x3333333333333333ULL;
  v = (v | (v << 1)) & 0x5555555555555555ULL;
  return v;
}

__device__ inline int axis_cell(double x, double mn, double mx, int N) {
  if (mx == mn) return 0;
  double f = floor((x - mn) / (mx - mn) * N);
  double v = f < (double)(N - 1) ? f : (double)(N - 1);
  int c = (int)v;
  return c < 0 ? 0 : c;
}

__device__ inline int cell_key(const DevParams& P, int a, int b, int c) {
  if (!P.z_index) return a + P.NI * b + P.NI * P.NJ * c;
  if (P.dim == 2) return (int)(spread2((uint64_t)a) | (spread2((uint64_t)b) << 1));
  return (int)(spread3((uint64_t)a) | spread3((uint64_t)b) << 1 | spread3((uint64_t)c) << 2);
}

__device__ inline void cell_of(const DevParams& P, double x, double y, double z, int& a, int& b, int& c) {
  a = axis_cell(x, P.x_min, P.x_max, P.NI);
  b = axis_cell(y, P.y_min, P.y_max, P.NJ);
  c = axis_cell(z, P.z_min, P.z_max, P.NK);
}

// Exact FP64 distance: products rounded, then summed left to right, then a
// correctly rounded sqrt -- the oracle's expression with no FMA contraction.
__device__ inline double dist_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  double a = dx * dx;
  double b = dy * dy;
  double c = dz * dz;
  double s = (a + b) + c;
  return sqrt(s);
}

// Cell index c along an axis of N cells: wrapped on a periodic axis (bit of
// pmask), else false outside 0..N-1.
__device__ inline bool cell_in(int pmask, int axis, int& c, int N) {
  if (pmask >> axis & 1) {
    c = c < 0 ? c + N : (c > N - 1 ? c - N : c);
    return true;
  }
  return c >= 0 && c <= N - 1;
}

// Visit the 27 cells around a point in the reference order (z, y, x),
// wrapping periodic axes.
template <class F>
__device__ inline void for_cells(const DevParams& P, double x, double y, double z, const CellView& cv, F&& f) {
  int ic, jc, kc;
  cell_of(P, x, y, z, ic, jc, kc);
  for (int dz = -1; dz <= 1; dz++) {
    int c = kc + dz;
    if (!cell_in(P.pmask, 2, c, P.NK)) continue;
    for (int dy = -1; dy <= 1; dy++) {
      int b = jc + dy;
      if (!cell_in(P.pmask, 1, b, P.NJ)) continue;
      for (int dx = -1; dx <= 1; dx++) {
        int a = ic + dx;
        if (!cell_in(P.pmask, 0, a, P.NI)) continue;
        int k = cell_key(P, a, b, c);
        uint32_t s = cv.start[k];
        if (s == kEmpty) continue;
        uint32_t e = cv.end[k];
        for (uint32_t j = s; j < e; j++) f((int)j);
      }
    }
  }
}

// The same 27 cells scanned by a group of G aligned lanes: every lane of the
// group passes the same point, and the candidates of each row of three cells
// (fixed dz, dy) are dealt out over the G lanes, so each lane sees a subset
// and the caller adds the lanes' partial sums with group_sum<G>.  For work
// items too few to fill the chip one thread each (DEM spheres, walls).
template <int G, class F>
__device__ inline void group_for_cells(const DevParams& P, double x, double y, double z, const CellView& cv, F&& f) {
  int ic, jc, kc;
  cell_of(P, x, y, z, ic, jc, kc);
  const uint32_t lane = threadIdx.x & (G - 1);
  for (int dz = -1; dz <= 1; dz++) {
    int c = kc + dz;
    if (c < 0 || c > P.NK - 1) continue;
    for (int dy = -1; dy <= 1; dy++) {
      int b = jc + dy;
      if (b < 0 || b > P.NJ - 1) continue;
      uint32_t s[3], len[3];
#pragma unroll
      for (int dx = 0; dx < 3; dx++) {
        s[dx] = 0;
        len[dx] = 0;
        int a = ic + dx - 1;
        if (a < 0 || a > P.NI - 1) continue;
        int k = cell_key(P, a, b, c);
        uint32_t st = cv.start[k];
        if (st == kEmpty) continue;
        s[dx] = st;
        len[dx] = cv.end[k] - st;
      }
      const uint32_t l01 = len[0] + len[1], tot = l01 + len[2];
      for (uint32_t t = lane; t < tot; t += G)
        f((int)(t < len[0] ? s[0] + t : t < l01 ? s[1] + (t - len[0]) : s[2] + (t - l01)));
    }
  }
}
template <class F>
__device__ inline void wave_for_cells(const DevParams& P, double x, double y, double z, const CellView& cv, F&& f) {
  group_for_cells<64>(P, x, y, z, cv, static_cast<F&&>(f));
}

// sum over a group of G aligned lanes; the butterfly gives every lane of the
// group the same bits
template <int G, typename T>
__device__ inline T group_sum(T a) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) a += __shfl_xor(a, off);
  return a;
}
__device__ inline double wave_sum(double a) { return group_sum<64>(a); }

// the (2 rng + 1)^3 cells around (x,y,z), same visiting order as for_cells
template <class F>
__device__ inline void for_cells_r(const DevParams& P, double x, double y, double z, const CellView& cv, int rng,
                                   F&& f) {
  int ic, jc, kc;
  cell_of(P, x, y, z, ic, jc, kc);
  for (int dz = -rng; dz <= rng; dz++) {
    int c = kc + dz;
    if (c < 0 || c > P.NK - 1) continue;
    for (int dy = -rng; dy <= rng; dy++) {
      int b = jc + dy;
      if (b < 0 || b > P.NJ - 1) continue;
      for (int dx = -rng; dx <= rng; dx++) {
        int a = ic + dx;
        if (a < 0 || a > P.NI - 1) continue;
        int k = cell_key(P, a, b, c);
        uint32_t s = cv.start[k];
        if (s == kEmpty) continue;
        uint32_t e = cv.end[k];
        for (uint32_t j = s; j < e; j++) f((int)j);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// kernel functions: functions_KNL.cuh:2-173 (Wendland uses q = r / 2h)
// ---------------------------------------------------------------------------
template <typename T>
__host__ __device__ inline T knl_tmpA(int kt, int dim, T tH) {
  const T PI = (T)3.141592653;
  T tA = 0;
  switch (kt) {
    case SOPHIA_KNL_GAUSSIAN:
      tA = dim == 2 ? (T)1.0 / (PI * tH * tH) : (T)1.0 / (pow(PI, (T)1.5) * tH * tH * tH);
      break;
    case SOPHIA_KNL_QUINTIC:
      tA = dim == 2 ? (T)7.0 / ((T)478.0 * PI * tH * tH) : (T)3.0 / ((T)359.0 * PI * tH * tH * tH);
      break;
    case SOPHIA_KNL_QUARTIC:
      tA = dim == 2 ? (T)15.0 / ((T)7.0 * PI * tH * tH) : (T)315.0 / ((T)208.0 * PI * tH * tH * tH);
      break;
    case SOPHIA_KNL_WENDLAND2:
      tA = dim == 2 ? (T)2.228169203286535 / (4 * tH * tH) : (T)3.342253804929802 / (8 * tH * tH * tH);
      break;
    case SOPHIA_KNL_WENDLAND4:
      tA = dim == 2 ? (T)2.864788975654116 / (4 * tH * tH) : (T)4.923856051905513 / (8 * tH * tH * tH);
      break;
    case SOPHIA_KNL_WENDLAND6:
      tA = dim == 2 ? (T)3.546881588905096 / (4 * tH * tH) : (T)6.788953041263660 / (8 * tH * tH * tH);
      break;
  }
  return tA;
}

template <typename T>
__host__ __device__ inline T ipow5(T v) { return v * v * v * v * v; }
template <typename T>
__host__ __device__ inline T ipow4(T v) { return v * v * v * v; }

template <typename T>
__host__ __device__ inline T knl_w(int kt, T tA, T tH, T rr) {
  T w = 0;
  switch (kt) {
    case SOPHIA_KNL_GAUSSIAN: {
      T tR = rr / tH;
      w = tA * exp(-tR * tR);
      break;
    }
    case SOPHIA_KNL_QUINTIC: {
      T tR = rr / tH;
      if (tR < 1) w = tA * (ipow5((T)3 - tR) - (T)6 * ipow5((T)2 - tR) + (T)15 * ipow5((T)1 - tR));
      else if (tR < 2) w = tA * (ipow5((T)3 - tR) - (T)6 * ipow5((T)2 - tR));
      else if (tR < 3) w = tA * ipow5((T)3 - tR);
      break;
    }
    case SOPHIA_KNL_QUARTIC: {
      T tR = rr / tH;
      if (tR < 2) w = tA * ((T)(2.0 / 3.0) - (T)(9.0 / 8.0) * tR * tR + (T)(19.0 / 24.0) * tR * tR * tR - (T)(5.0 / 32.0) * tR * tR * tR * tR);
      break;
    }
    case SOPHIA_KNL_WENDLAND2: {
      T tR = rr / tH * (T)0.5;
      T o = 1 - tR;
      if (tR < 1) w = tA * o * o * o * o * (1 + 4 * tR);
      break;
    }
    case SOPHIA_KNL_WENDLAND4: {
      T tR = rr / tH * (T)0.5;
      T o = 1 - tR;
      if (tR < 1) w = tA * o * o * o * o * o * o * (1 + 6 * tR + (T)11.666666666666666 * tR * tR);
      break;
    }
    case SOPHIA_KNL_WENDLAND6: {
      T tR = rr / tH * (T)0.5;
      T o = 1 - tR;
      if (tR < 1) w = tA * o * o * o * o * o * o * o * o * (1 + 8 * tR + 25 * tR * tR + 32 * tR * tR * tR);
      break;
    }
  }
  return w;
}

template <typename T>
__host__ __device__ inline T knl_dw(int kt, T tA, T tH, T rr) {
  T dw = 0;
  switch (kt) {
    case SOPHIA_KNL_GAUSSIAN: {
      T tR = rr / tH;
      dw = ((T)1 / tH) * tA * (T)(-2) * tR * exp(-tR * tR);
      break;
    }
    case SOPHIA_KNL_QUINTIC: {
      T tR = rr / tH;
      if (tR < 1) dw = ((T)1 / tH) * tA * (ipow5((T)3 - tR) + (T)30 * ipow4((T)2 - tR) - (T)75 * ipow4((T)1 - tR));
      else if (tR < 2) dw = ((T)1 / tH) * tA * ((T)(-5) * ipow4((T)3 - tR) + (T)30 * ipow4((T)2 - tR));
      else if (tR < 3) dw = ((T)1 / tH) * tA * ((T)(-5) * ipow4((T)3 - tR));
      break;
    }
    case SOPHIA_KNL_QUARTIC: {
      T tR = rr / tH;
      if (tR < 2) dw = ((T)1 / tH) * tA * ((T)(-9.0 / 8.0 * 2) * tR + (T)(19.0 / 24.0 * 3.0) * tR * tR - (T)(5.0 / 32.0 * 4.0) * tR * tR * tR);
      break;
    }
    case SOPHIA_KNL_WENDLAND2: {
      T tR = rr / tH * (T)0.5;
      T o = 1 - tR;
      if (tR < 1) dw = ((T)1 / (2 * tH)) * tA * ((T)(-20) * tR * o * o * o);
      break;
    }
    case SOPHIA_KNL_WENDLAND4: {
      T tR = rr / tH * (T)0.5;
      T o = 1 - tR;
      if (tR < 1) dw = ((T)1 / (2 * tH)) * tA * ((T)(-18.666666666666668) * tR * o * o * o * o * o * (5 * tR + 1));
      break;
    }
    case SOPHIA_KNL_WENDLAND6: {
      T tR = rr / tH * (T)0.5;
      T o = 1 - tR;
      if (tR < 1) dw = ((T)1 / (2 * tH)) * tA * ((T)(-22) * tR * o * o * o * o * o * o * o * (16 * tR * tR + 7 * tR + 1));
      break;
    }
  }
  return dw;
}

// knl_w / knl_dw with 1 / h supplied by the caller (hoisted out of the pair
// loop): the Wendland C2 branch multiplies where knl_* divides by h, which
// moves the result by at most an ulp; other kernels take the general path
template <typename T>
__host__ __device__ inline T knl_w_ih(int kt, T tA, T tH, T ih, T rr) {
  if (kt == SOPHIA_KNL_WENDLAND2) {
    const T tR = rr * ih * (T)0.5;
    const T o = 1 - tR;
    return tR < 1 ? tA * o * o * o * o * (1 + 4 * tR) : (T)0;
  }
  return knl_w<T>(kt, tA, tH, rr);
}
template <typename T>
__host__ __device__ inline T knl_dw_ih(int kt, T tA, T tH, T ih, T rr) {
  if (kt == SOPHIA_KNL_WENDLAND2) {
    const T tR = rr * ih * (T)0.5;
    const T o = 1 - tR;
    return tR < 1 ? (ih * (T)0.5) * tA * ((T)(-20) * tR * o * o * o) : (T)0;
  }
  return knl_dw<T>(kt, tA, tH, rr);
}

// KGC-corrected gradient with the symmetric inverse stored as 6 entries
template <typename R, typename T>
__device__ inline void kgc_apply(const DevParams& P, const R* cm, int n, int i, T& gx, T& gy, T& gz) {
  if (P.kgc != 1) return;
  T a = (T)cm[i], b = (T)cm[n + i], c = (T)cm[2 * n + i];
  T d = (T)cm[3 * n + i], e = (T)cm[4 * n + i], f = (T)cm[5 * n + i];
  T x = a * gx + b * gy + c * gz;
  T y = b * gx + d * gy + e * gz;
  T z = c * gx + e * gy + f * gz;
  gx = x;
  gy = y;
  gz = z;
}

inline int grid_for(int n, int block = kBlock) { return (n + block - 1) / block; }
// blocks for a per-particle kernel that gathers neighbour data: grid_for
// rounded up to a multiple of 8, for particle_index()
inline int grid_xcd(int n, int block = kBlock) {
  const int g = grid_for(n, block);
  return g < 8 ? g : (g + 7) & ~7;
}

// Particle of this thread.  Workgroups are dealt to the eight XCDs round
// robin; with a grid that is a multiple of 8, block b is given logical block
// (b mod 8) * (G / 8) + b / 8, so each XCD works on one contiguous eighth of
// the cell-sorted particles and the neighbour fields its waves gather stay
// in its own L2.  Other grids map linearly.
__device__ inline int particle_index() {
  const int G = gridDim.x;
  const int b = (G & 7) ? (int)blockIdx.x : (int)((blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3));
  return b * (int)blockDim.x + (int)threadIdx.x;
}

// ---------------------------------------------------------------------------
// launch wrappers implemented in the .hip files
// ---------------------------------------------------------------------------
struct SortScratch {
  uint32_t *keys_a, *keys_b, *vals_a, *vals_b;
  uint32_t* hist;   // 256 * max_tiles, then 256 digit totals
  int max_tiles;
};

int nnps_keys(hipStream_t s, const DevParams& P, int n, const double* x, const double* y, const double* z,
              const int* itype, uint32_t* keys, uint32_t* vals);
int radix_sort_pairs(hipStream_t s, int n, int nbits, SortScratch& sc, uint32_t** keys_out, uint32_t** vals_out);
int cell_ranges(hipStream_t s, int n, int num_cells, const uint32_t* keys, uint32_t* start, uint32_t* end);
struct PermuteField {
  void* src;
  void* dst;
  int bytes;  // 4 or 8
};
int permute_fields(hipStream_t s, int n, const uint32_t* order, const PermuteField* fields, int nf);
int fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n);
int fill_zero_bytes(hipStream_t s, void* p, size_t bytes);
// several fills in one launch (one grid row per span)
struct FillSpan {
  uint32_t* p;
  size_t n;  // 32-bit words
  uint32_t v;
};
int fill_u32_many(hipStream_t s, const FillSpan* spans, int count);

template <typename R>
int build_sph_neighbors(hipStream_t s, const DevParams& P, const SphView<R>& v, const CellView& cv, int* d_maxcnt);
// pos4 records of the current positions
template <typename R>
int refresh_pos4(hipStream_t s, const SphView<R>& v);

// SPH stages
template <typename R> int sph_mass_init(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int sph_kgc(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int sph_porosity(hipStream_t s, const DevParams& P, const SphView<R>& v, const DemView& d, const CellView& dcells);
template <typename R> int sph_prep(hipStream_t s, const DevParams& P, const SphView<R>& v);
// KERNEL_clc_prep3D_sph alone, without the ISPH mass/density reset (the WCSPH
// driver integrates density by continuity)
template <typename R> int sph_prep_sums(hipStream_t s, const DevParams& P, const SphView<R>& v);
// sph_kgc (when kgc) and sph_porosity in one pass over the list on the mixed path
template <typename R>
int sph_kgc_porosity(hipStream_t s, const DevParams& P, const SphView<R>& v, const DemView& d, const CellView& dc,
                     bool kgc);
// k_prep_prep alone (the ISPH density reset that sph_prep runs before its sums)
template <typename R> int sph_prep_prep(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int sph_advection(hipStream_t s, const DevParams& P, const SphView<R>& v);
// fluid Shepard sums of the flux walls (no-op without a wall heat flux)
template <typename R> int sph_flux_walls(hipStream_t s, const DevParams& P, const SphView<R>& v);
// per-particle viscosity and conductivity from the property tables (prop_table)
template <typename R> int sph_props(hipStream_t s, const DevParams& P, const SphView<R>& v);
// free-surface labels and their Dirichlet pressure before the PPE (free_surface > 0)
template <typename R> int sph_surface(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int sph_neumann(hipStream_t s, const DevParams& P, const SphView<R>& v, const CellView& cv);
template <typename R> int sph_pressure_force(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int sph_predictor(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int sph_corrector(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int sph_xsph(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int sph_xsph_only(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int sph_set_alpha_noop(hipStream_t s, const DevParams& P, const SphView<R>& v);

// DEM stages
template <typename R> int dem_porosity(hipStream_t s, const DevParams& P, const DemView& d, const SphView<R>& v, const CellView& scells, const CellView& dcells);
int dem_contact(hipStream_t s, const DevParams& P, const DemView& d, const CellView& dcells);
template <typename R> int dem_coupling(hipStream_t s, const DevParams& P, const DemView& d, const SphView<R>& v, const CellView& scells);
template <typename R> int sph_dem_reaction(hipStream_t s, const DevParams& P, const SphView<R>& v, const DemView& d, const CellView& dcells);
int dem_predictor(hipStream_t s, const DevParams& P, const DemView& d);
int dem_corrector(hipStream_t s, const DevParams& P, const DemView& d);
// WCSPH half-step predictor and corrector of the spheres
int wc_dem_predictor(hipStream_t s, const DevParams& P, const DemView& d);
int wc_dem_corrector(hipStream_t s, const DevParams& P, const DemView& d);
// open boundaries (sophia_params.open_boundary = 1, ob_*)
struct OpenBoundary {
  double L[4];  // z bands: inlet [L1, L2), outlet [L3, L4)
  double inlet_u, inlet_temp;
};
template <typename R>
int sph_open_boundary(hipStream_t s, const DevParams& P, const SphView<R>& v, const CellView& cv,
                      const OpenBoundary& ob);
// DEM wall boundaries (sophia_params.dem_walls and wall_*)
struct DemWallParams {
  int mask;                  // bit 0 cylinder, 1 box, 2 truncated cone
  double cyl[3], box[6], cone[5];
  double rest_n, rest_s, mu_s, mu_rs;
};
int dem_walls(hipStream_t s, const DevParams& P, const DemView& d, const DemWallParams& W);

// WCSPH stages (wcsph.hip)
template <typename R> int wc_predictor(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc_reset_p3(hipStream_t s, const SphView<R>& v);
template <typename R> int wc_volume_mass(hipStream_t s, const DevParams& P, const SphView<R>& v, const CellView& cv, int count);
template <typename R> int wc_continuity(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc_eos(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc_interaction(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc_corrector(hipStream_t s, const DevParams& P, const SphView<R>& v);
// 2D WCSPH (SOPHIA_multi's 2D kernels)
template <typename R> int wc2_kgc(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc2_prep(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc2_boundary(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc2_continuity(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc2_interaction(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc2_time_update(hipStream_t s, const DevParams& P, const SphView<R>& v);
template <typename R> int wc2_xsph(hipStream_t s, const DevParams& P, const SphView<R>& v);
// max |u|, rho, ftotal of one species (scratch: 3 * 1024 + 3 doubles)
template <typename U>
int monitor_max(hipStream_t s, int n, const int* ptype, const int* itype, const U* ux, const U* uy, const U* uz,
                const double* rho, const U* ft, double* scratch, double* out);
int knl_eval_device(hipStream_t s, int kt, int dim, double h, const double* r, int n, double* w, double* dw);
// allocation-free variant on a file-scope device array (nnps.hip): at most
// kKnlScratchMax radii per call; calls that overlap on one device race
constexpr int kKnlScratchMax = 65536;
constexpr int kMaxDevices = 64;
int knl_scratch_init(int device);
int knl_eval_noalloc(int kt, int dim, double h, const double* r, int n, double* w, double* dw);

}  // namespace sophia
