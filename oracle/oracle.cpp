/*
 * FP64 CPU oracle of the SOPHIA ISPH-DEM step (TEST INFRASTRUCTURE ONLY).
 *
 * A plain C++/OpenMP restatement of the reference kernels, run on the
 * reference's own AoS records with the reference's buffer roles
 * (P1 = carried state, SP1/SP2 = sorted work copies, P3 = per-step
 * accumulators zeroed at NNPS, DEM history read from SP1 and written to P1).
 * Each function cites the reference file:line it follows.  The oracle is the
 * checker of the HIP library and the CPU baseline of the benchmark; nothing
 * in the product links it.
 *
 * Built with -O3 -fopenmp -ffp-contract=off: no FMA contraction, so every
 * expression is the rounded IEEE sequence written here.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   cell geometry / keys      closed form of find_minmax + floor((x-min)/span*NI),
 *                             brute recomputation, Morton bit pattern vs a
 *                             bit-by-bit interleave
 *   sort / cell ranges        std::stable_sort contract (CUB SortPairs is
 *                             stable); ranges recomputed from sorted keys
 *   neighbour counts          O(N^2) brute force with the same FP64 test
 *   kernel functions          closed forms, normalisation (quadrature of
 *                             W over R^3 = 1), dW = numerical derivative of W
 *   KGC                       Cm = I on a uniform lattice interior; corrected
 *                             gradient of a linear field is exact
 *   porosity                  one sphere in a uniform lattice: 1 - eps equals
 *                             V_sphere * W / Shepard sum (closed form)
 *   PPE operator              manufactured p = |x|^2 on a uniform lattice with
 *                             eps = 1, KGC off: (A p)_i ~ -6 * 2 (interior)
 *   BiCGStab                  dense numpy solve of the explicitly assembled
 *                             tiny system
 *   Hertz-Mindlin             two-sphere head-on: F = 1.33333 E* sqrt(R*) d^1.5,
 *                             pair antisymmetry, tangential cap
 *   Di Felice drag            single sphere at eps -> 1: (1/8) Cd rho pi d^2
 *                             |u| u with Dallavalle's Cd, Re 1 to 1e4; air and
 *                             mixture regimes: Rong et al. (2013) porosity
 *                             function at the reported porosity
 *   DEM pressure gradient     uniform grad p gives -grad p V_a / m_a
 *   coupling reaction         momentum balance sum_i m_i da_i = -m_a F_a and
 *                             energy balance sum_i m_i c_p dT_i = -sum_a Q_sd
 *   Shepard / delta-SPH       Shepard sum = closed-form lattice quadrature of
 *                             W (within 1 % of 1); delta-SPH gradient exact
 *                             for a linear density (KGC)
 *   advection terms           viscous: 0 for linear shear, mu/rho lap(u) for
 *                             a quadratic (exact with KGC on a lattice);
 *                             Eulerian -(u.grad)u exact for linear u;
 *                             conduction 0 for linear T, Morris lattice sum
 *                             for quadratic T; artificial viscosity 0 in
 *                             expansion, dissipative in compression;
 *                             symmetric pressure gradient exact in a lattice
 *                             interior
 *   PPE right-hand side       -rho div(u*)/dt exact for linear u* (KGC); the
 *                             porosity rate = central difference of eps
 *   Neumann walls             brute force over all pairs (2 kappa h, +-1
 *                             cell, W at 2h)
 *   DEM porosity              brute force; one sphere: 1 - V W(0)/flt_sd
 *   heat                      Ranz-Marshall Q = Nu k pi d dT on one sphere;
 *                             contact conduction 2 k a dT (Batchelor-O'Brien)
 *                             equal and opposite; radiation equal and
 *                             opposite, ~T^4; wall flux delivers q A
 *   free surface              lattice labels: the open top only (walls
 *                             counted), plus wall-adjacent fluid (same-type
 *                             rule); a column stays at rest for 1000
 *                             converged steps and drifts without it
 *   predictor/corrector/XSPH  exact restatement, checked against hand-written
 *                             per-particle formulas in the tests
 *   open boundary             uniform interior fluid: the outlet copies it,
 *                             the inlet mirrors its velocity about u_in;
 *                             ghost node mirrored about the band plane
 *   DEM walls                 head-on impact: u_n -> -rest_n u_n, position
 *                             on the wall, spin unchanged; resting contact:
 *                             normal velocity removed, friction opposes slip;
 *                             all three geometries
 *   2D WCSPH                  KGC_2D: sum_j V_j x_ji (x) C grad W_ij = I for
 *   (tests/test_wcsph2d_      every particle; prep2D: rho_ref = m/(h/h_over_dx)^2,
 *    oracle.py)               Shepard sum = lattice quadrature, gradient of a
 *                             linear density exact (stored with the
 *                             reference's minus sign); continuity: -rho^2/rho_ref
 *                             div u for linear u, Molteni = delta h c0 lap(rho)
 *                             for a quadratic density, the reference's
 *                             Antuono term = twice Molteni there; boundary2D:
 *                             walls = minus a uniform fluid velocity;
 *                             interaction2D: -G/rho in a linear pressure,
 *                             Morris mu/rho lap(u), k lap(T)/(rho c_p m_i),
 *                             gravity (0, -g), artificial viscosity off in
 *                             expansion and dissipative; time update + XSPH:
 *                             symplectic-Euler free fall, exact translation
 *                             of a uniform flow
 *   parity unpinned           the 3D WCSPH interaction's heat and virial
 *                             terms beyond the gravity-only and uniform-flow
 *                             checks
 */
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr uint32_t kEmpty = 0xffffffffu;  // cu_memset (Parameters.cuh:1)
constexpr double kPI = 3.141592653;       // PI (Parameters.cuh:63)
constexpr double kMinDet = 1e-4;          // Min_det (Parameters.cuh:125)
constexpr int kITypeCrt = 2;              // i_type_crt (Parameters.cuh:128)

struct Species {
  std::vector<part1> P1, SP1;
  std::vector<part2> P2, SP2;
  std::vector<part3> P3;
  std::vector<int32_t> key_in, key, order;       // g_idx_in, g_idx, p_idx
  std::vector<uint32_t> g_str, g_end;            // cell ranges
  int n() const { return (int)P1.size(); }
};

struct PPERow {
  int i;
  std::vector<int> j;
  std::vector<double> a;
};

}  // namespace

struct orc_state {
  sophia_params p;
  sophia_domain d;
  // periodic axes (sophia_params.periodic_mask) and their periods
  int pmask = 0;
  double plen[3] = {0, 0, 0};
  Species sph, dem;
  int64_t count = 0;
  double time = 0.0;
  bool after_corrector = false;
  // fluid Shepard sum of each flux wall (sorted SPH order; 0 elsewhere)
  std::vector<double> flux_S;
  // data.txt property tables (orc_set_table)
  sophia_table table{};
  bool has_table = false;
  // PPE system of the last assembly
  std::vector<int32_t> row_mask;
  std::vector<PPERow> rows;
  std::vector<double> rhs;
};

static int g_threads = 0;

/* ------------------------------------------------------------------------ */
/* kernel functions: calc_tmpA, calc_kernel_wij, calc_kernel_dwij           */
/* (functions_KNL.cuh:2-32, 47-109, 111-173)                                */
/* ------------------------------------------------------------------------ */
static double knl_tmpA(int kt, int dim, double tH) {
  double tA = 0.0;
  if (kt == SOPHIA_KNL_GAUSSIAN) {
    if (dim == 2) tA = 1.0 / (kPI * std::pow(tH, 2));
    else if (dim == 3) tA = 1.0 / (std::pow(kPI, 1.5) * std::pow(tH, 3));
  } else if (kt == SOPHIA_KNL_QUINTIC) {
    if (dim == 2) tA = 7.0 / (478.0 * kPI * std::pow(tH, 2));
    else if (dim == 3) tA = 3.0 / (359.0 * kPI * std::pow(tH, 3));
  } else if (kt == SOPHIA_KNL_QUARTIC) {
    if (dim == 2) tA = 15.0 / (7.0 * kPI * std::pow(tH, 2));
    else if (dim == 3) tA = 315.0 / (208.0 * kPI * std::pow(tH, 3));
  } else if (kt == SOPHIA_KNL_WENDLAND2) {
    if (dim == 2) tA = 2.228169203286535 / (4 * tH * tH);
    else if (dim == 3) tA = 3.342253804929802 / (8 * tH * tH * tH);
  } else if (kt == SOPHIA_KNL_WENDLAND4) {
    if (dim == 2) tA = 2.864788975654116 / (4 * tH * tH);
    else if (dim == 3) tA = 4.923856051905513 / (8 * tH * tH * tH);
  } else if (kt == SOPHIA_KNL_WENDLAND6) {
    if (dim == 2) tA = 3.546881588905096 / (4 * tH * tH);
    else if (dim == 3) tA = 6.788953041263660 / (8 * tH * tH * tH);
  }
  return tA;
}

static inline double p5(double v) { return v * v * v * v * v; }
static inline double p4(double v) { return v * v * v * v; }

static double knl_w(int kt, double tA, double tH, double rr) {
  double tR, w = 0.0;
  switch (kt) {
    case SOPHIA_KNL_GAUSSIAN:
      tR = rr / tH;
      w = tA * std::exp(-std::pow(tR, 2));
      break;
    case SOPHIA_KNL_QUINTIC:
      tR = rr / tH;
      if (tR < 1) w = tA * (std::pow(3.0 - tR, 5) - 6.0 * std::pow(2.0 - tR, 5) + 15.0 * std::pow(1.0 - tR, 5));
      else if (1 <= tR && tR < 2) w = tA * (std::pow(3.0 - tR, 5) - 6.0 * std::pow(2.0 - tR, 5));
      else if (2 <= tR && tR < 3) w = tA * (std::pow(3.0 - tR, 5));
      break;
    case SOPHIA_KNL_QUARTIC:
      tR = rr / tH;
      w = (tR < 2) * tA * (2.0 / 3.0 - 9.0 / 8.0 * std::pow(tR, 2) + 19.0 / 24.0 * std::pow(tR, 3) - 5.0 / 32.0 * std::pow(tR, 4));
      break;
    case SOPHIA_KNL_WENDLAND2: {
      tR = rr / tH * 0.5;
      double o = 1 - tR;
      w = (tR < 1) * tA * o * o * o * o * (1 + 4 * tR);
      break;
    }
    case SOPHIA_KNL_WENDLAND4: {
      tR = rr / tH * 0.5;
      double o = 1 - tR;
      w = (tR < 1) * tA * o * o * o * o * o * o * (1 + 6 * tR + 11.666666666666666 * tR * tR);
      break;
    }
    case SOPHIA_KNL_WENDLAND6: {
      tR = rr / tH * 0.5;
      double o = 1 - tR;
      w = (tR < 1) * tA * o * o * o * o * o * o * o * o * (1 + 8 * tR + 25 * tR * tR + 32 * tR * tR * tR);
      break;
    }
  }
  return w;
}

static double knl_dw(int kt, double tA, double tH, double rr) {
  double tR, dw = 0.0;
  switch (kt) {
    case SOPHIA_KNL_GAUSSIAN:
      tR = rr / tH;
      dw = (1.0 / tH) * tA * (-2.0) * tR * std::exp(-std::pow(tR, 2));
      break;
    case SOPHIA_KNL_QUINTIC:
      tR = rr / tH;
      if (tR < 1) dw = (1.0 / tH) * tA * (std::pow(3.0 - tR, 5) + 30.0 * std::pow(2.0 - tR, 4) - 75.0 * std::pow(1.0 - tR, 4));
      else if (1 <= tR && tR < 2) dw = (1.0 / tH) * tA * (-5.0 * std::pow(3.0 - tR, 4) + 30.0 * std::pow(2.0 - tR, 4));
      else if (2 <= tR && tR < 3) dw = (1.0 / tH) * tA * (-5.0 * std::pow(3.0 - tR, 4));
      break;
    case SOPHIA_KNL_QUARTIC:
      tR = rr / tH;
      dw = (tR < 2) * (1.0 / tH) * tA * (-9.0 / 8.0 * 2 * tR + 19.0 / 24.0 * 3.0 * std::pow(tR, 2) - 5.0 / 32.0 * 4.0 * std::pow(tR, 3));
      break;
    case SOPHIA_KNL_WENDLAND2: {
      tR = rr / tH * 0.5;
      double o = 1 - tR;
      dw = (tR < 1) * (1 / (2 * tH)) * tA * (-20 * tR * o * o * o);
      break;
    }
    case SOPHIA_KNL_WENDLAND4: {
      tR = rr / tH * 0.5;
      double o = 1 - tR;
      dw = (tR < 1) * (1 / (2 * tH)) * tA * (-18.666666666666668 * tR * o * o * o * o * o * (5 * tR + 1));
      break;
    }
    case SOPHIA_KNL_WENDLAND6: {
      tR = rr / tH * 0.5;
      double o = 1 - tR;
      dw = (tR < 1) * (1 / (2 * tH)) * tA * (-22 * tR * o * o * o * o * o * o * o * (16 * tR * tR + 7 * tR + 1));
      break;
    }
  }
  return dw;
}

/* ------------------------------------------------------------------------ */
/* material properties with the reference constants as parameters           */
/* (functions_PROP.cuh:228-514)                                             */
/* ------------------------------------------------------------------------ */
/* interp2 (functions_PROP.cuh:143-175): piecewise linear over ascending
 * xs, the end segments extended beyond both ends */
static double interp2(const double* xs, const double* ys, int n, double x) {
  int a;
  if (xs[n - 1] < x) {
    a = n - 2;
  } else if (x <= xs[0]) {
    a = 0;
  } else {
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
static const sophia_table* g_no_table = nullptr;
/* viscosity(temp, p_type) and conductivity(temp, p_type) (functions_PROP.cuh:
 * 228-264): the data.txt block of |p_type| when property_table is on and the
 * type has one, else the constants (the reference's VISCOSITY_AIR for every
 * type, 0.0518 for fluid, 2.5 for spheres, as parameters) */
static const double* table_block(const sophia_params& p, const sophia_table* t, int ptype, int& n) {
  const int a = std::abs(ptype);
  if (!p.prop_table || !t || a >= t->n_tables || t->size[a] < 2) return nullptr;
  n = t->size[a];
  return t->T + t->index[a];
}
static double viscosity(const sophia_params& p, const sophia_table* t, int ptype, double T) {
  int n = 0;
  if (const double* xs = table_block(p, t, ptype, n)) return interp2(xs, t->mu + (xs - t->T), n, T);
  return p.mu_fluid;
}
static double conductivity(const sophia_params& p, const sophia_table* t, int ptype, double T) {
  int n = 0;
  if (const double* xs = table_block(p, t, ptype, n)) return interp2(xs, t->k + (xs - t->T), n, T);
  return std::abs(ptype) > 1000 ? p.k_dem : p.k_fluid;
}
static double heat_capacity(const sophia_params& p, int ptype) {
  return std::abs(ptype) > 1000 ? p.cp_dem : p.cp_fluid;
}
// emissivity of spheres (functions_PROP.cuh:420-440, graphite fit above 700 K)
static double dem_emissivity(double T) {
  if (T >= 700.0) return 0.505586 + 7.92943e-4 * T - 5.228643e-7 * T * T + 1.10479e-10 * T * T * T;
  return 0.842337;
}
// Pr^(1/3) of the Nusselt correlation; the reference's integer 1/3 gives Pr^0
static double pr_third(const sophia_params& p, double pr) {
  return (p.reference_compat & SOPHIA_COMPAT_NUSSELT) ? 1.0 : std::cbrt(pr);
}
static bool in_flux_box(const sophia_params& p, const part1& q) {
  const double* b = p.heat_flux_box;
  return q.p_type == 0 && q.x >= b[0] && q.y >= b[1] && q.z >= b[2] && q.x <= b[3] && q.y <= b[4] && q.z <= b[5];
}

/* ------------------------------------------------------------------------ */
/* cell geometry (function_init.cuh:913-950; ISPH.cuh:93-113;               */
/* functions_NNPS.cuh:306-390, 776-790)                                     */
/* ------------------------------------------------------------------------ */
static uint64_t morton3(unsigned a, unsigned b, unsigned c) {
  auto spread = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  return spread(a) | spread(b) << 1 | spread(c) << 2;
}

static uint64_t morton2(uint64_t x, uint64_t y) {
  auto spread = [](uint64_t v) {
    v = (v | (v << 16)) & 0x0000FFFF0000FFFFULL;
    v = (v | (v << 8)) & 0x00FF00FF00FF00FFULL;
    v = (v | (v << 4)) & 0x0F0F0F0F0F0F0F0FULL;
    v = (v | (v << 2)) & 0x3333333333333333ULL;
    v = (v | (v << 1)) & 0x5555555555555555ULL;
    return v;
  };
  return spread(x) | (spread(y) << 1);
}

static int axis_cell(double x, double mn, double mx, int N) {
  if (mx == mn) return 0;
  double f = std::floor((x - mn) / (mx - mn) * N);
  double v = f < (N - 1) ? f : (double)(N - 1);
  int c = (int)v;
  if (c < 0) c = 0;
  return c;
}

static int idx_cell(const orc_state& s, int ix, int iy, int iz) {
  if (s.p.flag_z_index == 0) return ix + s.d.NI * iy + s.d.NI * s.d.NJ * iz;
  if (s.p.dim == 2) return (int)morton2((uint64_t)ix, (uint64_t)iy);
  return (int)morton3((unsigned)ix, (unsigned)iy, (unsigned)iz);
}

/* periodic minimum image of the coordinate difference a - b along `axis`
 * (no build of the reference is periodic: its KERNEL_periodic_3D,
 * functions_BC.cuh:1522, is a velocity copy that is never called) */
static inline double pd(const orc_state& s, double a, double b, int axis) {
  double d = a - b;
  if (s.pmask >> axis & 1) {
    const double L = s.plen[axis];
    if (d > 0.5 * L) d -= L;
    else if (d < -0.5 * L) d += L;
  }
  return d;
}
/* cell index c along `axis` of N cells: wrapped on a periodic axis, else
 * false outside 0..N-1 */
static inline bool cell_in(const orc_state& s, int axis, int& c, int N) {
  if (s.pmask >> axis & 1) {
    if (c < 0) c += N;
    else if (c > N - 1) c -= N;
    return true;
  }
  return c >= 0 && c <= N - 1;
}
/* a position moved by one step back into [min, min + L) on a periodic axis */
static inline double wrap_pos(const orc_state& s, double x, int axis) {
  if (!(s.pmask >> axis & 1)) return x;
  const double mn = axis == 0 ? s.d.x_min : axis == 1 ? s.d.y_min : s.d.z_min;
  if (x >= mn + s.plen[axis]) return x - s.plen[axis];
  if (x < mn) return x + s.plen[axis];
  return x;
}

static void cell_of(const orc_state& s, double x, double y, double z, int& ic, int& jc, int& kc) {
  ic = axis_cell(x, s.d.x_min, s.d.x_max, s.d.NI);
  jc = axis_cell(y, s.d.y_min, s.d.y_max, s.d.NJ);
  kc = axis_cell(z, s.d.z_min, s.d.z_max, s.d.NK);
}

/* Visit candidates of the 27 cells around (x,y,z) in reference order:
 * z, then y, then x in {-1,0,1}, each cell's particles in sorted order. */
template <class F>
static void for_candidates(const orc_state& s, const Species& sp, double x, double y, double z, F&& f) {
  int ic, jc, kc;
  cell_of(s, x, y, z, ic, jc, kc);
  for (int dz = -1; dz <= 1; dz++)
    for (int dy = -1; dy <= 1; dy++)
      for (int dx = -1; dx <= 1; dx++) {
        int a = ic + dx, b = jc + dy, c = kc + dz;
        if (!cell_in(s, 0, a, s.d.NI) || !cell_in(s, 1, b, s.d.NJ) || !cell_in(s, 2, c, s.d.NK)) continue;
        int k = idx_cell(s, a, b, c);
        if (sp.g_str[k] == kEmpty) continue;
        uint32_t e = sp.g_end[k];
        for (uint32_t j = sp.g_str[k]; j < e; j++) f((int)j);
      }
}

/* the (2 rng + 1)^3 cells around (x,y,z), same order (KERNEL_clc_volume_update3D
 * scans +-3 cells, functions_MASS.cuh:190-192) */
template <class F>
static void for_candidates_r(const orc_state& s, const Species& sp, double x, double y, double z, int rng, F&& f) {
  int ic, jc, kc;
  cell_of(s, x, y, z, ic, jc, kc);
  for (int dz = -rng; dz <= rng; dz++)
    for (int dy = -rng; dy <= rng; dy++)
      for (int dx = -rng; dx <= rng; dx++) {
        int a = ic + dx, b = jc + dy, c = kc + dz;
        if (a < 0 || a > s.d.NI - 1 || b < 0 || b > s.d.NJ - 1 || c < 0 || c > s.d.NK - 1) continue;
        int k = idx_cell(s, a, b, c);
        if (sp.g_str[k] == kEmpty) continue;
        uint32_t e = sp.g_end[k];
        for (uint32_t j = sp.g_str[k]; j < e; j++) f((int)j);
      }
}

static inline double dist_of(double dx, double dy, double dz) {
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

/* setup: find_minmax over all particles but the last (function_init.cuh:
 * 916-924), margins applied in the reference's order, dcell = 1.1*kappa*h0
 * with h0 the first particle's h (ISPH.cuh:93-108), NI = (int)(span/dcell)+1
 * (ISPH.cuh:111-113), num_cells = NI*NJ*NK+1 (functions_NNPS.cuh:357-375). */
static void setup_domain(orc_state& s, const part1* P, int n) {
  sophia_domain& d = s.d;
  const sophia_params& p = s.p;
  if (p.domain_fixed) {
    d.x_min = p.domain_min[0]; d.x_max = p.domain_max[0];
    d.y_min = p.domain_min[1]; d.y_max = p.domain_max[1];
    d.z_min = p.domain_min[2]; d.z_max = p.domain_max[2];
  } else {
    double mnx = P[0].x, mxx = P[0].x, mny = P[0].y, mxy = P[0].y, mnz = P[0].z, mxz = P[0].z;
    for (int i = 0; i < n - 1; i++) {
      if (P[i].x < mnx) mnx = P[i].x;
      if (P[i].x > mxx) mxx = P[i].x;
      if (P[i].y < mny) mny = P[i].y;
      if (P[i].y > mxy) mxy = P[i].y;
      if (P[i].z < mnz) mnz = P[i].z;
      if (P[i].z > mxz) mxz = P[i].z;
    }
    mnx -= (mxx - mnx) * p.xmargin_m;
    mxx += (mxx - mnx) * p.xmargin_p;
    mny -= (mxy - mny) * p.ymargin_m;
    mxy += (mxy - mny) * p.ymargin_p;
    mnz -= (mxz - mnz) * p.zmargin_m;
    mxz += (mxz - mnz) * p.zmargin_p;
    d.x_min = mnx; d.x_max = mxx; d.y_min = mny; d.y_max = mxy; d.z_min = mnz; d.z_max = mxz;
  }
  double h0 = P[0].h;
  d.dcell = 1.1 * p.kappa * h0;
  d.NI = (int)((d.x_max - d.x_min) / d.dcell) + 1;
  d.NJ = (int)((d.y_max - d.y_min) / d.dcell) + 1;
  d.NK = (int)((d.z_max - d.z_min) / d.dcell) + 1;
  // periodic axes (fixed domain): whole cells of width >= dcell tile the
  // period, so the 27-cell stencil with wrapped indices covers kappa h
  s.pmask = p.domain_fixed ? (p.periodic_mask & (p.dim == 2 ? 3 : 7)) : 0;
  int* N[3] = {&d.NI, &d.NJ, &d.NK};
  const double span[3] = {d.x_max - d.x_min, d.y_max - d.y_min, d.z_max - d.z_min};
  for (int a = 0; a < 3; a++)
    if (s.pmask >> a & 1) {
      s.plen[a] = span[a];
      *N[a] = std::max(1, (int)(span[a] / d.dcell));
    }
  if (p.flag_z_index == 0) {
    d.num_cells = d.NI * d.NJ * d.NK + 1;
  } else {
    int m = std::max(std::max(d.NI, d.NJ), d.NK);
    uint64_t last = p.dim == 2 ? morton2(m - 1, m - 1) : morton3(m - 1, m - 1, m - 1);
    d.num_cells = (int)last + 2;  // every key plus one dummy slot
  }
}

/* ------------------------------------------------------------------------ */
extern "C" {

void orc_set_threads(int32_t n) {
  g_threads = n;
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#endif
}
int32_t orc_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

orc_state* orc_create(const sophia_params* p, const part1* particles, int32_t n) {
  if (!p || !particles || n <= 0) return nullptr;
  orc_state* s = new orc_state();
  s->p = *p;
  s->count = 0;
  s->time = p->time;
  setup_domain(*s, particles, n);
  // read_sph_dem (function_init.cuh:893-911): SPH = p_type <= 1000
  for (int i = 0; i < n; i++) {
    part1 q = particles[i];
    if (p->open_boundary == 1) {
      // c_initial_inner_outer_particle_single_sph/_dem (functions_NNPS.cuh:
      // 8-76, 138-200): reset, then the inlet [L1, L2) and outlet [L3, L4)
      // z bands of the fluid become buffer particles
      const double* L = p->ob_band;
      q.i_type = 1;
      q.buffer_type = 0;
      if (q.p_type < 1000 && q.p_type >= 1) {
        if (q.z >= (L[0] - 1e-6) && q.z < (L[1] - 1e-6)) { q.buffer_type = 1; q.i_type = 2; }
        if (q.z >= (L[2] - 1e-6) && q.z < (L[3] - 1e-6)) { q.buffer_type = 2; q.i_type = 2; }
      }
    }
    if (q.p_type <= 1000) s->sph.P1.push_back(q);
    else s->dem.P1.push_back(q);
  }
  for (Species* sp : {&s->sph, &s->dem}) {
    int m = sp->n();
    sp->SP1 = sp->P1;
    sp->P2.assign(m, part2{});
    sp->SP2.assign(m, part2{});
    sp->P3.assign(m, part3{});
    sp->key_in.assign(m, 0);
    sp->key.assign(m, 0);
    sp->order.assign(m, 0);
    sp->g_str.assign(s->d.num_cells + 1, kEmpty);
    sp->g_end.assign(s->d.num_cells + 1, 0);
  }
  s->d.n_sph = s->sph.n();
  s->d.n_dem = s->dem.n();
  return s;
}

void orc_destroy(orc_state* s) { delete s; }
void orc_set_table(orc_state* s, const sophia_table* t) {
  s->table = *t;
  s->has_table = true;
}
void orc_get_domain(const orc_state* s, sophia_domain* d) { *d = s->d; }
void orc_get_clock(const orc_state* s, int64_t* count, double* time) {
  if (count) *count = s->count;
  if (time) *time = s->time;
}
int orc_count(const orc_state* s, int32_t species) {
  return species == SOPHIA_SPH ? s->sph.n() : s->dem.n();
}

/* ------------------------------------------------------------------------ */
/* NNPS: KERNEL_set_alpha_sph/_dem (functions_ALE.cuh:26-69), cell keys     */
/* (functions_NNPS.cuh:762-821), stable SortPairs (ISPH_Calc_seperate.cuh:  */
/* 82-85), KERNEL_reorder (functions_NNPS.cuh:670-730), P3 reset and        */
/* P1 <- SP1 (ISPH_Calc_seperate.cuh:102-113)                               */
/* ------------------------------------------------------------------------ */
static void nnps_species(orc_state& s, Species& sp, bool is_sph) {
  int n = sp.n();
  for (int i = 0; i < n; i++) {
    part1& q = sp.P1[i];
    if (q.i_type == 3) continue;
    double e = (q.p_type <= 1000 && s.p.eulerian_sph) ? 0.0 : 1.0;
    q.elix = q.eliy = q.eliz = e;
  }
  for (int i = 0; i < n; i++) {
    const part1& q = sp.P1[i];
    if (q.i_type == 3) {
      sp.key_in[i] = s.d.num_cells - 1;
    } else {
      int a, b, c;
      cell_of(s, q.x, q.y, q.z, a, b, c);
      sp.key_in[i] = idx_cell(s, a, b, c);
    }
  }
  std::vector<int32_t> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return sp.key_in[a] < sp.key_in[b]; });
  std::fill(sp.g_str.begin(), sp.g_str.end(), kEmpty);
  for (int i = 0; i < n; i++) {
    sp.order[i] = idx[i];
    sp.key[i] = sp.key_in[idx[i]];
  }
  for (int i = 0; i < n; i++) {
    int h = sp.key[i];
    if (i == 0 || h != sp.key[i - 1]) {
      sp.g_str[h] = i;
      if (i > 0) sp.g_end[sp.key[i - 1]] = i;
    }
    if (i == n - 1) sp.g_end[h] = i + 1;
  }
  for (int i = 0; i < n; i++) {
    sp.SP1[i] = sp.P1[sp.order[i]];
    sp.SP2[i] = sp.P2[sp.order[i]];
  }
  std::fill(sp.P3.begin(), sp.P3.end(), part3{});
  sp.P1 = sp.SP1;
}

void orc_nnps(orc_state* s) {
  int stride = std::max(1, s->p.decouple_stride);
  bool sph_on = (s->count % stride) == 0;
  if (sph_on) nnps_species(*s, s->sph, true);
  nnps_species(*s, s->dem, false);
  s->after_corrector = false;
}

/* ------------------------------------------------------------------------ */
/* PREP                                                                     */
/* ------------------------------------------------------------------------ */
static inline double vol_of(const sophia_params& p, double h) {
  double hv = h / p.h_over_dx;
  return hv * hv * hv;
}

/* KERNEL_clc_mass_init_sph (functions_MASS.cuh:24-43) */
static void mass_init_sph(orc_state& s) {
  Species& sp = s.sph;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    part1& q = sp.SP1[i];
    if (q.p_type > 1000 || q.i_type > kITypeCrt) continue;
    q.vol0 = std::pow(q.h / s.p.h_over_dx, s.p.dim);
    q.vol = q.vol0;
    sp.SP2[i].rho0 = q.rho;
    q.cond = conductivity(s.p, s.has_table ? &s.table : g_no_table, q.p_type, q.temp);
  }
}

/* KERNEL_clc_correction_KGC_3D_sph (functions_PREP2.cuh:452-568) */
static void kgc_3d(orc_state& s) {
  Species& sp = s.sph;
  const int kt = s.p.kernel_type, dim = s.p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt || qi.p_type > 1000) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = s.p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z;
    double xx = 0, yy = 0, zz = 0, xy = 0, yz = 0, zx = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (qj.p_type > 1000) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1), dz = pd(s, zi, qj.z, 2);
      double r = dist_of(dx, dy, dz);
      if (!(r > 0 && r < R)) return;
      double dw = knl_dw(kt, A, h, r);
      double mtd = qj.m * dw;
      double rtd = 1.0 / (qj.rho * r);
      double t;
      t = -mtd * dx * dx; t *= rtd; xx += t;
      t = -mtd * dy * dx; t *= rtd; xy += t;
      t = -mtd * dy * dy; t *= rtd; yy += t;
      t = -mtd * dx * dz; t *= rtd; zx += t;
      t = -mtd * dy * dz; t *= rtd; yz += t;
      t = -mtd * dz * dz; t *= rtd; zz += t;
    });
    double det = xx * (yy * zz - yz * yz);
    det -= xy * (xy * zz - yz * zx);
    det += zx * (xy * yz - yy * zx);
    double (*C)[3] = sp.P3[i].Cm;
    if (std::abs(det) > kMinDet) {
      double r = 1.0 / det;
      C[0][0] = (yy * zz - yz * yz) * r;
      C[0][1] = (zx * yz - xy * zz) * r;
      C[0][2] = (xy * yz - zx * yy) * r;
      C[1][0] = (zx * yz - xy * zz) * r;
      C[1][1] = (xx * zz - zx * zx) * r;
      C[1][2] = (zx * xy - xx * yz) * r;
      C[2][0] = (xy * yz - zx * yy) * r;
      C[2][1] = (zx * xy - xx * yz) * r;
      C[2][2] = (xx * yy - xy * xy) * r;
    } else {
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) C[a][b] = (a == b);
    }
  }
}

/* apply_gradient_correction_3D (functions_KNL.cuh:763-783), KGC branch */
static inline void kgc_apply(const orc_state& s, const double C[3][3], double& gx, double& gy, double& gz) {
  if (s.p.kgc_solve != 1) return;
  double x = C[0][0] * gx + C[0][1] * gy + C[0][2] * gz;
  double y = C[1][0] * gx + C[1][1] * gy + C[1][2] * gz;
  double z = C[2][0] * gx + C[2][1] * gy + C[2][2] * gz;
  gx = x; gy = y; gz = z;
}

/* KERNEL_clc_prep3D_coupling_sph (functions_PREP2.cuh:3190-3364) */
static void coupling_prep_sph(orc_state& s) {
  Species& sp = s.sph;
  Species& dm = s.dem;
  const int kt = s.p.kernel_type, dim = s.p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt) continue;
    qi.dDEMpor_prev = qi.dDEMpor;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = s.p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z;
    double flt = 0, por = 0, dfx = 0, dfy = 0, dfz = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      double r = dist_of(pd(s, xi, qj.x, 0), pd(s, yi, qj.y, 1), pd(s, zi, qj.z, 2)) + 1e-20;
      if (!(r < R)) return;
      double w = knl_w(kt, A, h, r);
      if (qi.p_type <= 1000 && qj.p_type <= 1000) flt += qj.m / qj.rho * w * (qj.p_type == 1);
    });
    if (dm.n() > 0) {
      for_candidates(s, dm, xi, yi, zi, [&](int j) {
        const part1& qj = dm.SP1[j];
        double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1), dz = pd(s, zi, qj.z, 2);
        double r = dist_of(dx, dy, dz) + 1e-20;
        if (!(r < R)) return;
        double w = knl_w(kt, A, h, r);
        double dw = knl_dw(kt, A, h, r);
        double gx = dw * dx / r, gy = dw * dy / r, gz = dw * dz / r;
        if (qi.p_type <= 1000 && qj.p_type > 1000) {
          double rj = qj.rad;
          por += w * 4.18879 * rj * rj * rj;
          dfx += gx * qj.ux * 4.18879 * rj * rj * rj;
          dfy += gy * qj.uy * 4.18879 * rj * rj * rj;
          dfz += gz * qj.uz * 4.18879 * rj * rj * rj;
        }
      });
    }
    if (qi.p_type <= 1000) {
      qi.DEMpor = 1.0 - (por / (flt + 1e-20));
      qi.dDEMpor = (dfx + dfy + dfz) / (flt + 1e-20);
    }
  }
}

/* KERNEL_clc_prep3D_coupling_dem (functions_PREP2.cuh:3367-3523) */
static void coupling_prep_dem(orc_state& s) {
  Species& sp = s.sph;
  Species& dm = s.dem;
  const int kt = s.p.kernel_type, dim = s.p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < dm.n(); i++) {
    part1& qi = dm.SP1[i];
    if (qi.i_type > kITypeCrt) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = s.p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z;
    double fltd = 0, fltd2 = 0, flt = 0, por = 0;
    if (sp.n() > 0) {
      for_candidates(s, sp, xi, yi, zi, [&](int j) {
        const part1& qj = sp.SP1[j];
        double r = dist_of(xi - qj.x, yi - qj.y, zi - qj.z) + 1e-20;
        if (!(r < R)) return;
        double w = knl_w(kt, A, h, r);
        if (qi.p_type > 1000) {
          int t = qj.p_type;
          if (t <= 1000 && t != 0 && t != SOPHIA_PT_MOVING && t == 1) {
            fltd += qj.m / qj.rho * w;
            fltd2 += qj.m / qj.rho * w;
          }
          if (t <= 1000 && t != 0 && t != SOPHIA_PT_MOVING && t == 3) flt += qj.m / qj.rho * w;
        }
      });
    }
    for_candidates(s, dm, xi, yi, zi, [&](int j) {
      const part1& qj = dm.SP1[j];
      double r = dist_of(xi - qj.x, yi - qj.y, zi - qj.z) + 1e-20;
      if (!(r < R)) return;
      double w = knl_w(kt, A, h, r);
      if (qi.p_type > 1000 && qj.p_type > 1000) por += w * 4.18879 * qj.rad * qj.rad * qj.rad;
    });
    if (qi.p_type > 1000) {
      qi.flt_s = flt;
      qi.flt_sd = fltd;
      qi.flt_sd_2 = fltd2;
      qi.DEMpor = 1.0 - (por / (flt + fltd + 1.0e-20));
    }
  }
}

/* KERNEL_clc_prep3D_prep_sph (functions_PREP2.cuh:1623-1679) */
static void prep_prep_sph(orc_state& s) {
  Species& sp = s.sph;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    part1& q = sp.SP1[i];
    if (q.i_type > kITypeCrt || q.p_type > 1000) continue;
    int t = q.p_type;
    double hv = q.h / s.p.h_over_dx;
    double rt = (t == -3 || t == 0 || t == 3 || t == 1) ? s.p.rho_fluid : s.p.rho_water;
    if (t == 9 && q.z < 0.05) rt = s.p.rho_fluid;
    double m_ref = rt * hv * hv * hv * q.DEMpor;
    if (s.p.dim == 2) sp.SP2[i].rho_ref = m_ref / (hv * hv);
    if (s.p.dim == 3) sp.SP2[i].rho_ref = m_ref / (hv * hv * hv);
    if (s.p.rho_type == 1) {
      sp.SP2[i].rho0 = m_ref / (hv * hv * hv);
      q.rho = m_ref / (hv * hv * hv);
      q.m = m_ref;
    }
  }
}

/* KERNEL_clc_prep3D_sph (functions_PREP2.cuh:1985-2227): Shepard filter and
 * Antuono delta-SPH density gradient (turbulence / surface terms off) */
static void prep3d_sph(orc_state& s) {
  Species& sp = s.sph;
  const int kt = s.p.kernel_type, dim = s.p.dim;
  int stride = std::max(1, s.p.decouple_stride);
  bool filt = (s.count % ((int64_t)s.p.freq_filt * stride)) == 0;
  std::vector<double> fltv(sp.n()), gx(sp.n()), gy(sp.n()), gz(sp.n());
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt || qi.p_type > 1000) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = s.p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z;
    double rhoi = qi.rho, rrefi = sp.SP2[i].rho_ref;
    double flt = 0, rx = 0, ry = 0, rz = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (qj.p_type > 1000) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1), dz = pd(s, zi, qj.z, 2);
      double r = dist_of(dx, dy, dz) + 1e-20;
      if (!(r < R)) return;
      double w = knl_w(kt, A, h, r), dw = knl_dw(kt, A, h, r);
      double tx = dw * dx / r, ty = dw * dy / r, tz = dw * dz / r;
      if (filt) flt += qj.m / qj.rho * w;
      if (s.p.delSPH_solve == 2) {
        kgc_apply(s, sp.P3[i].Cm, tx, ty, tz);
        double c = -(qj.rho / sp.SP2[j].rho_ref - rhoi / rrefi) * (qj.m / qj.rho);
        rx += c * tx;
        ry += c * ty;
        rz += c * tz;
      }
    });
    fltv[i] = flt; gx[i] = rx; gy[i] = ry; gz[i] = rz;
  }
  for (int i = 0; i < sp.n(); i++) {
    part1& q = sp.SP1[i];
    if (q.i_type > kITypeCrt || q.p_type > 1000) continue;
    if (filt) q.flt_s = fltv[i];
    if (s.p.delSPH_solve == 2) {
      q.grad_rhox = gx[i];
      q.grad_rhoy = gy[i];
      q.grad_rhoz = gz[i];
      q.test1 = gy[i];
      q.test5 = gz[i];
    }
  }
}

/* Flux walls (sophia_params.heat_flux; no counterpart in the compiled      */
/* reference): S_w = sum over fluid k within kappa*h_w of V_k W(r_kw, h_w),  */
/* the normaliser that makes the fluid receive heat_flux times the walls'   */
/* face area exactly (sum_i V_i q A_w W_iw / S_w = q A_w).                  */
static bool flux_receiver(const sophia_params& p, const part1& q) {
  return q.i_type <= kITypeCrt && !(p.open_boundary > 0 && q.buffer_type > 0) && q.p_type > 0 && q.p_type < 1000;
}
static void flux_walls(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  s.flux_S.assign(sp.n(), 0.0);
  if (!(p.con_solve == 1 && p.heat_flux != 0.0)) return;
  const int kt = p.kernel_type, dim = p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& qi = sp.SP1[i];
    if (!in_flux_box(p, qi)) continue;
    const double hi = qi.h, A = knl_tmpA(kt, dim, hi), R = p.kappa * hi;
    double S = 0;
    for_candidates(s, sp, qi.x, qi.y, qi.z, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (!flux_receiver(p, qj)) return;
      double r = dist_of(pd(s, qi.x, qj.x, 0), pd(s, qi.y, qj.y, 1), pd(s, qi.z, qj.z, 2));
      if (!(r < R)) return;
      S += qj.m / qj.rho * knl_w(kt, A, hi, r);
    });
    s.flux_S[i] = S;
  }
}

void orc_prep(orc_state* s) {
  int stride = std::max(1, s->p.decouple_stride);
  bool sph_on = (s->count % stride) == 0;
  if (sph_on && s->p.dim == 3) mass_init_sph(*s);
  if (sph_on && s->p.kgc_solve > 0 && s->p.dim == 3) kgc_3d(*s);
  if (s->p.dim == 3) {
    if (sph_on) coupling_prep_sph(*s);
    coupling_prep_dem(*s);
    if (sph_on) {
      prep_prep_sph(*s);
      prep3d_sph(*s);
      flux_walls(*s);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* INTERACTION: KERNEL_advection_force3D_sph (functions_PPE2.cuh:1251-1607) */
/* terms: pressure-gradient density for coupling, laminar viscous,          */
/* artificial viscosity, Eulerian advection u.grad(u), conduction, gravity  */
/* ------------------------------------------------------------------------ */
void orc_interaction(orc_state* s) {
  int stride = std::max(1, s->p.decouple_stride);
  if ((s->count % stride) != 0 || s->p.dim != 3) return;
  Species& sp = s->sph;
  const sophia_params& p = s->p;
  const int kt = p.kernel_type, dim = p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt) continue;
    if (p.open_boundary > 0 && qi.buffer_type > 0) continue;
    if (qi.p_type >= 1000 || qi.p_type <= 0) continue;
    double xi = qi.x, yi = qi.y, zi = qi.z, uxi = qi.ux, uyi = qi.uy, uzi = qi.uz;
    double hi = qi.h, pi = qi.pres, mi = qi.m, rhoi = qi.rho, pori = qi.DEMpor, tempi = qi.temp;
    double A = knl_tmpA(kt, dim, hi), R = p.kappa * hi;
    const sophia_table* tab = s->has_table ? &s->table : g_no_table;
    double kci = conductivity(p, tab, qi.p_type, qi.temp), cpi = heat_capacity(p, qi.p_type);
    double visi = viscosity(p, tab, qi.p_type, qi.temp) + sp.P3[i].vis_t;
    double tx = 0, ty = 0, tz = 0, Rc = 0, px = 0, py = 0, pz = 0;
    double ex = 0, ey = 0, ez = 0, et = 0, fl = 0;
    const bool flux = p.con_solve == 1 && p.heat_flux != 0.0 && (int)s->flux_S.size() == sp.n();
    for_candidates(*s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (!(qj.p_type < 1000)) return;
      double dx = pd(*s, xi, qj.x, 0), dy = pd(*s, yi, qj.y, 1), dz = pd(*s, zi, qj.z, 2);
      if (flux && s->flux_S[j] > 0.0) {
        // wall heat flux: q A_w W_iw / S_w (the wall's face area from its own h)
        double rw = dist_of(dx, dy, dz);
        if (rw < R) fl += p.heat_flux * std::pow(qj.h / p.h_over_dx, 2) * knl_w(kt, A, hi, rw) / s->flux_S[j];
      }
      double r = dist_of(dx, dy, dz) + 1e-20;
      if (!(r < R)) return;
      double dw = knl_dw(kt, A, hi, r);
      double gx = dw * dx / r, gy = dw * dy / r, gz = dw * dz / r;
      kgc_apply(*s, sp.P3[i].Cm, gx, gy, gz);
      double mj = qj.m, rhoj = qj.rho, pj = qj.pres;
      if (p.fp_solve) {
        double Cp = -pori * mj * (p.pgrad_form ? pj - pi : pi + pj) / (rhoi * rhoj);
        px += rhoi / pori * Cp * gx;
        py += rhoi / pori * Cp * gy;
        pz += rhoi / pori * Cp * gz;
      }
      if (p.fv_solve) {
        double visj = viscosity(p, tab, qj.p_type, qj.temp) + sp.P3[j].vis_t;
        double Cv = dx * gx + dy * gy + dz * gz;
        Cv *= (visi * visj) / (visi + visj);
        Cv *= 4 * pori * (mj / (rhoi * rhoj));
        Cv /= r;
        Cv /= r;
        tx += Cv * (uxi - qj.ux);
        ty += Cv * (uyi - qj.uy);
        tz += Cv * (uzi - qj.uz);
      }
      if (p.fva_solve) {
        double uij_xij = (uxi - qj.ux) * dx + (uyi - qj.uy) * dy + (uzi - qj.uz) * dz;
        if (uij_xij < 0) {
          double h_ij = (hi + qj.h) * 0.5;
          double phi = h_ij * uij_xij;
          phi /= (r * r + 0.01 * h_ij * h_ij);
          double P = mi * phi * (-0.1 * p.soundspeed + 0.02 * phi);
          P /= (rhoi + rhoj);
          P *= 0.5;
          tx += -P * gx;
          ty += -P * gy;
          tz += -P * gz;
        }
      }
      if (qi.elix < 1.0 || qi.eliy < 1.0) {
        double c = mj / rhoj;
        ex += (uxi * (qj.ux - uxi) * gx + uyi * (qj.ux - uxi) * gy + uzi * (qj.ux - uxi) * gz) * c * (1 - qi.elix);
        ey += (uxi * (qj.uy - uyi) * gx + uyi * (qj.uy - uyi) * gy + uzi * (qj.uy - uyi) * gz) * c * (1 - qi.eliy);
        ez += (uxi * (qj.uz - uzi) * gx + uyi * (qj.uz - uzi) * gy + uzi * (qj.uz - uzi) * gz) * c * (1 - qi.eliz);
        if (p.con_solve) {
          double tj = qj.temp;
          et += (uxi * (tj - tempi) * gx + uyi * (tj - tempi) * gy + uzi * (tj - tempi) * gz) * c * (1 - qi.elix);
        }
      }
      if (p.con_solve) {
        double kcj = conductivity(p, tab, qj.p_type, qj.temp);
        double H = 4.0 * pori * mj * kcj * kci * (tempi - qj.temp) * dw;
        H /= (r + 1e-10) * rhoi * rhoj * (kci + kcj);
        Rc += H;
      }
    });
    if (p.fg_solve) tz += -p.gravity;
    part3& o = sp.P3[i];
    o.ftotalx = 1 * (tx - ex);
    o.ftotaly = 1 * (ty - ey);
    o.ftotalz = 1 * (tz - ez);
    if (p.con_solve) o.dtemp = Rc / cpi * (qi.p_type != -1) - et + fl / (rhoi * cpi);
    o.fsx = o.fsy = o.fsz = 0;
    qi.pgf_x = px;
    qi.pgf_y = py;
    qi.pgf_z = pz;
    o.ftotal = std::sqrt(tx * tx + ty * ty + tz * tz);
    if (p.con_solve) o.denthalpy = Rc;
  }
}

/* ------------------------------------------------------------------------ */
/* DEM_INTERACTION: KERNEL_DEM_interaction3D_dem                            */
/* (functions_DEM_INTERACTION2.cuh:888-1766), Hertz-Mindlin with 16-slot    */
/* shear history keyed by dem_idx; history read from SP1, written to P1     */
/* ------------------------------------------------------------------------ */
void orc_dem_interaction(orc_state* s) {
  if (s->p.dim != 3) return;
  Species& dm = s->dem;
  const sophia_params& p = s->p;
  const int slots = std::min(16, std::max(1, p.dem_max_contacts));
  const double G = p.dem_shear_modulus, nu = p.dem_poisson;
  const double E = G * (2 * (1 + nu));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < dm.n(); i++) {
    const part1& qi = dm.SP1[i];
    part1& ti = dm.P1[i];
    if (qi.i_type != 1) continue;
    if (qi.p_type <= 1000) continue;
    double xi = qi.x, yi = qi.y, zi = qi.z, uxi = qi.ux, uyi = qi.uy, uzi = qi.uz;
    double wxi = qi.wx, wyi = qi.wy, wzi = qi.wz, radi = qi.rad, mi = qi.m, rii = qi.ri;
    double tx = 0, ty = 0, tz = 0, qx = 0, qy = 0, qz = 0;
    // heat (con_solve, :948-1030, 1080-1194, 1300-1315): the outer shell of
    // thickness rad/6 exchanges heat with neighbouring spheres
    const bool heat = p.con_solve == 1;
    const double tempi = qi.temp, rhoi = qi.rho, kci = p.k_dem, cpi = p.cp_dem, emi = dem_emissivity(tempi);
    const double dr2 = radi / 6.0;
    const double vol_matrix = (4.0 / 3.0) * 3.141592 * (radi * radi * radi - (radi - dr2) * (radi - dr2) * (radi - dr2));
    double dt_rad = 0, dt_con = 0;
    for_candidates(*s, dm, xi, yi, zi, [&](int j) {
      const part1& qj = dm.SP1[j];
      if (!(qj.p_type > 1000)) return;
      double radj = qj.rad;
      int didx = qj.dem_idx;
      double dist = dist_of(xi - qj.x, yi - qj.y, zi - qj.z);
      double ovlp = radi + radj - dist;
      if (!(dist < 3.5 * radi && dist > 0.0)) return;
      const double radij_h = 1 / (1 / radi + 1 / radj);
      if (heat) {
        // radiation between the facing caps of a double cone (:1085-1152);
        // its radius 0.56 (2 R*) (1 - 0.401038)^(-1/3), with the reference's
        // integer -1/3 giving exponent 0 in compat mode
        const double tempj = qj.temp, emj = dem_emissivity(tempj);
        const double ai = radi / dist, aj = radj / dist, dA = aj - ai;
        const double cone = (p.reference_compat & SOPHIA_COMPAT_NUSSELT) ? 1.0 : std::pow(1 - 0.401038, -1.0 / 3.0);
        const double dc_r = 0.56 * (2 * radij_h) * cone;
        const double dc_ar = 3.14159 * dc_r * dc_r;
        const double di_max = 0.5 * (std::sqrt(1 + 4.0 * dc_ar / (3.14159 * dist * dist * (1 - dA * dA))) - dA);
        const double cos_i = (1 - dA * dA) / (2 * di_max) - dA;
        const double cos_j = (1 - dA * dA) / (2 * (di_max + dA)) + dA;
        const double sin_i = std::sqrt(1 - cos_i * cos_i), sin_j = std::sqrt(1 - cos_j * cos_j);
        const double ri_cr = radi * sin_i, rj_cr = radj * sin_j;
        const double d_cr = dist - radi * cos_i - radj * cos_j;
        const double z_cr = 1 + (4 * rj_cr * rj_cr + d_cr * d_cr) / (ri_cr * ri_cr);
        const double F_cr = 0.5 * (z_cr - std::sqrt(z_cr * z_cr - 4.0 * rj_cr * rj_cr / (ri_cr * ri_cr)));
        const double F_ij = (1 + cos_i) * F_cr * 0.5;
        const double ar_i = 2.0 * 3.14159 * radi * radi * (1 - cos_i);
        const double ar_j = 2.0 * 3.14159 * radj * radj * (1 - cos_j);
        const double ar_ratio = ar_j / ar_i;
        const double Y_ij = F_ij + (1 - F_ij) * (ar_ratio / (1 + ar_ratio));
        const double em_inv = (1 - emi) / (emi * ar_i) + (1 - emj) / (emj * ar_j) + 1 / (Y_ij * ar_i);
        const double t4 = tempi * tempi * tempi * tempi - tempj * tempj * tempj * tempj;
        dt_rad += -5.6704e-8 * t4 / (em_inv * (rhoi * vol_matrix) * cpi);
      }
      if (!(dist < 2.01 * radi)) return;
      if (dist > 0 && ovlp > 0) {
        if (heat) {
          // contact conduction through the overlap circle (:1300-1315)
          const double kc_ij = 1 / (1 / kci + 1 / p.k_dem);
          dt_con += -4.0 * kc_ij * std::sqrt((2 * ovlp) * (radij_h - 0.5 * ovlp)) * (tempi - qj.temp) /
                    ((rhoi * vol_matrix) * cpi);
        }
        double dsx = 0, dsy = 0, dsz = 0;
        int slot = -1;
        for (int k = 0; k < slots; k++)
          if (qi.ct_idx[k] == didx) {
            dsx = qi.del_s[k][0];
            dsy = qi.del_s[k][1];
            dsz = qi.del_s[k][2];
            slot = k;
            break;
          }
        if (slot < 0) {
          for (int k = 0; k < slots; k++)
            if (qi.ct_idx[k] < 1) {
              ti.ct_idx[k] = didx;
              slot = k;
              break;
            }
        }
        double nx = (qj.x - xi) / dist, ny = (qj.y - yi) / dist, nz = (qj.z - zi) / dist;
        double ov = ovlp;
        double rcxi = (radi - 0.5 * ov) * nx, rcyi = (radi - 0.5 * ov) * ny, rczi = (radi - 0.5 * ov) * nz;
        double rcxj = -(radj - 0.5 * ov) * nx, rcyj = -(radj - 0.5 * ov) * ny, rczj = -(radj - 0.5 * ov) * nz;
        double uroxi = wyi * rczi - wzi * rcyi, uroyi = wzi * rcxi - wxi * rczi, urozi = wxi * rcyi - wyi * rcxi;
        double uroxj = qj.wy * rczj - qj.wz * rcyj, uroyj = qj.wz * rcxj - qj.wx * rczj, urozj = qj.wx * rcyj - qj.wy * rcxj;
        double urx = (qj.ux + uroxj) - (uxi + uroxi);
        double ury = (qj.uy + uroyj) - (uyi + uroyi);
        double urz = (qj.uz + urozj) - (uzi + urozi);
        double un = nx * urx + ny * ury + nz * urz;
        double unx = un * nx, uny = un * ny, unz = un * nz;
        double usx = urx - unx, usy = ury - uny, usz = urz - unz;
        dsx += usx * p.dt;
        dsy += usy * p.dt;
        dsz += usz * p.dt;
        double usm = std::sqrt(usx * usx + usy * usy + usz * usz) + 1e-20;
        double sx = usx / usm, sy = usy / usm, sz = usz / usm;
        double ymij = 1 / (((1 - nu * nu) / E) + ((1 - nu * nu) / E));
        double Gij = 1 / (((2 - nu) / G) + ((2 - nu) / G));
        double radij = 1 / (1 / radi + 1 / radj);
        double mij = 1 / (1 / mi + 1 / qj.m);
        double Cspr = p.dem_hertz_factor * ymij * std::sqrt(radij) * std::sqrt(ov) * ov / mi;
        double Csprs = 8.0 * Gij * std::sqrt(radij * ov) / mi;
        double Kn = 2 * ymij * std::sqrt(radij * ov);
        double Kd = std::sqrt(mij * Kn);
        double Cd = p.dem_damping * Kd / mi;
        double Kds = std::sqrt(mij * Csprs * mi);
        double Cds = p.dem_damping * Kds / mi;
        tx += Cspr * (xi - qj.x) / dist;
        ty += Cspr * (yi - qj.y) / dist;
        tz += Cspr * (zi - qj.z) / dist;
        tx += Cd * unx;
        ty += Cd * uny;
        tz += Cd * unz;
        double fssx = Csprs * dsx + Cds * usx, fssy = Csprs * dsy + Cds * usy, fssz = Csprs * dsz + Cds * usz;
        double sgx = 1.0 * (fssx >= 0) - 1.0 * (fssx < 0);
        double sgy = 1.0 * (fssy >= 0) - 1.0 * (fssy < 0);
        double sgz = 1.0 * (fssz >= 0) - 1.0 * (fssz < 0);
        double mfss = std::sqrt(fssx * fssx + fssy * fssy + fssz * fssz);
        double ffx = p.dem_friction * Cspr * sx, ffy = p.dem_friction * Cspr * sy, ffz = p.dem_friction * Cspr * sz;
        double mfsf = std::sqrt(ffx * ffx + ffy * ffy + ffz * ffz);
        double fsx = fssx * (mfss < mfsf) + sgx * std::abs(ffx) * (mfss >= mfsf);
        double fsy = fssy * (mfss < mfsf) + sgy * std::abs(ffy) * (mfss >= mfsf);
        double fsz = fssz * (mfss < mfsf) + sgz * std::abs(ffz) * (mfss >= mfsf);
        tx += fsx;
        ty += fsy;
        tz += fsz;
        qx += mi * (rcyi * fsz - rczi * fsy);
        qy += mi * (rczi * fsx - rcxi * fsz);
        qz += mi * (rcxi * fsy - rcyi * fsx);
        if (slot >= 0) {
          ti.ct_idx[slot] = didx;
          ti.del_s[slot][0] = dsx;
          ti.del_s[slot][1] = dsy;
          ti.del_s[slot][2] = dsz;
        }
      } else {
        for (int k = 0; k < slots; k++)
          if (qi.ct_idx[k] == didx) {
            ti.del_s[k][0] = ti.del_s[k][1] = ti.del_s[k][2] = 0.0;
            ti.ct_idx[k] = 0;
            break;
          }
      }
    });
    if (p.fg_solve) tz += -p.gravity;
    part3& o = dm.P3[i];
    if (heat) {
      // (:1725-1729) without the inner-shell terms: this path never advances
      // the shell temperatures temp1..3, so their conduction is left out
      dm.SP1[i].test1 = dt_rad;  // the kernel writes its read buffer (P1[i].test1, :1723-1724)
      dm.SP1[i].test2 = dt_con;
      o.dtemp = dt_rad + dt_con;
    }
    o.ftotalx = tx;
    o.ftotaly = ty;
    o.ftotalz = tz;
    o.torqx = qx / rii;
    o.torqy = qy / rii;
    o.torqz = qz / rii;
  }
}

/* ------------------------------------------------------------------------ */
/* SPH_DEM_COUPLING: KERNEL_DEM_coupling3D_dem (functions_SPH_DEM_COUPLING2 */
/* .cuh:627-1129) then KERNEL_SPH_coupling3D_sph (:1815-2019)               */
/* ------------------------------------------------------------------------ */
static void dem_coupling(orc_state& s) {
  Species& sp = s.sph;
  Species& dm = s.dem;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < dm.n(); i++) {
    part1& qi = dm.SP1[i];
    if (qi.i_type != 1 || qi.p_type <= 1000) continue;
    double xi = qi.x, yi = qi.y, zi = qi.z, uxi = qi.ux, uyi = qi.uy, uzi = qi.uz;
    double radi = qi.rad, di = 2 * radi, voli = (4.0 / 3.0) * 3.14159 * radi * radi * radi;
    double mi = qi.m;
    double flt_air = qi.flt_s, flt_fluid = qi.flt_sd;
    part3& o = dm.P3[i];
    double fx0 = o.ftotalx, fy0 = o.ftotaly, fz0 = o.ftotalz;
    double hi = qi.h, A = knl_tmpA(kt, dim, hi), R = p.kappa * hi;
    double tx = 0, ty = 0, tz = 0, ufx = 0, ufy = 0, ufz = 0, uax = 0, uay = 0, uaz = 0, por = 0;
    double tf = 0, ta = 0;  // kernel-weighted fluid and air temperature sums (:806-867)
    if (sp.n() > 0) {
      for_candidates(s, sp, xi, yi, zi, [&](int j) {
        const part1& qj = sp.SP1[j];
        if (!(qj.p_type <= 1000 && qj.DEMpor < 1.0)) return;
        double r = dist_of(xi - qj.x, yi - qj.y, zi - qj.z);
        if (!(r < R)) return;
        double w = knl_w(kt, A, hi, r);
        double mj = qj.m, rhoj = qj.rho;
        int t = qj.p_type;
        if (t == 1 && flt_fluid > 0.000001) {
          ufx += qj.ux * mj * w / rhoj;
          ufy += qj.uy * mj * w / rhoj;
          ufz += qj.uz * mj * w / rhoj;
          por += qj.DEMpor * mj * w / rhoj;
          tf += qj.temp * mj * w / rhoj;
        }
        if (t == 3 && flt_air > 0.000001) {
          uax += qj.ux * mj * w / rhoj;
          uay += qj.uy * mj * w / rhoj;
          uaz += qj.uz * mj * w / rhoj;
          por += qj.DEMpor * mj * w / rhoj;
          ta += qj.temp * mj * w / rhoj;
        }
        if (t != 0 && t != SOPHIA_PT_MOVING && t != -3 && t != 3) {
          if (flt_fluid > 0.01) {
            tx += qj.pgf_x * (mj / rhoj) * w * voli / (mi * flt_fluid + 1.0e-15);
            ty += qj.pgf_y * (mj / rhoj) * w * voli / (mi * flt_fluid + 1.0e-15);
            tz += qj.pgf_z * (mj / rhoj) * w * voli / (mi * flt_fluid + 1.0e-15);
          }
        }
      });
    }
    double pori = por / (qi.flt_sd + 1.0e-15);
    double vfx = ufx / (flt_fluid + 1.0e-15), vfy = ufy / (flt_fluid + 1.0e-15), vfz = ufz / (flt_fluid + 1.0e-15);
    double vax = uax / (flt_air + 1.0e-15), vay = uay / (flt_air + 1.0e-15), vaz = uaz / (flt_air + 1.0e-15);
    double vmx = (ufx + uax) / (flt_fluid + flt_air + 1.0e-15);
    double vmy = (ufy + uay) / (flt_fluid + flt_air + 1.0e-15);
    double vmz = (ufz + uaz) / (flt_fluid + flt_air + 1.0e-15);
    double rfx = vfx - uxi, rfy = vfy - uyi, rfz = vfz - uzi;
    double rax = vax - uxi, ray = vay - uyi, raz = vaz - uzi;
    double rmx = vmx - uxi, rmy = vmy - uyi, rmz = vmz - uzi;
    double f_fluid = flt_fluid / (flt_fluid + flt_air + 1.0e-15);
    double f_air = flt_air / (flt_fluid + flt_air + 1.0e-15);
    double magf = std::sqrt(rfx * rfx + rfy * rfy + rfz * rfz);
    double maga = std::sqrt(rax * rax + ray * ray + raz * raz);
    double magm = std::sqrt(rmx * rmx + rmy * rmy + rmz * rmz);
    qi.test6 = magf;
    double dfx = 0, dfy = 0, dfz = 0, dax = 0, day = 0, daz = 0;
    qi.DEMpor = pori;
    // convective heat (:996-1106): Ranz-Marshall Nu = 2 + 0.6 Re^0.5 Pr^(1/3),
    // h = Nu k / d over the specific surface 6/d; the sphere's k and
    // diffusivity by material (:708-718); no branch taken means no exchange
    const double tempi = qi.temp, rhoi = qi.rho;
    const int mat = radi < p.dem_mat_radius ? 1 : 0;
    const double ki = p.dem_k_mat[mat], cpi = ki / (p.dem_alpha_mat[mat] * rhoi);
    const double a_ht = 6.0 / di;
    const double T_f = tf / (flt_fluid + 1.0e-15) - tempi, T_a = ta / (flt_air + 1.0e-15) - tempi;
    const double T_m = (tf + ta) / (flt_fluid + flt_air + 1.0e-15) - tempi;
    double dq_vol = 0;
    double nu_f = p.mu_fluid / p.rho_fluid;  // vis_kif = VISCOSITY_AIR/DENSITY_AIR
    double rho_f = p.rho_fluid;
    if (magf > 1.0e-8 && (flt_fluid + flt_air) > 0.00001 && f_fluid > p.drag_switch) {
      double Re = di * magf * pori / (nu_f + 1.0e-15);
      double Cd = (0.63 + 4.8 / std::sqrt(Re)) * (0.63 + 4.8 / std::sqrt(Re));
      // ISPH (KERNEL_DEM_coupling3D_dem, :979-1009): Di Felice, eps^(2-beta);
      // WCSPH (KERNEL_DEM_coupling3D, :298-310): the Rong form of the other
      // regimes, eps^(-1-beta)
      const bool wc = p.solver_type == 0;
      double beta = wc ? 2.65 * (1.0 + pori) - (5.3 - 3.5 * pori) * pori * pori *
                             std::exp(-(1.5 - std::log10(Re)) * (1.5 - std::log10(Re)) / 2.0)
                       : 3.7 - 0.65 * std::exp(-(1.5 - std::log10(Re)) * (1.5 - std::log10(Re)) / 2.0);
      double k = (1.0 / 8.0) * Cd * rho_f * 3.14159 * di * di;
      double e = std::pow(pori, wc ? -1 - beta : 2 - beta);
      dfx = k * rfx * magf * e / mi;
      dfy = k * rfy * magf * e / mi;
      dfz = k * rfz * magf * e / mi;
      qi.Fdx_df = dfx;
      qi.Fdy_df = dfy;
      qi.Fdz_df = dfz;
      const double Nu = 2.0 + 0.6 * std::pow(Re, 0.5) * pr_third(p, p.pr_fluid);
      dq_vol = Nu * p.k_fluid / di * a_ht * T_f;
    }
    if (maga > 1.0e-8 && (flt_fluid + flt_air) > 0.00001 && f_air > p.drag_switch) {
      double Re = di * pori * maga / (nu_f + 1.0e-15);
      double Cd = (0.63 + 4.8 / std::sqrt(Re)) * (0.63 + 4.8 / std::sqrt(Re));
      double beta = 2.65 * (1.0 + pori) - (5.3 - 3.5 * pori) * pori * pori * std::exp(-(1.5 - std::log10(Re)) * (1.5 - std::log10(Re)) / 2.0);
      double k = (1.0 / 8.0) * Cd * p.rho_fluid * 3.14159 * di * di;
      double e = std::pow(pori, -1 - beta);
      dax = k * rax * maga * e / mi;
      day = k * ray * maga * e / mi;
      daz = k * raz * maga * e / mi;
      const double Nu = 2.0 + 0.6 * std::pow(Re, 0.5) * pr_third(p, p.pr_air);
      dq_vol = Nu * ki / di * a_ht * T_a;
    }
    if ((maga > 1.0e-8 || magf > 1.0e-8) && (flt_fluid + flt_air) > 0.00001 && f_air <= p.drag_switch &&
        f_fluid <= p.drag_switch) {
      double rho_m = p.rho_fluid * f_air + p.rho_water * f_fluid;
      double vis_m = p.mu_fluid * f_air + p.mu_water * f_fluid;
      double nu_m = vis_m / (rho_m + 1.0e-15);
      double Re = di * pori * magm / (nu_m + 1.0e-15);
      double Cd = (0.63 + 4.8 / std::sqrt(Re)) * (0.63 + 4.8 / std::sqrt(Re));
      double beta = 2.65 * (1.0 + pori) - (5.3 - 3.5 * pori) * pori * pori * std::exp(-(1.5 - std::log10(Re)) * (1.5 - std::log10(Re)) / 2.0);
      double k = (1.0 / 8.0) * Cd * rho_m * 3.14159 * di * di;
      double e = std::pow(pori, -1 - beta);
      dax = k * rmx * magm * e / mi;
      day = k * rmy * magm * e / mi;
      daz = k * rmz * magm * e / mi;
      const double Nu = 2.0 + 0.6 * std::pow(Re, 0.5) * pr_third(p, p.pr_air);
      dq_vol = Nu * ki / di * a_ht * T_m;
    }
    // the ISPH kernel divides the fluid drag by eps (:1099-1101), the WCSPH one does not (:338-340)
    const double dscale = p.solver_type == 0 ? 1.0 : pori;
    o.ftotalx = fx0 + dfx / dscale + tx;
    o.ftotaly = fy0 + dfy / dscale + ty;
    o.ftotalz = fz0 + dfz / dscale + tz;
    qi.Fdx_da = dax;
    qi.Fdy_da = day;
    qi.Fdz_da = daz;
    qi.Fdx_b = tx;
    qi.Fdy_b = ty;
    qi.Fdz_b = tz;
    if (p.con_solve == 1) {
      const double conv = dq_vol / (rhoi * cpi);  // [K/s]
      o.dtemp = (p.reference_compat & SOPHIA_COMPAT_HEAT_OVERWRITE) ? conv : o.dtemp + conv;
      qi.Q_sd = dq_vol * mi / rhoi;  // [W]
    }
  }
}

static void sph_coupling(orc_state& s) {
  Species& sp = s.sph;
  Species& dm = s.dem;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    part1& qi = sp.SP1[i];
    if (qi.i_type != 1 || qi.p_type > 1000) continue;
    double xi = qi.x, yi = qi.y, zi = qi.z, rhoi = qi.rho;
    int t = qi.p_type;
    part3& o = sp.P3[i];
    double fx0 = o.ftotalx, fy0 = o.ftotaly, fz0 = o.ftotalz, dti = o.dtemp;
    // heat_capacity(ptypei, tempi) at :1860 swaps its arguments: the
    // temperature is read as the particle type
    double cpi = (p.reference_compat & SOPHIA_COMPAT_CP_ARGS)
                     ? (std::abs((long)(unsigned)qi.temp) > 1000 ? p.cp_dem : p.cp_fluid)
                     : heat_capacity(p, t);
    double hi = qi.h, A = knl_tmpA(kt, dim, hi), R = p.kappa * hi;
    double tx = 0, ty = 0, tz = 0, tt = 0;
    if (dm.n() > 0) {
      for_candidates(s, dm, xi, yi, zi, [&](int j) {
        const part1& qj = dm.SP1[j];
        if (!(qj.p_type > 1000)) return;
        double r = dist_of(xi - qj.x, yi - qj.y, zi - qj.z);
        if (!(r < R)) return;
        if (t == SOPHIA_PT_MOVING || t == -3 || t == 0) return;
        double w = knl_w(kt, A, hi, r);
        double flt = qj.flt_sd_2 + qj.flt_s;
        if ((t == 1 || t == 3) && flt > 0.0) {
          // the WCSPH reaction carries the fluid particle's porosity
          // (KERNEL_SPH_coupling3D, :1677-1679); the ISPH one does not
          const double mr = p.solver_type == 0 ? qj.m * qi.DEMpor / rhoi : qj.m / rhoi;
          tx += -mr * qj.Fdx_df * w / (flt + 1.0e-20);
          ty += -mr * qj.Fdy_df * w / (flt + 1.0e-20);
          tz += -mr * qj.Fdz_df * w / (flt + 1.0e-20);
          tt += -qj.Q_sd / (rhoi * cpi) * w / (flt + 1.0e-20);
        }
      });
    }
    o.ftotalx = fx0 + tx;
    o.ftotaly = fy0 + ty;
    o.ftotalz = fz0 + tz;
    o.ftotal = std::sqrt((fx0 + tx) * (fx0 + tx) + (fy0 + ty) * (fy0 + ty) + (fz0 + tz) * (fz0 + tz));
    qi.test4 = o.ftotal;
    qi.test7 = tz;
    if (p.con_solve == 1) o.dtemp = dti + tt;
  }
}

void orc_sph_dem_coupling(orc_state* s) {
  if (s->p.dim != 3) return;
  dem_coupling(*s);
  int stride = std::max(1, s->p.decouple_stride);
  if ((s->count % stride) == 0) sph_coupling(*s);
}

/* ------------------------------------------------------------------------ */
/* TIME: KERNEL_clc_projection_sph/_dem (functions_TIME_ISPH2.cuh:104-308)  */
/* and KERNEL_time_update_projection_sph/_dem (:467-776)                    */
/* ------------------------------------------------------------------------ */
static void predictor(orc_state& s, Species& sp, double dt) {
  const sophia_params& p = s.p;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    part1& q = sp.SP1[i];
    part2& q2 = sp.SP2[i];
    const part3& o = sp.P3[i];
    if (q.i_type > kITypeCrt) continue;
    int t = q.p_type;
    double x0 = q.x, y0 = q.y, z0 = q.z, ux0 = q.ux, uy0 = q.uy, uz0 = q.uz;
    double xp, yp, zp, uxp, uyp, uzp;
    if (t >= 0) {
      double bz = (q.buffer_type == 0);
      uxp = ux0 + o.ftotalx * bz * dt;
      uyp = uy0 + o.ftotaly * bz * dt;
      uzp = uz0 + o.ftotalz * bz * dt;
      xp = x0 + uxp * dt * (t > 0) * q.elix;
      yp = y0 + uyp * dt * (t > 0) * q.eliy;
      zp = z0 + uzp * dt * (t > 0) * q.eliz;
    } else {
      xp = x0; yp = y0; zp = z0;
      uxp = q.ux; uyp = q.uy; uzp = q.uz;
    }
    q.x_star = xp; q.y_star = yp; q.z_star = zp;
    q.ux = uxp; q.uy = uyp; q.uz = uzp;
    q2.x0 = x0; q2.y0 = y0; q2.z0 = z0;
    q2.ux0 = ux0; q2.uy0 = uy0; q2.uz0 = uz0;
    if (s.time == 0) q2.rho_ref = q.rho;
    if (p.con_solve == 1 && t > 0) q2.temp0 = q.temp;
  }
}

static void corrector(orc_state& s, Species& sp, double dt, bool is_sph) {
  const sophia_params& p = s.p;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& q = sp.SP1[i];
    part1& T = sp.P1[i];
    const part2& q2 = sp.SP2[i];
    part2& T2 = sp.P2[i];
    const part3& o = sp.P3[i];
    if (q.i_type > kITypeCrt) continue;
    int t = q.p_type;
    double ux0, uy0, uz0, ax, ay, az;
    if (t > 0) {
      ux0 = q.ux; uy0 = q.uy; uz0 = q.uz;
      double bz = (q.buffer_type == 0);
      ax = o.fpx * bz + q.fbx / q.rho;
      ay = o.fpy * bz + q.fby / q.rho;
      az = o.fpz * bz + q.fbz / q.rho;
    } else {
      ux0 = q2.ux0; uy0 = q2.uy0; uz0 = q2.uz0;
      ax = ay = az = 0.0;
    }
    double uxc = ux0 + ax * dt, uyc = uy0 + ay * dt, uzc = uz0 + az * dt;
    if ((uxc * uxc + uyc * uyc + uzc * uzc) >= p.u_limit * p.u_limit) {
      uxc = ux0; uyc = uy0; uzc = uz0;
    }
    double xc = q.x + (q2.ux0 + uxc) / 2.0 * dt * (t > 0) * q.elix;
    double yc = q.y + (q2.uy0 + uyc) / 2.0 * dt * (t > 0) * q.eliy;
    double zc = q.z + (q2.uz0 + uzc) / 2.0 * dt * (t > 0) * q.eliz;
    if (!p.xsph_solve || t > 1000) {
      T.x = wrap_pos(s, xc, 0); T.y = wrap_pos(s, yc, 1); T.z = wrap_pos(s, zc, 2);
    }
    T.ux = uxc; T.uy = uyc; T.uz = uzc;
    if (!is_sph && t > 1000 && !(p.reference_compat & SOPHIA_COMPAT_NO_SPIN)) {
      // spin from the contact torque per inertia, as the WCSPH corrector
      // does (functions_TIME.cuh:753-755); the ISPH path leaves it out
      const double bz = (q.buffer_type == 0);
      T.wx = q.wx + o.torqx * bz * dt;
      T.wy = q.wy + o.torqy * bz * dt;
      T.wz = q.wz + o.torqz * bz * dt;
    }
    if (p.con_solve == 1 && t > 0) {
      T.temp = q2.temp0 + o.dtemp * dt * (t != -1);
    } else {
      T.enthalpy = q.enthalpy;
      T.temp = q.temp;
    }
    T.pres = q.pres; T.flt_s = q.flt_s; T.m = q.m; T.h = q.h;
    T.grad_rhox = q.grad_rhox; T.grad_rhoy = q.grad_rhoy; T.grad_rhoz = q.grad_rhoz;
    T.k_turb = q.k_turb; T.e_turb = q.e_turb;
    T.elix = q.elix; T.eliy = q.eliy; T.eliz = q.eliz;
    T.vol = q.vol; T.vol0 = q.vol0;
    T2.rho_ref = q2.rho_ref;
    T.flt_sd = q.flt_sd; T.flt_sd_2 = q.flt_sd_2; T.Q_sd = q.Q_sd;
    T.fcx = q.fcx; T.fcy = q.fcy; T.fcz = q.fcz;
    T.Fdx_b = q.Fdx_b; T.Fdy_b = q.Fdy_b; T.Fdz_b = q.Fdz_b;
    T.Fdx_da = q.Fdx_da; T.Fdy_da = q.Fdy_da; T.Fdz_da = q.Fdz_da;
    T.Fdx_df = q.Fdx_df; T.Fdy_df = q.Fdy_df; T.Fdz_df = q.Fdz_df;
    T.pgf_x = q.pgf_x; T.pgf_y = q.pgf_y; T.pgf_z = q.pgf_z;
    T.DEMpor = q.DEMpor;
    if (is_sph) {
      T.dDEMpor = q.dDEMpor;
      T.dDEMpor_prev = q.dDEMpor_prev;
    }
    T.DEMvf = q.DEMvf;
    T.PPE1 = q.PPE1; T.PPE2 = q.PPE2;
    T.test1 = q.test1; T.test2 = q.test2; T.test3 = q.test3; T.test4 = q.test4;
    T.test5 = q.test5; T.test6 = q.test6; T.test7 = q.test7; T.test8 = q.test8;
    T.rho = q.rho;
  }
}

/* ------------------------------------------------------------------------ */
/* DEM walls: KERNEL_treat_DEM_cylinder_z_dem (functions_DEM_BC.cuh:491-729),*/
/* KERNEL_treat_DEM_box_z_dem (:1770-2027), KERNEL_treat_DEM_truncated_      */
/* cone_z_dem (:4485-4725), called in this order after the DEM time update   */
/* (ISPH_Calc_seperate.cuh:439-441).  Geometry from sophia_params.wall_*.    */
/* ------------------------------------------------------------------------ */
namespace {
struct V3 {
  double x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double k, V3 a) { return {k * a.x, k * a.y, k * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double zero_small(double v) { return std::fabs(v) < 1e-10 ? 0.0 : v; }

// the tail shared by the three kernels (e.g. :609-725); n points inwards
void wall_contact(const sophia_params& p, double dt, V3 n, part1& T, part3& o) {
  const double rad = T.rad, m = T.m, ri = T.ri;
  V3 u{T.ux, T.uy, T.uz}, w{T.wx, T.wy, T.wz};
  const double un_mag = dot(u, n);
  V3 uc_w = (-rad) * cross(w, n);   // spin part of the contact velocity
  V3 uc = u + uc_w;
  const double ucn = dot(uc, n), ucwn = dot(uc_w, n);
  V3 ucs = uc - ucn * n, ucws = uc_w - ucwn * n;
  ucs = {zero_small(ucs.x), zero_small(ucs.y), zero_small(ucs.z)};
  ucws = {zero_small(ucws.x), zero_small(ucws.y), zero_small(ucws.z)};
  const double ucs_mag = std::sqrt(dot(ucs, ucs)), ucws_mag = std::sqrt(dot(ucws, ucws));
  const V3 s{ucs.x / (ucs_mag + 1e-20), ucs.y / (ucs_mag + 1e-20), ucs.z / (ucs_mag + 1e-20)};
  const V3 sw{ucws.x / (ucws_mag + 1e-20), ucws.y / (ucws_mag + 1e-20), ucws.z / (ucws_mag + 1e-20)};
  if (std::fabs(un_mag) < 50 * dt) {  // rolling on the surface
    const double fn = o.ftotalx * n.x + o.ftotaly * n.y + o.ftotalz * n.z;
    V3 ff = (p.wall_mu[0] * m * fn) * s, frf = (p.wall_mu[1] * m * fn) * sw;
    if (fn > 0.0) ff = frf = {0.0, 0.0, 0.0};
    o.ftotalx += ff.x / m; o.ftotaly += ff.y / m; o.ftotalz += ff.z / m;
    const V3 tq = (-(rad / ri)) * cross(n, ff + frf);
    o.torqx += tq.x; o.torqy += tq.y; o.torqz += tq.z;
    const V3 u_c{u.x + (ff.x / m) * dt, u.y + (ff.y / m) * dt, u.z + (ff.z / m) * dt};
    const V3 un = dot(u_c, n) * n;
    T.ux = u_c.x - un.x; T.uy = u_c.y - un.y; T.uz = u_c.z - un.z;
    T.wx = w.x + tq.x * dt; T.wy = w.y + tq.y * dt; T.wz = w.z + tq.z * dt;
  } else {  // impulse
    const V3 Jn = (-m * (1 + p.wall_rest[0]) * ucn) * n;
    const V3 Js = (-1 / ((1 / m) + (rad * rad / ri)) * (1 + p.wall_rest[1]) * ucs_mag) * s;
    const V3 J = Jn + Js;
    T.ux = u.x + (1 / m) * J.x; T.uy = u.y + (1 / m) * J.y; T.uz = u.z + (1 / m) * J.z;
    const V3 dw = (rad / ri) * cross(n, J);
    T.wx = w.x - dw.x; T.wy = w.y - dw.y; T.wz = w.z - dw.z;
  }
}

void wall_cylinder_z(const sophia_params& p, double dt, part1& T, part3& o) {
  const double xc = p.wall_cyl[0], yc = p.wall_cyl[1], R = p.wall_cyl[2], rad = T.rad;
  const double F = (T.x - xc) * (T.x - xc) + (T.y - yc) * (T.y - yc) - (R - rad) * (R - rad);
  if (F <= 0) return;
  const double dist = std::sqrt((T.x - xc) * (T.x - xc) + (T.y - yc) * (T.y - yc));
  const double cpx = xc + (R - rad) * (T.x - xc) / dist;
  const double cpy = yc + (R - rad) * (T.y - yc) / dist;
  T.x = cpx;
  T.y = cpy;
  const V3 n{-(cpx - xc) / (R - rad), -(cpy - yc) / (R - rad), 0.0};
  if (std::sqrt(dot(n, n)) < 0.5) return;
  wall_contact(p, dt, n, T, o);
}

void wall_box_z(const sophia_params& p, double dt, part1& T, part3& o) {
  const double* b = p.wall_box;
  const double rad = T.rad, x = T.x, y = T.y, z = T.z;
  const double ax = 0.5 * b[3] - rad, ay = 0.5 * b[4] - rad, az = 0.5 * b[5] - rad;
  const double F = std::fmax(std::fmax(std::fabs(x - b[0]) - ax, std::fabs(y - b[1]) - ay), std::fabs(z - b[2]) - az);
  if (F <= 0) return;
  const double cpx = b[0] + std::fmin(ax, std::fmax(-ax, x - b[0]));
  const double cpy = b[1] + std::fmin(ay, std::fmax(-ay, y - b[1]));
  const double cpz = b[2] + std::fmin(az, std::fmax(-az, z - b[2]));
  T.x = cpx; T.y = cpy; T.z = cpz;
  double sgx = (cpx - x) / (std::fabs(cpx - x) + 1e-20);
  double sgy = (cpy - y) / (std::fabs(cpy - y) + 1e-20);
  double sgz = (cpz - z) / (std::fabs(cpz - z) + 1e-20);
  if (std::fabs(cpy - y) < 1e-12) sgy = 0.0;
  if (std::fabs(cpy - x) < 1e-12) sgx = 0.0;  // cp_y against x, as at :1877
  if (std::fabs(cpz - z) < 1e-12) sgz = 0.0;
  const double mag = std::sqrt(sgx * sgx + sgy * sgy + sgz * sgz) + 1e-20;
  const V3 n{sgx / mag, sgy / mag, sgz / mag};
  if (std::sqrt(dot(n, n)) < 0.5) return;
  wall_contact(p, dt, n, T, o);
}

void wall_truncated_cone_z(const sophia_params& p, double dt, part1& T, part3& o) {
  const double xc = p.wall_cone[0], yc = p.wall_cone[1], r1 = p.wall_cone[2], r2 = p.wall_cone[3], h = p.wall_cone[4];
  const double rad = T.rad, x = T.x, y = T.y, z = T.z;
  if (z <= -h + rad) return;
  const double sin_t = h / std::sqrt((r1 - r2) * (r1 - r2) + h * h);
  const double cos_t = std::fabs(r1 - r2) / std::sqrt((r1 - r2) * (r1 - r2) + h * h);
  const double tan_t = sin_t / cos_t;
  const double rw = r2 + (h + z) / tan_t - rad / sin_t;
  const double rxy = std::sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc));
  if ((x - xc) * (x - xc) + (y - yc) * (y - yc) - rw * rw <= 0) return;
  T.z = tan_t * (rxy + rad / sin_t - r2) - h;
  const V3 n{-sin_t * (x - xc) / rxy, -sin_t * (y - yc) / rxy, cos_t};
  wall_contact(p, dt, n, T, o);
}
}  // namespace

static void dem_walls(orc_state& s) {
  Species& sp = s.dem;
  for (int i = 0; i < sp.n(); i++) {
    if (sp.SP1[i].i_type != 1 || sp.SP1[i].p_type <= 1000) continue;
    if (s.p.dem_walls & 1) wall_cylinder_z(s.p, s.p.dt, sp.P1[i], sp.P3[i]);
    if (s.p.dem_walls & 2) wall_box_z(s.p, s.p.dt, sp.P1[i], sp.P3[i]);
    if (s.p.dem_walls & 4) wall_truncated_cone_z(s.p, s.p.dt, sp.P1[i], sp.P3[i]);
  }
}

/* ------------------------------------------------------------------------ */
/* Open boundary: KERNEL_open_boundary_extrapolation3D_1_sph (functions_BC. */
/* cuh:787-984) then _calc_sph (:1204-1217), every decouple_stride steps    */
/* after the time update (ISPH_Calc_seperate.cuh:410-426)                   */
/* ------------------------------------------------------------------------ */
static void open_boundary(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int n = sp.n();
  struct OBC { double pres, rho, m, ux, uy, uz, temp; bool set; };
  std::vector<OBC> out(n, OBC{0, 0, 0, 0, 0, 0, 0, false});
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const part1& q = sp.P1[i];
    if (q.buffer_type <= 0 || q.p_type >= 1000) continue;
    const double h = q.h, A = knl_tmpA(p.kernel_type, p.dim, h), range = p.kappa * h;
    // ghost node mirrored about the inlet (L2) or outlet (L3) plane
    const double xg = q.x, yg = q.y;
    double zg = q.z;
    if (q.buffer_type == 1) zg = 2 * p.ob_band[1] - q.z;
    else if (q.buffer_type == 2) zg = 2 * p.ob_band[2] - q.z;
    double trho = 0, tpres = 0, tuz = 0, tux = 0, tuy = 0, ttemp = 0, tflt = 0;
    for_candidates(s, sp, xg, yg, zg, [&](int j) {
      const part1& r = sp.P1[j];
      if (!(r.p_type < 1000)) return;
      const double d = dist_of(xg - r.x, yg - r.y, zg - r.z) + 1e-20;
      if (!(d < range)) return;
      if (!(r.p_type == 1 && r.buffer_type == 0)) return;
      const double w = knl_w(p.kernel_type, A, h, d);
      trho += r.rho / r.DEMpor * w * r.vol;
      tpres += r.vol * r.pres * w;
      tuz += r.vol * r.uz * w;
      tux += r.vol * r.ux * w;
      tuy += r.vol * r.uy * w;
      ttemp += r.vol * r.temp * w;
      tflt += r.vol * w;
    });
    OBC o;
    o.set = true;
    o.m = q.rho * q.vol0;
    if (q.elix < 0.9999 && q.p_type > 0) o.m = q.rho * q.vol;
    if (q.buffer_type == 1) {
      o.pres = tpres / tflt;
      o.rho = q.DEMpor * p.rho_fluid;
      o.uz = p.ob_inlet_u * 2 - tuz / tflt;
      o.ux = 0 - tux / tflt;
      o.uy = 0 - tuy / tflt;
      o.temp = p.ob_inlet_temp;
    } else if (q.buffer_type == 2) {
      o.pres = 0.0;
      o.rho = q.DEMpor * (trho / tflt);
      o.uz = tuz / tflt;
      o.ux = tux / tflt;
      o.uy = tuy / tflt;
      o.temp = ttemp / tflt;
    } else {
      o.set = false;
    }
    out[i] = o;
  }
  for (int i = 0; i < n; i++) {
    if (!out[i].set) continue;
    part1& q = sp.P1[i];
    q.pres = out[i].pres; q.rho = out[i].rho; q.m = out[i].m;
    q.uz = out[i].uz; q.ux = out[i].ux; q.uy = out[i].uy; q.temp = out[i].temp;
  }
}

void orc_time(orc_state* s, int32_t phase) {
  int stride = std::max(1, s->p.decouple_stride);
  bool sph_on = (s->count % stride) == 0;
  if (phase == SOPHIA_PREDICT) {
    if (s->p.time_type != 1) return;
    if (sph_on) predictor(*s, s->sph, stride * s->p.dt);
    predictor(*s, s->dem, s->p.dt);
  } else {
    if (sph_on) corrector(*s, s->sph, stride * s->p.dt, true);
    corrector(*s, s->dem, s->p.dt, false);
    if (sph_on && s->p.open_boundary == 1) open_boundary(*s);
    if (s->p.dem_walls) dem_walls(*s);
    s->after_corrector = true;
  }
}

/* ------------------------------------------------------------------------ */
/* PPE: KERNEL_Neumann_boundary3D_sph (functions_BC.cuh:139-274),           */
/* KERNEL_PPE3D_sph (functions_PPE2.cuh:542-698), KERNEL_PPE3D_calc_sph     */
/* (:868-878), KERNEL_pressureforce3D_sph (:2154-2335); Krylov solve new    */
/* ------------------------------------------------------------------------ */
static void neumann_walls(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
  int n = sp.n();
  std::vector<double> np(n), nt(n), nr(n);
  std::vector<char> act(n, 0), zero(n, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const part1& qi = sp.SP1[i];
    if (qi.p_type > 0) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, 1.0 * h), R = p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z;
    double tp = 0, fl = 0, tt = 0, tr = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      double r = dist_of(pd(s, xi, qj.x, 0), pd(s, yi, qj.y, 1), pd(s, zi, qj.z, 2)) + 1e-20;
      if (!(qj.p_type < 1000)) return;
      if (!(r < 2.0 * R)) return;
      double w = knl_w(kt, A, 2.0 * h, r);
      double on = (qj.p_type >= 1);
      double v = qj.m / qj.rho;
      tp += v * qj.pres * w * on;
      fl += v * w * on;
      tt += v * qj.temp * w * on;
      tr += qj.rho / qj.DEMpor * w * v * on;
    });
    act[i] = 1;
    if (fl < 1e-6) {
      zero[i] = 1;
    } else {
      np[i] = tp / fl;
      nt[i] = tt / fl;
      nr[i] = qi.DEMpor * tr / fl;
    }
  }
  for (int i = 0; i < n; i++) {
    if (!act[i]) continue;
    part1& q = sp.SP1[i];
    if (zero[i]) {
      q.rho = sp.SP2[i].rho_ref;
      q.pres = 0.0;
    } else {
      q.pres = np[i];
      q.temp = nt[i];
      q.rho = nr[i];
    }
  }
}

/* fluid particle that takes a pressure force (KERNEL_pressureforce3D_sph's
 * filter, functions_PPE2.cuh:2154-2335) and, unless it is a free-surface
 * particle, a PPE row (KERNEL_PPE3D_sph, :542-698) */
static bool ppe_fluid(const orc_state& s, const part1& q) {
  if (q.i_type > kITypeCrt) return false;
  if (s.p.open_boundary > 0 && q.buffer_type > 0) return false;
  if (q.p_type >= 1000 || q.p_type <= 0) return false;
  return true;
}

static bool ppe_row(const orc_state& s, int i) {
  const part1& q = s.sph.SP1[i];
  if (!ppe_fluid(s, q)) return false;
  return !(s.p.free_surface > 0 && s.sph.P3[i].lbl_surf != 0.0);
}

/* KERNEL_clc_surface_detect3D (functions_PST.cuh:263-320), run before the
 * PPE as at ISPH_Calc_seperate.cuh:289-296: the mean unit vector from the
 * neighbours within kappa*h, |sum (x_i-x_j)/(r+1e-10)| / (non+1e-10) >
 * fs_threshold, or non < 10, marks a surface particle.  free_surface = 2
 * counts neighbours of the particle's own p_type as the reference does;
 * free_surface = 1 counts every SPH particle (walls included), so fluid next
 * to a wall is not taken for surface.  Surface particles take the Dirichlet
 * value p_ref (the commented-out rows of functions_PPE2.cuh:533-534). */
static void surface_detect(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int n = sp.n();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const part1& qi = sp.SP1[i];
    sp.P3[i].lbl_surf = 0.0;
    if (!ppe_fluid(s, qi)) continue;
    const double R = p.kappa * qi.h;
    double tx = 0, ty = 0, tz = 0;
    int non = 0;
    for_candidates(s, sp, qi.x, qi.y, qi.z, [&](int j) {
      const part1& qj = sp.SP1[j];
      double dx = pd(s, qi.x, qj.x, 0), dy = pd(s, qi.y, qj.y, 1), dz = pd(s, qi.z, qj.z, 2);
      double r = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (!(r < R)) return;
      bool same = p.free_surface == 2 ? qj.p_type == qi.p_type : qj.p_type < 1000;
      if (!same) return;
      double rd = 1.0 / (r + 1e-10);
      tx += dx * rd;
      ty += dy * rd;
      tz += dz * rd;
      non++;
    });
    double rn = 1.0 / (non + 1e-10);
    tx *= rn;
    ty *= rn;
    tz *= rn;
    bool surf = std::sqrt(tx * tx + ty * ty + tz * tz) > p.fs_threshold || non < 10;
    sp.P3[i].lbl_surf = surf ? 1.0 : 0.0;
  }
  for (int i = 0; i < n; i++)
    if (sp.P3[i].lbl_surf != 0.0) sp.SP1[i].pres = p.p_ref;
}

static void ppe_assemble(orc_state& s, double tdt) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
  int n = sp.n();
  s.row_mask.assign(n, 0);
  s.rows.assign(n, PPERow{});
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    part1& qi = sp.SP1[i];
    if (!ppe_row(s, i)) continue;
    s.row_mask[i] = 1;
    PPERow& row = s.rows[i];
    row.i = i;
    double hi = qi.h, A = knl_tmpA(kt, dim, hi), R = p.kappa * hi;
    double xi = qi.x, yi = qi.y, zi = qi.z, uxi = qi.ux, uyi = qi.uy, uzi = qi.uz;
    double rhoi = qi.rho, pori = qi.DEMpor;
    double bx = 0, by = 0, bz = 0, Aij = 0, AP = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (!(qj.p_type < 1000)) return;
      if (qj.i_type == 4) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1), dz = pd(s, zi, qj.z, 2);
      double r = dist_of(dx, dy, dz) + 1e-20;
      if (!(r < R)) return;
      double dw = knl_dw(kt, A, hi, r);
      double gx = dw * dx / r, gy = dw * dy / r, gz = dw * dz / r;
      kgc_apply(s, sp.P3[i].Cm, gx, gy, gz);
      double mj = qj.m, rhoj = qj.rho, porj = qj.DEMpor;
      bx += mj / rhoj * (pori * uxi - porj * qj.ux) * gx;
      by += mj / rhoj * (pori * uyi - porj * qj.uy) * gy;
      bz += mj / rhoj * (pori * uzi - porj * qj.uz) * gz;
      double c = -(dx * gx + dy * gy + dz * gz) / r / r;
      double a = 1 * mj * ((pori * rhoi + porj * rhoj) / (rhoj * rhoi)) * c;
      Aij += a;
      AP += 1 * mj * ((pori * rhoi + porj * rhoj) / (rhoj * rhoi)) * c * qj.pres;
      row.j.push_back(j);
      row.a.push_back(a);
    });
    double bi = p.rho_fluid * (bx + by + bz + 2 * qi.dDEMpor - qi.dDEMpor_prev) / tdt;
    qi.PPE3 = bi;
    qi.PPE4 = AP;
    qi.PPE1 = rhoi * (bx + by + bz) / tdt;
    qi.PPE2 = Aij;
  }
  // Krylov right-hand side: b_i plus the lagged non-row columns
  s.rhs.assign(n, 0.0);
  for (int i = 0; i < n; i++) {
    if (!s.row_mask[i]) continue;
    double r = sp.SP1[i].PPE3;
    const PPERow& row = s.rows[i];
    for (size_t k = 0; k < row.j.size(); k++)
      if (!s.row_mask[row.j[k]]) r += row.a[k] * sp.SP1[row.j[k]].pres;
    s.rhs[i] = r;
  }
}

/* y = A x over rows: (A x)_i = d_i x_i - sum_{j in rows} a_ij x_j */
static void ppe_matvec(const orc_state& s, const std::vector<double>& x, std::vector<double>& y) {
  int n = (int)x.size();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    if (!s.row_mask[i]) {
      y[i] = 0.0;
      continue;
    }
    const PPERow& row = s.rows[i];
    double acc = s.sph.SP1[i].PPE2 * x[i];
    for (size_t k = 0; k < row.j.size(); k++)
      if (s.row_mask[row.j[k]]) acc -= row.a[k] * x[row.j[k]];
    y[i] = acc;
  }
}

static double dotv(const std::vector<double>& a, const std::vector<double>& b) {
  double acc = 0.0;
  for (size_t i = 0; i < a.size(); i++) acc += a[i] * b[i];
  return acc;
}

static void project_mean(const orc_state& s, std::vector<double>& v) {
  double sum = 0;
  int cnt = 0;
  for (size_t i = 0; i < v.size(); i++)
    if (s.row_mask[i]) { sum += v[i]; cnt++; }
  if (!cnt) return;
  double m = sum / cnt;
  for (size_t i = 0; i < v.size(); i++)
    if (s.row_mask[i]) v[i] -= m;
}

/* BiCGStab (van der Vorst 1992) with right Jacobi preconditioning on the
 * fluid rows; walls and other non-rows enter through the right-hand side */
static int ppe_bicgstab(orc_state& s, const sophia_ppe_opts& o, sophia_ppe_stats& st) {
  Species& sp = s.sph;
  int n = sp.n();
  std::vector<double> x(n, 0.0), r(n), rh(n), v(n, 0.0), pv(n, 0.0), ph(n), sv(n), sh(n), t(n), Ax(n);
  for (int i = 0; i < n; i++)
    if (s.row_mask[i]) x[i] = (o.guess == SOPHIA_GUESS_ZERO) ? 0.0 : sp.SP1[i].pres;
  std::vector<double> b = s.rhs;
  if (o.zero_mean) project_mean(s, b);
  double bn = std::sqrt(dotv(b, b));
  if (bn == 0) bn = 1;
  ppe_matvec(s, x, Ax);
  for (int i = 0; i < n; i++) r[i] = s.row_mask[i] ? b[i] - Ax[i] : 0.0;
  rh = r;
  st.r0_rel = std::sqrt(dotv(r, r)) / bn;
  st.guess_used = o.guess == SOPHIA_GUESS_ZERO ? SOPHIA_GUESS_ZERO : SOPHIA_GUESS_PREVIOUS;
  double rho = 1, alpha = 1, omega = 1;
  int it = 0;
  double rn = st.r0_rel * bn;
  st.converged = rn <= o.rtol * bn;
  auto dinv = [&](int i) { return s.row_mask[i] && sp.SP1[i].PPE2 != 0 ? 1.0 / sp.SP1[i].PPE2 : 0.0; };
  while (!st.converged && it < o.max_iter) {
    double rho1 = dotv(rh, r);
    if (rho1 == 0) break;
    double beta = (rho1 / rho) * (alpha / omega);
    for (int i = 0; i < n; i++) pv[i] = r[i] + beta * (pv[i] - omega * v[i]);
    for (int i = 0; i < n; i++) ph[i] = pv[i] * dinv(i);
    ppe_matvec(s, ph, v);
    alpha = rho1 / dotv(rh, v);
    for (int i = 0; i < n; i++) sv[i] = r[i] - alpha * v[i];
    it++;
    double sn = std::sqrt(dotv(sv, sv));
    if (sn <= o.rtol * bn) {
      for (int i = 0; i < n; i++) x[i] += alpha * ph[i];
      r = sv;
      rn = sn;
      st.converged = 1;
      break;
    }
    for (int i = 0; i < n; i++) sh[i] = sv[i] * dinv(i);
    ppe_matvec(s, sh, t);
    double tt = dotv(t, t);
    omega = tt > 0 ? dotv(t, sv) / tt : 0.0;
    for (int i = 0; i < n; i++) x[i] += alpha * ph[i] + omega * sh[i];
    for (int i = 0; i < n; i++) r[i] = sv[i] - omega * t[i];
    rho = rho1;
    rn = std::sqrt(dotv(r, r));
    if (rn <= o.rtol * bn) st.converged = 1;
    if (omega == 0) break;
  }
  if (o.zero_mean) project_mean(s, x);
  // true residual
  ppe_matvec(s, x, Ax);
  double tr = 0;
  for (int i = 0; i < n; i++)
    if (s.row_mask[i]) tr += (b[i] - Ax[i]) * (b[i] - Ax[i]);
  st.r_rel = std::sqrt(tr) / bn;
  st.iterations = it;
  st.outer = 1;
  for (int i = 0; i < n; i++)
    if (s.row_mask[i]) sp.SP1[i].pres = x[i];
  return 0;
}

static void pressure_force(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& qi = sp.SP1[i];
    if (!ppe_fluid(s, qi)) continue;
    double hi = qi.h, A = knl_tmpA(kt, dim, hi), R = p.kappa * hi;
    double xi = qi.x, yi = qi.y, zi = qi.z, pi = qi.pres, rhoi = qi.rho, pori = qi.DEMpor;
    double tx = 0, ty = 0, tz = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (!(qj.p_type < 1000)) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1), dz = pd(s, zi, qj.z, 2);
      double r = dist_of(dx, dy, dz) + 1e-20;
      if (!(r < R)) return;
      double dw = knl_dw(kt, A, hi, r);
      double gx = dw * dx / r, gy = dw * dy / r, gz = dw * dz / r;
      kgc_apply(s, sp.P3[i].Cm, gx, gy, gz);
      if (p.fp_solve) {
        double Cp = -pori * 1 * qj.m * (p.pgrad_form ? qj.pres - pi : pi + qj.pres) / (rhoi * qj.rho);
        tx += Cp * gx;
        ty += Cp * gy;
        tz += Cp * gz;
      }
    });
    part3& o = sp.P3[i];
    o.fpx = tx;
    o.fpy = ty;
    o.fpz = tz;
  }
}

int orc_ppe(orc_state* s, const sophia_ppe_opts* o, sophia_ppe_stats* st) {
  sophia_ppe_stats local{};
  if (!st) st = &local;
  std::memset(st, 0, sizeof(*st));
  int stride = std::max(1, s->p.decouple_stride);
  if ((s->count % stride) != 0 || s->p.dim != 3) return 0;
  sophia_ppe_opts def{};
  def.method = SOPHIA_PPE_JACOBI;
  if (!o) o = &def;
  if (s->count == 0) neumann_walls(*s);
  if (s->p.free_surface > 0) surface_detect(*s);
  ppe_assemble(*s, stride * s->p.dt);
  int nrows = 0;
  for (int v : s->row_mask) nrows += v;
  st->n_rows = nrows;
  if (o->method == SOPHIA_PPE_JACOBI) {
    Species& sp = s->sph;
    for (int i = 0; i < sp.n(); i++)
      if (s->row_mask[i]) sp.SP1[i].pres = (sp.SP1[i].PPE3 + sp.SP1[i].PPE4) / sp.SP1[i].PPE2;
    st->iterations = 1;
    st->converged = 1;
  } else {
    ppe_bicgstab(*s, *o, *st);
  }
  neumann_walls(*s);
  pressure_force(*s);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* XSPH: KERNEL_xsph3D_sph + KERNEL_xsph3D_calc_sph (functions_TIME.cuh:    */
/* 1735-1850, 1985-1997), on the carried state P1                           */
/* ------------------------------------------------------------------------ */
static void xsph_sph(orc_state* s, bool apply);
void orc_xsph(orc_state* s) {
  int stride = std::max(1, s->p.decouple_stride);
  if (!s->p.xsph_solve || (s->count % stride) != 0 || s->p.dim != 3) return;
  xsph_sph(s, true);
}

/* KERNEL_xsph3D_sph computes XSPH_u/XSPH_temp; with apply the calc kernel
 * (KERNEL_xsph3D_calc_sph) copies them into u/temp.  The WCSPH driver calls
 * only KERNEL_xsph3D (WCSPH_Calc.cuh:313-314), which is the same text. */
static void xsph_sph(orc_state* s, bool apply) {
  int stride = std::max(1, s->p.decouple_stride);
  Species& sp = s->sph;
  const sophia_params& p = s->p;
  const int kt = p.kernel_type, dim = p.dim;
  double dt = stride * p.dt;
  int n = sp.n();
  std::vector<double> vx(n), vy(n), vz(n), vt(n);
  std::vector<char> act(n, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const part1& qi = sp.P1[i];
    if (qi.i_type != 1 || qi.p_type > 1000 || qi.p_type == 0 || qi.p_type == 9) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z, uxi = qi.ux, uyi = qi.uy, uzi = qi.uz, ti = qi.temp;
    double tx = 0, ty = 0, tz = 0, tt = 0, fl = 0;
    for_candidates(*s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.P1[j];
      if (!(qj.p_type == 1 || qj.p_type == 3)) return;
      double r = dist_of(pd(*s, xi, qj.x, 0), pd(*s, yi, qj.y, 1), pd(*s, zi, qj.z, 2));
      if (!(r < R)) return;
      double w = knl_w(kt, A, h, r);
      double v = qj.m / qj.rho;
      tx += p.c_xsph * v * (-uxi + qj.ux) * w;
      ty += p.c_xsph * v * (-uyi + qj.uy) * w;
      tz += p.c_xsph * v * (-uzi + qj.uz) * w;
      tt += 0.01 * p.c_xsph * v * (qj.temp - ti) * w;
      fl += v * w;
    });
    double on = (qi.p_type > 0);
    vx[i] = (uxi + tx / fl) * on;
    vy[i] = (uyi + ty / fl) * on;
    vz[i] = (uzi + tz / fl) * on;
    vt[i] = (ti + tt / fl) * on;
    act[i] = 1;
  }
  for (int i = 0; i < n; i++) {
    if (!act[i]) continue;
    part1& q = sp.P1[i];
    q.XSPH_ux = vx[i]; q.XSPH_uy = vy[i]; q.XSPH_uz = vz[i]; q.XSPH_temp = vt[i];
    if (!apply) continue;
    q.ux = vx[i]; q.uy = vy[i]; q.uz = vz[i]; q.temp = vt[i];
    if (p.fix_xsph_position && q.p_type > 0 && !p.eulerian_sph) {
      const part2& q2 = sp.SP2[i];
      q.x = wrap_pos(*s, q2.x0 + vx[i] * dt, 0);
      q.y = wrap_pos(*s, q2.y0 + vy[i] * dt, 1);
      q.z = wrap_pos(*s, q2.z0 + vz[i] * dt, 2);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* WCSPH step (SOPHIA_single, WCSPH_Calc.cuh:4-340): fluid only, 3D.        */
/* KERNEL_set_alpha makes every SPH particle Eulerian (functions_ALE.cuh:   */
/* 2-20), so NNPS runs at count 0 only (WCSPH_Calc.cuh:34); later steps     */
/* copy P1 -> SP1, P2 -> SP2.  P3 is reset after the predictor, which       */
/* therefore uses the previous step's ftotal, drho and dtemp.               */
/* ------------------------------------------------------------------------ */
constexpr int kMoving = 9;  // MOVING (Parameters.cuh:12)

/* KERNEL_set_alpha (functions_ALE.cuh:2-20) */
static void wc_set_alpha(Species& sp, std::vector<part1>& P) {
  for (int i = 0; i < sp.n(); i++) {
    part1& q = P[i];
    if (q.i_type == 3) continue;
    double e = q.p_type <= 1000 ? 0.0 : 1.0;
    q.elix = q.eliy = q.eliz = e;
  }
}

/* KERNEL_clc_predictor (functions_TIME.cuh:77-292) on P1/P2: half step of
 * x (Lagrangian only), u, rho (continuity) and T with the previous forces */
static void wc_predictor(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const double dt = p.dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    part1& q = sp.P1[i];
    part2& q2 = sp.P2[i];
    const part3& o = sp.P3[i];
    if (q.i_type == 3) continue;
    const int t = q.p_type;
    const double bz = (q.buffer_type == 0);
    if (t != kMoving) {
      double x0 = q.x, y0 = q.y, z0 = q.z;
      double ux0 = q.ux, uy0 = q.uy, uz0 = q.uz;
      double xp, yp, zp, uxp, uyp, uzp;
      if (t >= 0) {
        xp = x0 + ux0 * (dt * 0.5) * (t > 0) * q.elix;
        yp = y0 + uy0 * (dt * 0.5) * (t > 0) * q.eliy;
        zp = z0 + uz0 * (dt * 0.5) * (t > 0) * q.eliz;
        uxp = ux0 + o.ftotalx * bz * (dt * 0.5);
        uyp = uy0 + o.ftotaly * bz * (dt * 0.5);
        uzp = uz0 + o.ftotalz * bz * (dt * 0.5);
      } else {
        xp = x0; yp = y0; zp = z0;
        uxp = q.ux; uyp = q.uy; uzp = q.uz;
      }
      q.x = xp; q.y = yp; q.z = zp;
      q.ux = uxp; q.uy = uyp; q.uz = uzp;
      q2.x0 = x0; q2.y0 = y0; q2.z0 = z0;
      q2.ux0 = ux0; q2.uy0 = uy0; q2.uz0 = uz0;
    }
    if (s.time == 0) q2.rho_ref = q.rho;
    if (p.rho_type == 1 && t > 0) {
      double trho = q.rho;
      if (s.time < 1.0e-10) q.vol0 = q.m / trho;
      q2.rho0 = trho;
      q.rho = trho + o.drho * (dt * 0.5);
    } else {
      double trho = q.rho;
      double hv = q.h / p.h_over_dx;
      if (s.time < 1.0e-10) q.vol0 = hv * hv * hv;
      q2.rho0 = trho;
    }
    if (p.con_solve == 1) {
      q2.temp0 = q.temp;
      q.temp = q.temp + o.dtemp * (dt * 0.5);
    }
  }
}

/* KERNEL_clc_volume_update3D (functions_MASS.cuh:148-222): Shepard-corrected
 * volume over the +-3 cells; Eulerian particles only at count 0 */
static void wc_volume_update(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
  std::vector<double> vol(sp.n());
  std::vector<char> act(sp.n(), 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt) continue;
    if (qi.elix < 1e-8 && qi.eliy < 1e-8 && s.count > 0) continue;
    if (qi.p_type >= 1000) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = p.kappa * h;
    double tR = 0, filt = 0;
    for_candidates_r(s, sp, qi.x, qi.y, qi.z, 3, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (!(qj.p_type < 1000) || qj.i_type == 4) return;
      double r = dist_of(pd(s, qi.x, qj.x, 0), pd(s, qi.y, qj.y, 1), pd(s, qi.z, qj.z, 2)) + 1e-20;
      if (!(r < R)) return;
      double w = knl_w(kt, A, h, r);
      tR += w;
      filt += qj.m / qj.rho / qj.vol0 * w;
    });
    vol[i] = qi.vol0 * 1.0 / tR * filt;
    act[i] = 1;
  }
  for (int i = 0; i < sp.n(); i++)
    if (act[i]) sp.SP1[i].vol = vol[i];
}

/* KERNEL_clc_mass_update (functions_MASS.cuh:45-60): m = rho vol, Eulerian */
static void wc_mass_update(orc_state& s) {
  Species& sp = s.sph;
  for (int i = 0; i < sp.n(); i++) {
    part1& q = sp.SP1[i];
    if (q.p_type > 1000 || q.i_type > kITypeCrt) continue;
    if (q.elix == 1.0 && q.eliy == 1.0) continue;
    q.m = q.rho * q.vol;
  }
}

/* KERNEL_clc_continuity_norm3D (functions_MASS.cuh:771-914) -> P3.drho */
static void wc_continuity(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt || qi.p_type > 1000) continue;
    if (p.open_boundary > 0 && qi.buffer_type > 0) continue;
    if (qi.p_type <= 0) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z, uxi = qi.ux, uyi = qi.uy, uzi = qi.uz;
    double rhoi = qi.rho, pori = qi.DEMpor;
    double gxi = qi.grad_rhox, gyi = qi.grad_rhoy, gzi = qi.grad_rhoz;
    double tx = 0, ty = 0, tz = 0, tR = 0, px = 0, py = 0, pz = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (!(qj.p_type <= 1000)) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1), dz = pd(s, zi, qj.z, 2);
      double r = dist_of(dx, dy, dz) + 1e-20;
      if (!(r < R)) return;
      double dw = knl_dw(kt, A, h, r);
      double gx = dw * dx / r, gy = dw * dy / r, gz = dw * dz / r;
      if (p.kgc_solve > 0) kgc_apply(s, sp.P3[i].Cm, gx, gy, gz);
      double mj = qj.m, porj = qj.DEMpor, rhoj = qj.rho;
      double tmprho = mj * (rhoi / rhoj);
      tx += (uxi - qj.ux) * tmprho * gx;
      ty += (uyi - qj.uy) * tmprho * gy;
      tz += (uzi - qj.uz) * tmprho * gz;
      px += uxi * mj / rhoj * (rhoi / pori - rhoj / porj) * gx;
      py += uyi * mj / rhoj * (rhoi / pori - rhoj / porj) * gy;
      pz += uzi * mj / rhoj * (rhoi / pori - rhoj / porj) * gz;
      if (p.delSPH_solve > 0) {
        double rri = sp.SP2[i].rho_ref, rrj = sp.SP2[j].rho_ref;
        double phi = (gxi + qj.grad_rhox) * (-pd(s, xi, qj.x, 0));
        phi += (gyi + qj.grad_rhoy) * (-pd(s, yi, qj.y, 1));
        phi += (gzi + qj.grad_rhoz) * (-pd(s, zi, qj.z, 2));
        phi = (-0.5 * phi) * (p.delSPH_solve == 2) + (rhoj / rrj - rhoi / rri);
        double tr = -2 * p.delta_sph * h * rri * p.soundspeed * (mj / rhoj);
        tr *= phi * dw / r;
        tR += tr;
      }
    });
    sp.P3[i].drho = (tx + ty + tz) + pori * (px + py + pz) * (1 - qi.elix) + tR * (p.delSPH_solve > 0);
  }
}

/* KERNEL_EOS (functions_PROP.cuh:7-28): Tait, B = b c0^2 / gamma with the
 * reference's b = 1.186 unless eos_b_factor says otherwise */
static void wc_eos(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const double B = p.soundspeed * p.soundspeed * (p.eos_b_factor) / p.gamma;
  for (int i = 0; i < sp.n(); i++) {
    part1& q = sp.SP1[i];
    if (q.i_type > kITypeCrt || q.p_type > 1000 || q.buffer_type != 0) continue;
    q.pres = B * (std::pow(q.rho / sp.SP2[i].rho_ref, p.gamma) - 1.0);
  }
}

/* KERNEL_interaction3D (functions_INTERACTION.cuh:413-832), i_type 1:
 * pressure, laminar and artificial viscosity, gravity, conduction and the
 * Eulerian advection term (interface, repulsive, surface-tension,
 * Boussinesq and concentration terms off) */
static void wc_interaction(orc_state& s) {
  const sophia_table* tab = s.has_table ? &s.table : g_no_table;
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    part1& qi = sp.SP1[i];
    if (qi.i_type != 1 || qi.p_type > 1000) continue;
    const int pti = qi.p_type;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z, uxi = qi.ux, uyi = qi.uy, uzi = qi.uz;
    double pi = qi.pres, mi = qi.m, rhoi = qi.rho, pori = qi.DEMpor, tempi = qi.temp;
    double eta = 0.001 * h, kci = 0, cpi = 1;
    if (p.con_solve) {
      kci = conductivity(p, tab, pti, qi.temp);
      cpi = heat_capacity(p, pti);
    }
    double visi = viscosity(p, tab, pti, qi.temp) + sp.P3[i].vis_t;
    double tx = 0, ty = 0, tz = 0, Rc = 0, pgx = 0, pgy = 0, pgz = 0;
    double ex = 0, ey = 0, ez = 0, et = 0;
    const bool euler = qi.elix < 1.0 || qi.eliy < 1.0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (!(qj.p_type <= 1000)) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1), dz = pd(s, zi, qj.z, 2);
      double r = dist_of(dx, dy, dz) + 1e-20;
      if (!(r < R)) return;
      double dw = knl_dw(kt, A, h, r);
      double gx = dw * dx / r, gy = dw * dy / r, gz = dw * dz / r;
      if (p.kgc_solve > 0) kgc_apply(s, sp.P3[i].Cm, gx, gy, gz);
      const int ptj = qj.p_type;
      double uxj = qj.ux, uyj = qj.uy, uzj = qj.uz, mj = qj.m, tempj = qj.temp, rhoj = qj.rho;
      double pj = qj.pres, hj = qj.h;
      if (p.fp_solve) {
        double pp = p.pgrad_form ? pj - pi : pi + pj;
        double Cp = -pori * (mj) * pp / (rhoi * rhoj);
        tx += Cp * gx;
        ty += Cp * gy;
        tz += Cp * gz;
        pgx += rhoi / pori * Cp * gx;
        pgy += rhoi / pori * Cp * gy;
        pgz += rhoi / pori * Cp * gz;
      }
      if (p.fv_solve) {
        double visj = viscosity(p, tab, qj.p_type, qj.temp) + sp.P3[j].vis_t;
        double Cv = dx * gx + dy * gy + dz * gz;
        Cv *= (visi * visj) / (visi + visj);
        Cv *= 4 * pori * (mj / (rhoi * rhoj));
        Cv /= r;
        Cv /= r;
        if (pti <= 0) Cv = 0.0;
        tx += Cv * (uxi - uxj);
        ty += Cv * (uyi - uyj);
        tz += Cv * (uzi - uzj);
      }
      if (p.fva_solve) {
        double uij_xij = (uxi - uxj) * dx + (uyi - uyj) * dy + (uzi - uzj) * dz;
        if (uij_xij < 0) {
          double hij = (h + hj) * 0.5;
          double phi = hij * uij_xij;
          phi /= (r * r + 0.01 * hij * hij);
          double Pij = mi * phi * (-0.1 * p.soundspeed + 0.02 * phi);
          Pij /= (rhoi + rhoj);
          Pij *= 0.5;
          tx += -(Pij)*gx;
          ty += -(Pij)*gy;
          tz += -(Pij)*gz;
        }
      }
      if (p.con_solve) {
        double kcj = conductivity(p, tab, ptj, qj.temp);
        double H = 4.0 * mj * kcj * kci * (tempi - tempj) * dw;
        H /= cpi * (r + eta * eta) * rhoi * rhoj * (kci + kcj);
        Rc += H;
      }
      if (euler) {
        ex += (uxi * (uxj - uxi) * gx + uyi * (uxj - uxi) * gy + uzi * (uxj - uxi) * gz) * mj / rhoj * (1 - qi.elix);
        ey += (uxi * (uyj - uyi) * gx + uyi * (uyj - uyi) * gy + uzi * (uyj - uyi) * gz) * mj / rhoj * (1 - qi.eliy);
        ez += (uxi * (uzj - uzi) * gx + uyi * (uzj - uzi) * gy + uzi * (uzj - uzi) * gz) * mj / rhoj * (1 - qi.eliz);
        if (p.con_solve)
          et += (uxi * (tempj - tempi) * gx + uyi * (tempj - tempi) * gy + uzi * (tempj - tempi) * gz) * mj / rhoj *
                (1 - qi.elix);
      }
    });
    if (p.fg_solve) tz += -p.gravity;
    part3& o = sp.P3[i];
    o.fsx = o.fsy = o.fsz = 0;
    o.ftotalx = tx - ex;
    o.ftotaly = ty - ey;
    o.ftotalz = tz - ez;
    qi.pgf_x = pgx;
    qi.pgf_y = pgy;
    qi.pgf_z = pgz;
    o.ftotal = std::sqrt(tx * tx + ty * ty + tz * tz);
    if (p.con_solve) o.dtemp = Rc * (pti != -3) - et;
  }
}

/* KERNEL_time_update_single (functions_TIME.cuh:676-877): SP1 -> P1.  The
 * temperature branch reads `k_con_solve==1 * p_type_i>0`, which C parses as
 * k_con_solve == (p_type_i > 0): with conduction off, walls take
 * temp0 + dtemp dt and fluid keeps its temperature. */
static void wc_corrector(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const double dt = p.dt;
  for (int i = 0; i < sp.n(); i++) {
    const part1& q = sp.SP1[i];
    const part2& q2 = sp.SP2[i];
    const part3& o = sp.P3[i];
    part1& T = sp.P1[i];
    if (q.i_type == 3) continue;
    const int t = q.p_type;
    const double bz = (q.buffer_type == 0);
    T = q;  // every other field is copied from SP1
    if (t == kMoving) {
      T.x = q2.x0 + q.ux * dt;
      T.y = q2.y0 + q.uy * dt;
      T.z = q2.z0 + q.uz * dt;
    } else {
      double ux0 = q2.ux0, uy0 = q2.uy0, uz0 = q2.uz0;
      double ax = 0, ay = 0, az = 0;
      if (t > 0) {
        ax = o.ftotalx * bz;
        ay = o.ftotaly * bz;
        az = o.ftotalz * bz;
      }
      double uxc = ux0 + ax * dt, uyc = uy0 + ay * dt, uzc = uz0 + az * dt;
      if (uxc * uxc + uyc * uyc + uzc * uzc >= p.u_limit * p.u_limit) {
        uxc = ux0;
        uyc = uy0;
        uzc = uz0;
      }
      double xc = q2.x0 + uxc * dt * (t > 0) * q.elix;
      double yc = q2.y0 + uyc * dt * (t > 0) * q.eliy;
      double zc = q2.z0 + uzc * dt * (t > 0) * q.eliz;
      if (!p.xsph_solve) {
        T.x = xc;
        T.y = yc;
        T.z = zc;
      } else {
        T.x = sp.P1[i].x;
        T.y = sp.P1[i].y;
        T.z = sp.P1[i].z;
      }
      T.ux = uxc;
      T.uy = uyc;
      T.uz = uzc;
    }
    if (p.rho_type == 1 && t > 0) T.rho = q2.rho0 + o.drho * dt;
    else T.rho = q.rho;
    if (p.con_solve == (t > 0 ? 1 : 0)) T.temp = q2.temp0 + o.dtemp * dt;
    else T.temp = q.temp;
    sp.P2[i].rho_ref = q2.rho_ref;
  }
}

/* KERNEL_clc_predictor (functions_TIME.cuh:77-292) for the spheres, which
 * the reference keeps in the same arrays as the fluid: half step of
 * position, velocity and spin from the previous step's accelerations, the
 * start-of-step state saved in P2; runs before the spheres are re-sorted */
static void wc_dem_predictor(orc_state& s) {
  Species& dm = s.dem;
  const sophia_params& p = s.p;
  const double dt = p.dt;
  for (int i = 0; i < dm.n(); i++) {
    part1& q = dm.P1[i];
    part2& q2 = dm.P2[i];
    const part3& o = dm.P3[i];
    if (q.i_type == 3) continue;
    const double bz = (q.buffer_type == 0);
    const double x0 = q.x, y0 = q.y, z0 = q.z, ux0 = q.ux, uy0 = q.uy, uz0 = q.uz;
    const double wx0 = q.wx, wy0 = q.wy, wz0 = q.wz;
    q.x = x0 + ux0 * (dt * 0.5);
    q.y = y0 + uy0 * (dt * 0.5);
    q.z = z0 + uz0 * (dt * 0.5);
    q.ux = ux0 + o.ftotalx * bz * (dt * 0.5);
    q.uy = uy0 + o.ftotaly * bz * (dt * 0.5);
    q.uz = uz0 + o.ftotalz * bz * (dt * 0.5);
    q.wx = wx0 + o.torqx * bz * (dt * 0.5);
    q.wy = wy0 + o.torqy * bz * (dt * 0.5);
    q.wz = wz0 + o.torqz * bz * (dt * 0.5);
    q2.x0 = x0; q2.y0 = y0; q2.z0 = z0;
    q2.ux0 = ux0; q2.uy0 = uy0; q2.uz0 = uz0;
    q2.wx0 = wx0; q2.wy0 = wy0; q2.wz0 = wz0;
    if (p.con_solve == 1) {
      q2.temp0 = q.temp;
      q.temp = q.temp + o.dtemp * (dt * 0.5);
    }
  }
}

/* KERNEL_time_update_single (functions_TIME.cuh:676-877) for the spheres:
 * u = u0 + a dt (kept at u0 past u_limit), w = w0 + torque/inertia dt,
 * x = x0 + u dt, T = T0 + dT dt.  The reference moves nothing when XSPH is on
 * (:788-792), which freezes the spheres; here they always move. */
static void wc_dem_corrector(orc_state& s) {
  Species& dm = s.dem;
  const sophia_params& p = s.p;
  const double dt = p.dt;
  for (int i = 0; i < dm.n(); i++) {
    const part1& q = dm.SP1[i];
    const part2& q2 = dm.SP2[i];
    const part3& o = dm.P3[i];
    part1& T = dm.P1[i];
    if (q.i_type == 3) continue;
    const double bz = (q.buffer_type == 0);
    double uxc = q2.ux0 + o.ftotalx * bz * dt, uyc = q2.uy0 + o.ftotaly * bz * dt, uzc = q2.uz0 + o.ftotalz * bz * dt;
    if (uxc * uxc + uyc * uyc + uzc * uzc >= p.u_limit * p.u_limit) {
      uxc = q2.ux0;
      uyc = q2.uy0;
      uzc = q2.uz0;
    }
    T.x = q2.x0 + uxc * dt;
    T.y = q2.y0 + uyc * dt;
    T.z = q2.z0 + uzc * dt;
    T.ux = uxc; T.uy = uyc; T.uz = uzc;
    T.wx = q2.wx0 + o.torqx * bz * dt;
    T.wy = q2.wy0 + o.torqy * bz * dt;
    T.wz = q2.wz0 + o.torqz * bz * dt;
    T.temp = p.con_solve == 1 ? q2.temp0 + o.dtemp * dt : q.temp;
    T.flt_s = q.flt_s; T.flt_sd = q.flt_sd; T.flt_sd_2 = q.flt_sd_2; T.Q_sd = q.Q_sd;
    T.DEMpor = q.DEMpor;
    T.Fdx_df = q.Fdx_df; T.Fdy_df = q.Fdy_df; T.Fdz_df = q.Fdz_df;
    T.Fdx_da = q.Fdx_da; T.Fdy_da = q.Fdy_da; T.Fdz_da = q.Fdz_da;
    T.Fdx_b = q.Fdx_b; T.Fdy_b = q.Fdy_b; T.Fdz_b = q.Fdz_b;
    T.test1 = q.test1; T.test2 = q.test2; T.test6 = q.test6;
    dm.P2[i] = q2;
  }
}

/* ------------------------------------------------------------------------ */
/* 2D WCSPH.  SOPHIA_single comments out KERNEL_interaction2D               */
/* (WCSPH_Calc.cuh:187), so the 2D step follows SOPHIA_multi                */
/* (WCSPH_Calc.cuh:368-563) on one GPU's share: predictor, NNPS every step  */
/* (no KERNEL_set_alpha: Lagrangian), KGC, prep2D, prep2D_coupling,         */
/* boundary2D, continuity, EOS, interaction2D, KERNEL_time_update, xsph2D.  */
/* Spheres are refused in 2D (every 2D DEM call is commented out there).    */
/* ------------------------------------------------------------------------ */
static void wc2_set_lagrangian(std::vector<part1>& P) {
  for (part1& q : P)
    if (q.i_type != 3) q.elix = q.eliy = q.eliz = 1.0;
}

/* KERNEL_clc_correction_KGC_2D (functions_PREP2.cuh:73-159): inverse of the
 * 2x2 moment matrix, identity when |det| <= Min_det; the third row and
 * column stay 0 (P3 is reset every step) */
static void kgc_2d(orc_state& s) {
  Species& sp = s.sph;
  const int kt = s.p.kernel_type, dim = s.p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt || qi.p_type > 1000) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = s.p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z;
    double xx = 0, yy = 0, xy = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (qj.p_type > 1000) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1);
      double r = std::sqrt(dx * dx + dy * dy);
      if (!(r > 0 && r < R)) return;
      double dw = knl_dw(kt, A, h, r);
      double mtd = qj.m * dw;
      double rtd = 1.0 / (qj.rho * r);
      double t;
      t = -mtd * dx * dx; t *= rtd; xx += t;
      t = -mtd * dy * dx; t *= rtd; xy += t;
      t = -mtd * dy * dy; t *= rtd; yy += t;
    });
    double det = xx * yy - xy * xy;
    double (*C)[3] = sp.P3[i].Cm;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) C[a][b] = 0.0;
    if (std::abs(det) > kMinDet) {
      double r = 1.0 / det;
      C[0][0] = yy * r;
      C[0][1] = -xy * r;
      C[1][0] = -xy * r;
      C[1][1] = xx * r;
    } else {
      C[0][0] = 1;
      C[1][1] = 1;
    }
  }
}

/* KERNEL_clc_prep2D (functions_PREP2.cuh:1247-1470): Shepard sum every
 * freq_filt steps, reference density m / (h / 1.6)^2 every step, and the
 * Antuono density gradient -(rho_j - rho_i) m_j / rho_j grad~W (rho, not
 * rho / rho_ref as in 3D); turbulence, p_type switching and the surface-
 * tension normals are off in this build */
static void prep2d(orc_state& s) {
  Species& sp = s.sph;
  const int kt = s.p.kernel_type, dim = s.p.dim;
  bool filt = (s.count % (int64_t)std::max(1, s.p.freq_filt)) == 0;
  std::vector<double> fltv(sp.n()), gx(sp.n()), gy(sp.n());
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt || qi.p_type > 1000) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = s.p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z, rhoi = qi.rho;
    double flt = 0, rx = 0, ry = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (qj.p_type > 1000) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1);
      double r = std::sqrt(dx * dx + dy * dy) + 1e-20;
      if (!(r < R)) return;
      double w = knl_w(kt, A, h, r), dw = knl_dw(kt, A, h, r);
      double tx = dw * dx / r, ty = dw * dy / r, tz = 0.0;
      if (filt) flt += qj.m / qj.rho * w;
      if (s.p.delSPH_solve == 2) {
        kgc_apply(s, sp.P3[i].Cm, tx, ty, tz);
        rx += -(qj.rho - rhoi) * (qj.m / qj.rho) * tx;
        ry += -(qj.rho - rhoi) * (qj.m / qj.rho) * ty;
      }
    });
    fltv[i] = flt; gx[i] = rx; gy[i] = ry;
  }
  for (int i = 0; i < sp.n(); i++) {
    part1& q = sp.SP1[i];
    if (q.i_type > kITypeCrt || q.p_type > 1000) continue;
    double hv = q.h / s.p.h_over_dx;
    sp.SP2[i].rho_ref = q.m / (hv * hv);
    if (filt) q.flt_s = fltv[i];
    if (s.p.delSPH_solve == 2) {
      q.grad_rhox = gx[i];
      q.grad_rhoy = gy[i];
    }
  }
}

/* KERNEL_clc_prep2D_coupling (functions_PREP2.cuh:2944-3046): porosity
 * 1 - sum over spheres of W pi r^2; with no spheres in 2D it is 1 */
static void prep2d_coupling(orc_state& s) {
  for (part1& q : s.sph.SP1)
    if (q.i_type <= kITypeCrt) q.DEMpor = 1.0;
}

/* KERNEL_boundary2D (functions_BC.cuh:1234-1328), noslip_bc = 1: walls
 * (p_type 0) and moving walls (9) take 2 u_i [moving] minus the Shepard
 * average of the other particles' velocity (every p_type but 0, as the
 * reference's unsigned compare reads it).  All reads see the velocities
 * before the update. */
static void boundary2d(orc_state& s) {
  Species& sp = s.sph;
  const int kt = s.p.kernel_type, dim = s.p.dim;
  const int n = sp.n();
  std::vector<double> nx(n), ny(n);
  std::vector<char> act(n, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt || qi.p_type > 1000) continue;
    const uint32_t pti = (uint32_t)qi.p_type;
    if (pti > 0 && pti != (uint32_t)kMoving) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = s.p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z;
    double tx = 0, ty = 0, fl = 1.0e-10;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (qj.p_type > 1000) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1);
      double r = std::sqrt(dx * dx + dy * dy) + 1e-20;
      if (!(r < R)) return;
      double w = knl_w(kt, A, h, r);
      if ((uint32_t)qj.p_type > 0) {
        tx += qj.ux * w;
        ty += qj.uy * w;
        fl += w;
      }
    });
    nx[i] = 2 * qi.ux * (pti == (uint32_t)kMoving) - tx / fl;
    ny[i] = 2 * qi.uy * (pti == (uint32_t)kMoving) - ty / fl;
    act[i] = 1;
  }
  for (int i = 0; i < n; i++)
    if (act[i]) {
      sp.SP1[i].ux = nx[i];
      sp.SP1[i].uy = ny[i];
    }
}

/* KERNEL_clc_continuity_norm2D (functions_MASS.cuh:670-769): walls too,
 * velocities of walls taken as 0, the divergence scaled by rho_i / rho_ref_i,
 * delta-SPH on rho_j - rho_i */
static void continuity2d(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    const part1& qi = sp.SP1[i];
    if (qi.i_type > kITypeCrt || qi.p_type > 1000) continue;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z;
    double uxi = qi.ux * (qi.p_type > 0), uyi = qi.uy * (qi.p_type > 0);
    double rhoi = qi.rho, gxi = qi.grad_rhox, gyi = qi.grad_rhoy;
    double tx = 0, ty = 0, tR = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (!(qj.p_type <= 1000)) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1);
      double r = std::sqrt(dx * dx + dy * dy) + 1e-20;
      if (!(r < R)) return;
      double dw = knl_dw(kt, A, h, r);
      double gx = dw * dx / r, gy = dw * dy / r, gz = 0.0;
      if (p.kgc_solve > 0) kgc_apply(s, sp.P3[i].Cm, gx, gy, gz);
      double mj = qj.m, rhoj = qj.rho;
      double uxj = qj.ux * (qj.p_type > 0), uyj = qj.uy * (qj.p_type > 0);
      double tmprho = mj * (rhoi / rhoj);
      tx += (uxi - uxj) * tmprho * gx;
      ty += (uyi - uyj) * tmprho * gy;
      if (p.delSPH_solve > 0) {
        double phi = (gxi + qj.grad_rhox) * (-pd(s, xi, qj.x, 0));
        phi += (gyi + qj.grad_rhoy) * (-pd(s, yi, qj.y, 1));
        phi = (-0.5 * phi) * (p.delSPH_solve == 2) + (rhoj - rhoi);
        double tr = -2 * p.delta_sph * h * p.soundspeed * (mj / rhoj);
        tr *= phi * dw / r;
        tR += tr;
      }
    });
    sp.P3[i].drho = (tx + ty) * (rhoi / sp.SP2[i].rho_ref) + tR * (p.delSPH_solve > 0);
  }
}

/* KERNEL_interaction2D (functions_INTERACTION.cuh:64-411), i_type 1:
 * pressure m_j (p_i + p_j) / (eps_i rho_i eps_j rho_j), laminar viscosity
 * (no eps), artificial viscosity, gravity on y, conduction (divided by m_i
 * as the reference does); interface, repulsive, surface-tension, Boussinesq
 * and concentration terms are off in this build */
static void interaction2d(orc_state& s) {
  const sophia_table* tab = s.has_table ? &s.table : g_no_table;
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < sp.n(); i++) {
    part1& qi = sp.SP1[i];
    if (qi.i_type != 1 || qi.p_type > 1000) continue;
    const int pti = qi.p_type;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z, uxi = qi.ux, uyi = qi.uy;
    double pi = qi.pres, mi = qi.m, rhoi = qi.rho, pori = qi.DEMpor, tempi = qi.temp;
    double visi = viscosity(p, tab, pti, tempi) + sp.P3[i].vis_t;
    double kci = 0, cpi = 1, eta = 0;
    if (p.con_solve) {
      kci = conductivity(p, tab, pti, tempi);
      cpi = heat_capacity(p, pti);
      eta = 0.001 * h;
    }
    double tx = 0, ty = 0, Rc = 0, pgx = 0, pgy = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (!(qj.p_type <= 1000)) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1);
      double r = std::sqrt(dx * dx + dy * dy) + 1e-20;
      if (!(r < R)) return;
      double dw = knl_dw(kt, A, h, r);
      double gx = dw * dx / r, gy = dw * dy / r, gz = 0.0;
      if (p.kgc_solve > 0) kgc_apply(s, sp.P3[i].Cm, gx, gy, gz);
      const int ptj = qj.p_type;
      double uxj = qj.ux, uyj = qj.uy, mj = qj.m, tempj = qj.temp, rhoj = qj.rho;
      double pj = qj.pres, hj = qj.h, porj = qj.DEMpor;
      if (p.fp_solve) {
        double pp = p.pgrad_form ? pj - pi : pi + pj;
        double Cp = -mj * pp / (pori * rhoi * porj * rhoj);
        double tpx = Cp * gx, tpy = Cp * gy;
        tx += tpx;
        ty += tpy;
        pgx += rhoi * pori * tpx;
        pgy += rhoi * pori * tpy;
      }
      if (p.fv_solve) {
        double visj = viscosity(p, tab, ptj, tempj) + sp.P3[j].vis_t;
        double Cv = dx * gx + dy * gy;
        Cv *= 4 * (mj / (rhoi * rhoj));
        Cv *= (visi * visj) / (visi + visj);
        Cv /= r;
        Cv /= r;
        tx += Cv * (uxi - uxj);
        ty += Cv * (uyi - uyj);
      }
      if (p.fva_solve) {
        double uij_xij = (uxi - uxj) * dx + (uyi - uyj) * dy;
        if (uij_xij < 0) {
          double hij = (h + hj) * 0.5;
          double phi = hij * uij_xij;
          phi /= (r * r + 0.01 * hij * hij);
          double Pij = mi * phi * (-0.1 * p.soundspeed + 0.02 * phi);
          Pij /= (rhoi + rhoj);
          Pij *= 0.5;
          tx += -(Pij)*gx;
          ty += -(Pij)*gy;
        }
      }
      if (p.con_solve) {
        double kcj = conductivity(p, tab, ptj, tempj);
        double H = 4.0 * mj * kcj * kci * (tempi - tempj) * dw;
        H /= mi * cpi * (r + eta * eta) * rhoi * rhoj * (kci + kcj);
        Rc += H;
      }
    });
    if (p.fg_solve) ty += -p.gravity;
    part3& o = sp.P3[i];
    o.ftotalx = tx;
    o.ftotaly = ty;
    o.ftotalz = 0;
    o.ftotal = std::sqrt(tx * tx + ty * ty);
    qi.pgf_x = pgx;
    qi.pgf_y = pgy;
    if (p.con_solve) o.dtemp = Rc;
  }
}

/* KERNEL_time_update (functions_TIME.cuh:523-675), i_type 1: full step from
 * the start-of-step state; walls and other p_type <= 0 get u = 0, density
 * rho0 + drho dt for every p_type, temperature kept (the reference's
 * enthalpy path carries the heat); x written here only without XSPH.
 * Fields are carried from SP1 (this build keeps one copy of each field; the
 * reference's TP1 keeps its pre-boundary velocity for p_type 9). */
static void time_update2d(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const double dt = p.dt;
  for (int i = 0; i < sp.n(); i++) {
    const part1& q = sp.SP1[i];
    const part2& q2 = sp.SP2[i];
    const part3& o = sp.P3[i];
    part1& T = sp.P1[i];
    T = q;
    if (q.i_type != 1) continue;
    const int t = q.p_type;
    if (t == kMoving) {
      T.x = q2.x0 + q.ux * dt;
      T.y = q2.y0 + q.uy * dt;
      T.z = q2.z0 + q.uz * dt;
    } else {
      double ux0 = 0, uy0 = 0, uz0 = 0, ax = 0, ay = 0, az = 0;
      if (t > 0) {
        ux0 = q2.ux0; uy0 = q2.uy0; uz0 = q2.uz0;
        ax = o.ftotalx; ay = o.ftotaly; az = o.ftotalz;
      }
      double uxc = ux0 + ax * dt, uyc = uy0 + ay * dt, uzc = uz0 + az * dt;
      if (uxc * uxc + uyc * uyc + uzc * uzc >= p.u_limit * p.u_limit) {
        uxc = ux0;
        uyc = uy0;
        uzc = uz0;
      }
      if (!p.xsph_solve) {
        T.x = q2.x0 + uxc * dt;
        T.y = q2.y0 + uyc * dt;
        T.z = q2.z0 + uzc * dt;
      }
      T.ux = uxc;
      T.uy = uyc;
      T.uz = uzc;
    }
    if (p.rho_type == 1) T.rho = q2.rho0 + o.drho * dt;
  }
}

/* KERNEL_xsph2D (functions_TIME.cuh:1520-1607): x = x0 + (u + c sum_j
 * m_j / rho_j (u_j - u_i) W / flt_s_i) dt for fluid, x0 for walls, with
 * the corrected velocities (P1) and the predicted positions, masses and
 * densities (SP1); the velocity itself is not smoothed */
static void xsph2d(orc_state& s) {
  Species& sp = s.sph;
  const sophia_params& p = s.p;
  const int kt = p.kernel_type, dim = p.dim;
  const double dt = p.dt;
  const int n = sp.n();
  std::vector<double> nx(n), ny(n);
  std::vector<char> act(n, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const part1& qi = sp.SP1[i];
    if (qi.i_type != 1 || qi.p_type > 1000) continue;
    const int pti = qi.p_type;
    double h = qi.h, A = knl_tmpA(kt, dim, h), R = p.kappa * h;
    double xi = qi.x, yi = qi.y, zi = qi.z;
    double uxi = sp.P1[i].ux, uyi = sp.P1[i].uy, fl = qi.flt_s;
    double tx = 0, ty = 0;
    for_candidates(s, sp, xi, yi, zi, [&](int j) {
      const part1& qj = sp.SP1[j];
      if (qj.p_type > 1000) return;
      double dx = pd(s, xi, qj.x, 0), dy = pd(s, yi, qj.y, 1);
      double r = std::sqrt(dx * dx + dy * dy);
      if (!(r < R)) return;
      double w = knl_w(kt, A, h, r);
      double uxj = sp.P1[j].ux, uyj = sp.P1[j].uy;
      tx += p.c_xsph * qj.m / qj.rho * (-uxi + uxj) * w / fl;
      ty += p.c_xsph * qj.m / qj.rho * (-uyi + uyj) * w / fl;
    });
    const part2& q2 = sp.SP2[i];
    if (pti == kMoving) {
      nx[i] = q2.x0 + uxi * dt;
      ny[i] = q2.y0 + uyi * dt;
    } else {
      nx[i] = q2.x0 + (uxi + tx) * dt * (pti > 0);
      ny[i] = q2.y0 + (uyi + ty) * dt * (pti > 0);
    }
    act[i] = 1;
  }
  for (int i = 0; i < n; i++)
    if (act[i]) {
      sp.P1[i].x = nx[i];
      sp.P1[i].y = ny[i];
    }
}

static int wc_check(const orc_state& s) {
  if (s.p.rho_type != 1) return -1;
  if (s.p.dim == 2) return s.dem.n() == 0 ? 0 : -1;
  if (s.p.dim != 3) return -1;
  return 0;
}

static int wc2_stage(orc_state* s, int32_t stage) {
  Species& sp = s->sph;
  switch (stage) {
    case 0:
      wc2_set_lagrangian(sp.P1);
      if (s->p.time_type == 1) wc_predictor(*s);
      nnps_species(*s, sp, true);
      wc2_set_lagrangian(sp.SP1);
      wc2_set_lagrangian(sp.P1);
      s->after_corrector = false;
      break;
    case 1:
      break;  // SOPHIA_multi has no volume or mass update
    case 2:
      if (s->p.kgc_solve > 0 || s->p.delSPH_solve == 2) kgc_2d(*s);
      prep2d(*s);
      prep2d_coupling(*s);
      if (s->p.noslip_bc == 1) boundary2d(*s);
      break;
    case 3:
      continuity2d(*s);
      break;
    case 4:
      wc_eos(*s);
      break;
    case 5:
      break;  // no Neumann wall extrapolation in 2D
    case 6:
      interaction2d(*s);
      break;
    case 7:
      time_update2d(*s);
      s->after_corrector = true;
      break;
    case 8:
      if (s->p.xsph_solve) xsph2d(*s);
      s->time += s->p.dt;
      s->count += 1;
      break;
    default:
      return -1;
  }
  return 0;
}

/* the WCSPH stages in driver order; stage 0 also does the NNPS (count 0) or
 * the P1 -> SP1 copy (WCSPH_Calc.cuh:34-69) */
int orc_wcsph_stage(orc_state* s, int32_t stage) {
  if (wc_check(*s)) return -1;
  if (s->p.dim == 2) return wc2_stage(s, stage);
  Species& sp = s->sph;
  switch (stage) {
    case 0:
      wc_set_alpha(sp, sp.P1);
      if (s->p.time_type == 1) wc_predictor(*s);
      if (s->count == 0) {
        nnps_species(*s, sp, true);
        wc_set_alpha(sp, sp.SP1);
        wc_set_alpha(sp, sp.P1);
      } else {
        std::fill(sp.P3.begin(), sp.P3.end(), part3{});
        sp.SP1 = sp.P1;
        sp.SP2 = sp.P2;
      }
      if (s->dem.n() > 0) {
        // spheres move: half step, then a fresh sort every step (the
        // reference's WCSPH driver sorts once, which only its Eulerian
        // fluid allows)
        if (s->p.time_type == 1) wc_dem_predictor(*s);
        nnps_species(*s, s->dem, false);
      }
      s->after_corrector = false;
      break;
    case 1:
      wc_volume_update(*s);
      wc_mass_update(*s);
      break;
    case 2:
      if (s->p.kgc_solve > 0 || s->p.delSPH_solve == 2) kgc_3d(*s);
      coupling_prep_sph(*s);
      if (s->dem.n() > 0) coupling_prep_dem(*s);  // KERNEL_clc_prep3D_coupling (functions_PREP2.cuh:3048)
      prep3d_sph(*s);
      break;
    case 3:
      wc_continuity(*s);
      break;
    case 4:
      wc_eos(*s);
      break;
    case 5:
      neumann_walls(*s);
      break;
    case 6:
      wc_interaction(*s);
      if (s->dem.n() > 0) {
        // Hertz-Mindlin contact (commented out in the reference's WCSPH
        // driver, WCSPH_Calc.cuh:206), then KERNEL_DEM_coupling3D and
        // KERNEL_SPH_coupling3D (:225, 236)
        orc_dem_interaction(s);
        dem_coupling(*s);
        sph_coupling(*s);
      }
      break;
    case 7:
      wc_corrector(*s);
      if (s->dem.n() > 0) {
        wc_dem_corrector(*s);
        if (s->p.dem_walls) dem_walls(*s);
      }
      // buffer state applied after the time update (computed but not applied
      // by the reference's WCSPH driver, WCSPH_Calc.cuh:282)
      if (s->p.open_boundary == 1) open_boundary(*s);
      s->after_corrector = true;
      break;
    case 8:
      if (s->p.xsph_solve) xsph_sph(s, false);
      s->time += s->p.dt;
      s->count += 1;
      break;
    default:
      return -1;
  }
  return 0;
}

/* one ISPH-DEM time step in the order of SOPHIA_single_ISPH
 * (ISPH_Calc_seperate.cuh:13-503); open-boundary and DEM wall kernels are
 * outside this build's scope */
int orc_step(orc_state* s, const sophia_ppe_opts* o, sophia_ppe_stats* st) {
  if (s->p.solver_type == 0) {
    for (int k = 0; k <= 8; k++)
      if (orc_wcsph_stage(s, k)) return -1;
    return 0;
  }
  orc_nnps(s);
  orc_prep(s);
  orc_interaction(s);
  orc_dem_interaction(s);
  orc_sph_dem_coupling(s);
  orc_time(s, SOPHIA_PREDICT);
  orc_ppe(s, o, st);
  orc_time(s, SOPHIA_CORRECT);
  orc_xsph(s);
  s->time += s->p.dt;
  s->count += 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
void orc_get_nnps(const orc_state* s, int32_t species, int32_t* keys, int32_t* order,
                  uint32_t* cell_start, uint32_t* cell_end, int32_t* nbr_count) {
  const Species& sp = species == SOPHIA_SPH ? s->sph : s->dem;
  int n = sp.n();
  if (keys) std::copy(sp.key.begin(), sp.key.end(), keys);
  if (order) std::copy(sp.order.begin(), sp.order.end(), order);
  if (cell_start) std::copy(sp.g_str.begin(), sp.g_str.begin() + s->d.num_cells, cell_start);
  if (cell_end) std::copy(sp.g_end.begin(), sp.g_end.begin() + s->d.num_cells, cell_end);
  if (nbr_count) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      const part1& qi = sp.SP1[i];
      double R = s->p.kappa * qi.h;
      int c = 0;
      for_candidates(*s, sp, qi.x, qi.y, qi.z, [&](int j) {
        const part1& qj = sp.SP1[j];
        double r = dist_of(pd(*s, qi.x, qj.x, 0), pd(*s, qi.y, qj.y, 1), pd(*s, qi.z, qj.z, 2)) + 1e-20;
        if (r < R) c++;
      });
      nbr_count[i] = c;
    }
  }
}

void orc_get_particles(const orc_state* s, int32_t species, part1* p1, part2* p2, part3* p3) {
  const Species& sp = species == SOPHIA_SPH ? s->sph : s->dem;
  int n = sp.n();
  if (p1) {
    const std::vector<part1>& src = s->after_corrector ? sp.P1 : sp.SP1;
    std::copy(src.begin(), src.end(), p1);
    if (species == SOPHIA_DEM && !s->after_corrector) {
      for (int i = 0; i < n; i++) {
        std::memcpy(p1[i].ct_idx, sp.P1[i].ct_idx, sizeof(p1[i].ct_idx));
        std::memcpy(p1[i].del_s, sp.P1[i].del_s, sizeof(p1[i].del_s));
      }
    }
  }
  if (p2) std::copy(sp.SP2.begin(), sp.SP2.end(), p2);
  if (p3) std::copy(sp.P3.begin(), sp.P3.end(), p3);
}

void orc_set_pressure(orc_state* s, const double* p, int32_t n) {
  Species& sp = s->sph;
  for (int i = 0; i < n && i < sp.n(); i++) {
    sp.SP1[i].pres = p[i];
    sp.P1[i].pres = p[i];
  }
}

void orc_get_ppe_system(const orc_state* s, double* diag, double* rhs, int32_t* row_mask) {
  const Species& sp = s->sph;
  int n = sp.n();
  for (int i = 0; i < n; i++) {
    bool on = i < (int)s->row_mask.size() && s->row_mask[i];
    if (diag) diag[i] = on ? sp.SP1[i].PPE2 : 0.0;
    if (rhs) rhs[i] = on ? s->rhs[i] : 0.0;
    if (row_mask) row_mask[i] = on;
  }
}

void orc_ppe_apply(const orc_state* s, const double* x, double* y) {
  int n = s->sph.n();
  std::vector<double> xv(x, x + n), yv(n);
  ppe_matvec(*s, xv, yv);
  std::copy(yv.begin(), yv.end(), y);
}

double orc_knl_tmpA(int32_t kt, int32_t dim, double h) { return knl_tmpA(kt, dim, h); }

void orc_knl_eval(int32_t kt, int32_t dim, double h, const double* r, int32_t n, double* w, double* dw) {
  double A = knl_tmpA(kt, dim, h);
  for (int i = 0; i < n; i++) {
    if (w) w[i] = knl_w(kt, A, h, r[i]);
    if (dw) dw[i] = knl_dw(kt, A, h, r[i]);
  }
}

int32_t orc_cell_key(const orc_state* s, double x, double y, double z) {
  int a, b, c;
  cell_of(*s, x, y, z, a, b, c);
  return idx_cell(*s, a, b, c);
}

void orc_brute_neighbor_counts(const orc_state* s, int32_t* counts) {
  const Species& sp = s->sph;
  int n = sp.n();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    const part1& qi = sp.SP1[i];
    double R = s->p.kappa * qi.h;
    int c = 0;
    for (int j = 0; j < n; j++) {
      const part1& qj = sp.SP1[j];
      double r = dist_of(pd(*s, qi.x, qj.x, 0), pd(*s, qi.y, qj.y, 1), pd(*s, qi.z, qj.z, 2)) + 1e-20;
      if (r < R) c++;
    }
    counts[i] = c;
  }
}

}  // extern "C"
