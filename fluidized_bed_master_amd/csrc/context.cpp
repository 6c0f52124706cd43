// Host side of the C ABI: one context owns the device state of one run and
// drives the stage kernels in the order of SOPHIA_single_ISPH
// (ISPH_Calc_seperate.cuh:13-503) on a single HIP stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "learned.hpp"
#include "ppe.hpp"
#include "coarse.hpp"
#include "timers.hpp"

static_assert(sizeof(part1) == 1272, "part1 layout must match class_Cuda_Particle_Array2.cuh:11-85");
static_assert(sizeof(part2) == 176, "part2 layout must match class_Cuda_Particle_Array2.cuh:87-106");
static_assert(sizeof(part3) == 496, "part3 layout must match class_Cuda_Particle_Array2.cuh:108-167");

namespace sophia {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
  set_error("HIP error %d (%s) in %s at %s:%d", (int)e, hipGetErrorString(e), what, file, line);
  return e == hipErrorOutOfMemory ? SOPHIA_E_NOMEM : SOPHIA_E_HIP;
}

// ---------------------------------------------------------------------------
// device memory bookkeeping
// ---------------------------------------------------------------------------
struct Arena {
  std::vector<void*> ptrs;
  ~Arena() { release(); }
  void release() {
    for (void* p : ptrs) (void)hipFree(p);
    ptrs.clear();
  }
  template <typename T>
  int alloc(T** p, size_t count) {
    void* q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    // every array starts at zero: fields a stage leaves untouched for some
    // particles read as the reference's zero-initialised records
    e = hipMemset(q, 0, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMemset", __FILE__, __LINE__);
    ptrs.push_back(q);
    *p = (T*)q;
    return 0;
  }
  void drop(void* p) {
    auto it = std::find(ptrs.begin(), ptrs.end(), p);
    if (it != ptrs.end()) {
      (void)hipFree(p);
      ptrs.erase(it);
    }
  }
};


// persistent field with its permutation spare
template <typename T>
struct PField {
  T* cur = nullptr;
  T* spare = nullptr;
  int alloc(Arena& a, size_t n) {
    TRY(a.alloc(&cur, n));
    return a.alloc(&spare, n);
  }
  void swap() { std::swap(cur, spare); }
};

struct BaseCtx {
  virtual ~BaseCtx() {}
};

}  // namespace sophia

using namespace sophia;

// ---------------------------------------------------------------------------
// the context
// ---------------------------------------------------------------------------
struct sophia_ctx {
  virtual ~sophia_ctx() {}
  sophia_params p{};
  sophia_domain d{};
  double pgrid = 1.0;  // pos4 grid spacing (power of two; period / 1024 when periodic)
  // periodic axes (bit 0 x, 1 y, 2 z) of a fixed domain and their periods
  int pmask = 0;
  double plen[3] = {0, 0, 0};
  int device = 0;
  int precision = SOPHIA_PREC_MIXED;
  hipStream_t stream = nullptr;
  int64_t count = 0;
  double time = 0;
  bool nnps_done = false;
  bool after_corrector = false;
  // The SPH cell order, ranges, neighbour list, position records and surface
  // labels stay valid while no SPH position changes: ISPH with Eulerian SPH
  // (eulerian_sph = 1) never moves an SPH particle after the first sort, so
  // a later sort of the already sorted keys is the identity and is skipped.
  bool sph_sorted = false;
  bool sph_order_identity = false;
  virtual int nnps() = 0;
  virtual int prep() = 0;
  virtual int interaction() = 0;
  virtual int dem_interaction() = 0;
  // interaction + dem_interaction with the sphere contact beside advection
  virtual int interaction_both() = 0;
  virtual int coupling() = 0;
  virtual int time_stage(int phase) = 0;
  virtual int ppe(const sophia_ppe_opts* o, sophia_ppe_stats* st) = 0;
  virtual int xsph() = 0;
  virtual int get_nnps(int species, int32_t* keys, int32_t* order, uint32_t* cs, uint32_t* ce, int32_t* nc) const = 0;
  virtual int get_particles(int species, part1* p1, part2* p2, part3* p3) const = 0;
  virtual int set_pressure(const double* p, int n) = 0;
  virtual int get_ppe_system(double* diag, double* rhs, int32_t* row) const = 0;
  virtual int ppe_apply(const double* x, double* y, int n) = 0;
  virtual int matvec_probe(int variant, int reps, const double* x, double* y, int n, float* ms) = 0;
  virtual int ppe_layout(int64_t* nnz, int32_t* ib, int32_t* cb) const = 0;
  virtual int ppe_coarse(int32_t* dims, float* E, float* X) const = 0;
  virtual int monitors(int species, double* out) = 0;
  virtual int wcsph_stage(int stage) = 0;
  virtual int set_table(const sophia_table& t) = 0;
  virtual int mlp_predict(double* out, int n) = 0;
  virtual int mlp_train(const double* target, int n, double lr, double* loss, float* grad, int n_params, bool step) = 0;
  virtual int mlp_set(const float* w, int n) = 0;
  virtual int mlp_get(float* w, int n) const = 0;
  virtual int mlp_reset() = 0;
  virtual int mlp_feat(float* nf, int n, double* scale) = 0;
  Timers timers;
  Timers* tm() { return timers.on ? &timers : nullptr; }
};

namespace {

// c_initial_inner_outer_particle_single_sph/_dem (functions_NNPS.cuh:8-76,
// 138-200), with open boundaries: every particle starts as i_type 1,
// buffer_type 0; fluid (1 <= p_type < 1000) with z in [L1, L2) becomes an
// inlet buffer particle (buffer_type 1, i_type 2), in [L3, L4) an outlet
// one (buffer_type 2, i_type 2); both tests carry the reference's -1e-6
void tag_open_boundary(const sophia_params& p, std::vector<part1>& P) {
  const double* L = p.ob_band;
  for (part1& q : P) {
    q.i_type = 1;
    q.buffer_type = 0;
    if (q.p_type < 1000 && q.p_type >= 1) {
      if (q.z >= (L[0] - 1e-6) && q.z < (L[1] - 1e-6)) {
        q.buffer_type = 1;
        q.i_type = 2;
      }
      if (q.z >= (L[2] - 1e-6) && q.z < (L[3] - 1e-6)) {
        q.buffer_type = 2;
        q.i_type = 2;
      }
    }
  }
}

int ceil_log2(uint32_t v) {
  int b = 0;
  while ((1ull << b) < (uint64_t)v) b++;
  return b;
}

template <typename R>
struct Ctx : public sophia_ctx {
  Arena arena;
  int ns = 0, nd = 0;
  // SPH persistent
  PField<double> sx, sy, sz, sh, spres, sm, srho;
  PField<double> sph1, sph2;  // pressure of the two solves before the last one
  int p_hist = 0;             // valid history levels (0..2)
  double* pbase = nullptr;    // extrapolated PPE guess
  double* mon_scratch = nullptr;
  PField<R> sux, suy, suz, stemp, sflt, sdpor;
  PField<int> sptype, sbtype, sitype;
  // DEM persistent
  PField<double> dx_, dy_, dz_, dux, duy, duz, dwx, dwy, dwz, dm, dri, drad, dh, drho, dtemp, dfdx, dfdy, dfdz, dqsd;
  PField<int> dptype, dbtype, ditype, didx;
  // WCSPH with spheres: start-of-step x0, y0, z0, ux0, uy0, uz0, wx0, wy0,
  // wz0, temp0, written before the re-sort and permuted with the spheres
  PField<double> dws[10];
  bool wc_dem = false;
  // 2D WCSPH re-sorts the fluid every step after the predictor: the
  // start-of-step state it wrote (x0, y0, z0, rho0, rho_ref; ux0, uy0, uz0,
  // temp0) is permuted with the particles through these spares
  bool wc_lag = false;
  double* lag_d[5] = {};
  R* lag_r[4] = {};
  int* hist_idx[2] = {nullptr, nullptr};
  double* hist_ds[2] = {nullptr, nullptr};
  SphView<R> sv{};
  DemView dv{};
  // cells and sort
  uint32_t *scs = nullptr, *sce = nullptr, *dcs = nullptr, *dce = nullptr;
  uint32_t* scoff = nullptr;  // SPH first particle of each cell, dense (num_cells + 1)
  SortScratch ssort{}, dsort{};
  uint32_t *skeys = nullptr, *sorder = nullptr, *dkeys = nullptr, *dorder = nullptr;
  int* maxcnt = nullptr;
  PpeWork ppe_w;
  bool ppe_assembled = false;
  int n_rows = 0;
  // property tables (data.txt) and their device copy: T, h, k, mu blocks of
  // SOPHIA_TABLE_VALUES each
  sophia_table table{};
  bool has_table = false;
  double* dtab = nullptr;
  Mlp mlp;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // second stream for sphere-side kernels that overlap SPH kernels they do
  // not share data with (default priority, like the main stream)
  hipStream_t dem_stream = nullptr;
  hipEvent_t ev_dem_in = nullptr, ev_dem_out = nullptr;

  ~Ctx() override {
    if (dem_stream) (void)hipStreamSynchronize(dem_stream);
    ppe_workspace_free(ppe_w);
    mlp_free(mlp);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_dem_in) (void)hipEventDestroy(ev_dem_in);
    if (ev_dem_out) (void)hipEventDestroy(ev_dem_out);
    arena.release();
    if (dem_stream) (void)hipStreamDestroy(dem_stream);
    if (stream) (void)hipStreamDestroy(stream);
  }

  DevParams devparams() const {
    DevParams P{};
    P.dim = p.dim;
    P.kernel_type = p.kernel_type;
    P.kgc = p.kgc_solve;
    P.delsph = p.delSPH_solve;
    P.rho_type = p.rho_type;
    P.fp_solve = p.fp_solve;
    P.fv_solve = p.fv_solve;
    P.fva_solve = p.fva_solve;
    P.fg_solve = p.fg_solve;
    P.con_solve = p.con_solve;
    P.xsph_solve = p.xsph_solve;
    P.open_boundary = p.open_boundary;
    P.z_index = p.flag_z_index;
    // 3D WCSPH: KERNEL_set_alpha (functions_ALE.cuh:2-20) makes SPH
    // Eulerian; the 2D step (SOPHIA_multi) has no set_alpha: Lagrangian
    P.eulerian = p.solver_type == 0 ? (p.dim == 3 ? 1 : 0) : p.eulerian_sph;
    P.fix_xsph = p.fix_xsph_position;
    P.pgrad_form = p.pgrad_form;
    P.NI = d.NI;
    P.NJ = d.NJ;
    P.NK = d.NK;
    P.num_cells = d.num_cells;
    P.dem_slots = std::min(kMaxHist, std::max(1, p.dem_max_contacts));
    int stride = std::max(1, p.decouple_stride);
    P.filt_step = (count % ((int64_t)std::max(1, p.freq_filt) * stride)) == 0;
    P.first_step = time == 0.0;
    P.x_min = d.x_min; P.x_max = d.x_max;
    P.y_min = d.y_min; P.y_max = d.y_max;
    P.z_min = d.z_min; P.z_max = d.z_max;
    P.kappa = p.kappa;
    P.dt_sph = stride * p.dt;
    P.dt_dem = p.dt;
    P.u_limit = p.u_limit;
    P.c_xsph = p.c_xsph;
    P.soundspeed = p.soundspeed;
    P.gravity = p.gravity;
    P.rho_fluid = p.rho_fluid;
    P.rho_water = p.rho_water;
    P.h_over_dx = p.h_over_dx;
    P.mu_fluid = p.mu_fluid;
    P.mu_water = p.mu_water;
    P.k_fluid = p.k_fluid;
    P.k_dem = p.k_dem;
    P.cp_fluid = p.cp_fluid;
    P.cp_dem = p.cp_dem;
    P.dem_G = p.dem_shear_modulus;
    P.dem_nu = p.dem_poisson;
    P.dem_friction = p.dem_friction;
    P.dem_damping = p.dem_damping;
    P.dem_hertz = p.dem_hertz_factor;
    P.drag_switch = p.drag_switch;
    P.eos_b = p.eos_b_factor;
    P.gamma = p.gamma;
    P.delta_sph = p.delta_sph;
    P.p_ref = p.p_ref;
    P.compat = p.reference_compat;
    P.free_surface = p.free_surface;
    P.fs_threshold = p.fs_threshold;
    P.pr_fluid = p.pr_fluid;
    P.pr_air = p.pr_air;
    for (int k = 0; k < 2; k++) {
      P.dem_k_mat[k] = p.dem_k_mat[k];
      P.dem_alpha_mat[k] = p.dem_alpha_mat[k];
    }
    P.dem_mat_radius = p.dem_mat_radius;
    P.heat_flux = p.con_solve == 1 ? p.heat_flux : 0.0;
    for (int k = 0; k < 6; k++) P.heat_flux_box[k] = p.heat_flux_box[k];
    P.wcsph = p.solver_type == 0 ? 1 : 0;
    P.noslip = p.noslip_bc;
    P.pmask = pmask;
    for (int k = 0; k < 3; k++) P.plen[k] = plen[k];
    P.prop_table = (p.prop_table && dtab) ? 1 : 0;
    P.tab_types = has_table ? SOPHIA_TABLE_TYPES : 0;
    for (int k = 0; k < SOPHIA_TABLE_TYPES; k++) {
      P.tab_index[k] = table.index[k];
      P.tab_size[k] = table.size[k];
    }
    P.tab = dtab;
    return P;
  }

  bool sph_on() const { return (count % std::max(1, p.decouple_stride)) == 0; }

  void refresh_views() {
    SphView<R>& v = sv;
    v.n = ns;
    v.x = sx.cur; v.y = sy.cur; v.z = sz.cur; v.h = sh.cur; v.pres = spres.cur;
    v.ux = sux.cur; v.uy = suy.cur; v.uz = suz.cur; v.m = sm.cur; v.rho = srho.cur; v.temp = stemp.cur;
    v.flt_s = sflt.cur; v.dpor = sdpor.cur;
    v.ptype = sptype.cur; v.btype = sbtype.cur; v.itype = sitype.cur;
    v.pg0x = d.x_min; v.pg0y = d.y_min; v.pg0z = d.z_min;
    v.pgrid = (float)pgrid;
    v.pmask = pmask;
    v.plx = plen[0]; v.ply = plen[1]; v.plz = plen[2];
    DemView& w = dv;
    w.n = nd;
    w.x = dx_.cur; w.y = dy_.cur; w.z = dz_.cur; w.ux = dux.cur; w.uy = duy.cur; w.uz = duz.cur;
    w.wx = dwx.cur; w.wy = dwy.cur; w.wz = dwz.cur; w.m = dm.cur; w.ri = dri.cur; w.rad = drad.cur;
    w.h = dh.cur; w.rho = drho.cur; w.temp = dtemp.cur; w.fdx = dfdx.cur; w.fdy = dfdy.cur; w.fdz = dfdz.cur;
    w.qsd = dqsd.cur; w.ptype = dptype.cur; w.btype = dbtype.cur; w.itype = ditype.cur; w.dem_idx = didx.cur;
    w.hist_idx = hist_idx[0]; w.hist_ds = hist_ds[0];
    w.hist_idx_new = hist_idx[1]; w.hist_ds_new = hist_ds[1];
    if (wc_dem) {
      w.x0 = dws[0].cur; w.y0 = dws[1].cur; w.z0 = dws[2].cur;
      w.ux0 = dws[3].cur; w.uy0 = dws[4].cur; w.uz0 = dws[5].cur;
      w.wx0 = dws[6].cur; w.wy0 = dws[7].cur; w.wz0 = dws[8].cur;
      w.temp0 = dws[9].cur;
    }
  }

  int alloc_nbr(int cap) {
    if (sv.nbr) arena.drop(sv.nbr);
    cap = (cap + kNbrChunk - 1) / kNbrChunk * kNbrChunk;  // whole chunks
    sv.nbr_cap = cap;
    TRY(arena.alloc(&sv.nbr, (size_t)cap * std::max(ns, 1)));
    ppe_workspace_free(ppe_w);
    TRY(ppe_workspace_alloc(ppe_w, ns, cap, precision == SOPHIA_PREC_MIXED));
    ppe_assembled = false;
    return 0;
  }

  int init(const sophia_params& prm, const part1* P, int n) {
    p = prm;
    count = 0;
    time = prm.time;
    HIP_TRY(hipSetDevice(device));
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&dem_stream, hipStreamNonBlocking));
    timers.stream = stream;
    HIP_TRY(hipEventCreate(&ev0));
    HIP_TRY(hipEventCreate(&ev1));
    HIP_TRY(hipEventCreateWithFlags(&ev_dem_in, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&ev_dem_out, hipEventDisableTiming));
    setup_domain(P, n);
    std::vector<part1> tagged;
    if (p.open_boundary == 1) {
      tagged.assign(P, P + n);
      tag_open_boundary(p, tagged);
      P = tagged.data();
    }
    std::vector<int> si, di;
    for (int i = 0; i < n; i++) (P[i].p_type <= 1000 ? si : di).push_back(i);
    ns = (int)si.size();
    nd = (int)di.size();
    for (int k : si) {
      const part1& q = P[k];
      if (q.i_type <= 2 && !(p.open_boundary > 0 && q.buffer_type > 0) && q.p_type > 0 && q.p_type < 1000) n_rows++;
    }
    d.n_sph = ns;
    d.n_dem = nd;
    size_t NS = ns, ND = nd;
    // SPH persistent
    for (auto* f : {&sx, &sy, &sz, &sh, &spres, &sm, &srho, &sph1, &sph2}) TRY(f->alloc(arena, NS));
    TRY(arena.alloc(&pbase, NS));
    TRY(arena.alloc(&mon_scratch, (size_t)3 * 1024 + 3));
    for (auto* f : {&sux, &suy, &suz, &stemp, &sflt, &sdpor}) TRY(f->alloc(arena, NS));
    for (auto* f : {&sptype, &sbtype, &sitype}) TRY(f->alloc(arena, NS));
    // SPH scratch
    SphView<R>& v = sv;
    TRY(arena.alloc(&v.por, NS));
    TRY(arena.alloc(&v.flux_s, NS));
    TRY(arena.alloc(&v.surf, NS));
    if (NS) TRY(fill_u32(stream, (uint32_t*)v.surf, 0u, NS));
    TRY(arena.alloc(&v.visp, NS));
    TRY(arena.alloc(&v.condp, NS));
    for (double** f : {&v.rho_ref, &v.rho0, &v.vol, &v.vol0, &v.x0, &v.y0, &v.z0}) TRY(arena.alloc(f, NS));
    wc_lag = p.solver_type == 0 && p.dim == 2;
    if (wc_lag) {
      for (auto& f : lag_d) TRY(arena.alloc(&f, NS));
      for (auto& f : lag_r) TRY(arena.alloc(&f, NS));
      for (double** f : {&v.xp, &v.yp, &v.rhop}) TRY(arena.alloc(f, NS));
    }
    for (R** f : {&v.dpor_prev, &v.grx, &v.gry, &v.grz, &v.fx, &v.fy, &v.fz, &v.fmag,
                  &v.fpx, &v.fpy, &v.fpz, &v.pgx, &v.pgy, &v.pgz, &v.dtemp, &v.temp0, &v.ux0, &v.uy0, &v.uz0,
                  &v.drho, &v.xsx, &v.xsy, &v.xsz, &v.xst})
      TRY(arena.alloc(f, NS));
    TRY(arena.alloc(&v.cm, 6 * NS));
    for (double** f : {&v.ppe1, &v.ppe2, &v.ppe3, &v.ppe4}) TRY(arena.alloc(f, NS));
    TRY(arena.alloc(&v.nbr_cnt, NS));
    TRY(arena.alloc(&v.pk, NS));
    TRY(arena.alloc(&v.pk2, NS));
    TRY(arena.alloc(&v.pkv, NS));
    TRY(arena.alloc(&v.pkh, NS));
    TRY(arena.alloc(&v.r32, NS));
    TRY(arena.alloc(&v.ra, NS));
    TRY(arena.alloc(&v.rn, NS));
    TRY(arena.alloc(&v.pkb, NS));
    TRY(arena.alloc(&v.pkc, NS));
    TRY(arena.alloc(&v.pos4, NS));
    // DEM persistent
    for (auto* f : {&dx_, &dy_, &dz_, &dux, &duy, &duz, &dwx, &dwy, &dwz, &dm, &dri, &drad, &dh, &drho, &dtemp, &dfdx,
                    &dfdy, &dfdz, &dqsd})
      TRY(f->alloc(arena, ND));
    for (auto* f : {&dptype, &dbtype, &ditype, &didx}) TRY(f->alloc(arena, ND));
    wc_dem = p.solver_type == 0 && nd > 0;
    if (wc_dem)
      for (auto& f : dws) {
        TRY(f.alloc(arena, ND));
        TRY(fill_u32(stream, (uint32_t*)f.cur, 0u, ND * 2));
      }
    for (int b = 0; b < 2; b++) {
      TRY(arena.alloc(&hist_idx[b], (size_t)kMaxHist * ND));
      TRY(arena.alloc(&hist_ds[b], (size_t)kMaxHist * 3 * ND));
    }
    DemView& w = dv;
    for (double** f : {&w.flt_s, &w.flt_sd, &w.flt_sd2, &w.por, &w.fax, &w.fay, &w.faz, &w.fbx, &w.fby, &w.fbz,
                       &w.fx, &w.fy, &w.fz, &w.tqx, &w.tqy, &w.tqz, &w.ux0, &w.uy0, &w.uz0, &w.test6,
                       &w.dtemp, &w.temp0, &w.hrad, &w.hcon})
      TRY(arena.alloc(f, ND));
    // cells and sort scratch
    size_t NC = (size_t)d.num_cells;
    TRY(arena.alloc(&scs, NC));
    TRY(arena.alloc(&sce, NC));
    TRY(arena.alloc(&scoff, NC + 1));
    TRY(arena.alloc(&dcs, NC));
    TRY(arena.alloc(&dce, NC));
    for (auto pr : {std::make_pair(&ssort, ns), std::make_pair(&dsort, nd)}) {
      SortScratch& sc = *pr.first;
      size_t m = std::max(pr.second, 1);
      TRY(arena.alloc(&sc.keys_a, m));
      TRY(arena.alloc(&sc.keys_b, m));
      TRY(arena.alloc(&sc.vals_a, m));
      TRY(arena.alloc(&sc.vals_b, m));
      sc.max_tiles = (int)((m + 4095) / 4096);
      TRY(arena.alloc(&sc.hist, (size_t)256 * sc.max_tiles + 256));
    }
    TRY(arena.alloc(&maxcnt, 1));
    int cap = p.sph_max_neighbors > 0 ? p.sph_max_neighbors : 96;
    TRY(alloc_nbr(cap));
    TRY(mlp_alloc(mlp, ns, sv.nbr_cap));
    // upload
    refresh_views();
    TRY(upload(P, si, di));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
  }

  // find_minmax (function_init.cuh:913-950) over all particles except the
  // last, margins in the reference's order; dcell = 1.1*kappa*h0 with h0 the
  // first particle's h; NI = (int)(span/dcell)+1 (ISPH.cuh:93-113)
  void setup_domain(const part1* P, int n) {
    if (p.domain_fixed) {
      d.x_min = p.domain_min[0]; d.x_max = p.domain_max[0];
      d.y_min = p.domain_min[1]; d.y_max = p.domain_max[1];
      d.z_min = p.domain_min[2]; d.z_max = p.domain_max[2];
    } else {
      double a = P[0].x, b = P[0].x, c = P[0].y, e = P[0].y, f = P[0].z, g = P[0].z;
      for (int i = 0; i < n - 1; i++) {
        a = P[i].x < a ? P[i].x : a; b = P[i].x > b ? P[i].x : b;
        c = P[i].y < c ? P[i].y : c; e = P[i].y > e ? P[i].y : e;
        f = P[i].z < f ? P[i].z : f; g = P[i].z > g ? P[i].z : g;
      }
      a -= (b - a) * p.xmargin_m; b += (b - a) * p.xmargin_p;
      c -= (e - c) * p.ymargin_m; e += (e - c) * p.ymargin_p;
      f -= (g - f) * p.zmargin_m; g += (g - f) * p.zmargin_p;
      d.x_min = a; d.x_max = b; d.y_min = c; d.y_max = e; d.z_min = f; d.z_max = g;
    }
    d.dcell = 1.1 * p.kappa * P[0].h;
    // pos4 grid: the smallest power-of-two spacing with at most 1000 cells
    // along the longest side
    const double span = std::max({d.x_max - d.x_min, d.y_max - d.y_min, d.z_max - d.z_min, 1e-30});
    pgrid = std::ldexp(1.0, (int)std::ceil(std::log2(span / 1000.0)));
    d.NI = (int)((d.x_max - d.x_min) / d.dcell) + 1;
    d.NJ = (int)((d.y_max - d.y_min) / d.dcell) + 1;
    d.NK = (int)((d.z_max - d.z_min) / d.dcell) + 1;
    // periodic axes of a fixed domain: whole cells of width >= dcell tile the
    // period (the wrapped 27-cell stencil then covers kappa h), and the pos4
    // grid has 1024 cells per period (the image is an integer wrap)
    pmask = p.domain_fixed ? (p.periodic_mask & (p.dim == 2 ? 3 : 7)) : 0;
    int* N[3] = {&d.NI, &d.NJ, &d.NK};
    const double sp[3] = {d.x_max - d.x_min, d.y_max - d.y_min, d.z_max - d.z_min};
    for (int a = 0; a < 3; a++)
      if (pmask >> a & 1) {
        plen[a] = sp[a];
        *N[a] = std::max(1, (int)(sp[a] / d.dcell));
        pgrid = sp[a] / 1024.0;
      }
    if (p.flag_z_index == 0) {
      d.num_cells = d.NI * d.NJ * d.NK + 1;
    } else {
      int m = std::max(std::max(d.NI, d.NJ), d.NK);
      uint64_t last = p.dim == 2 ? (spread2(m - 1) | spread2(m - 1) << 1)
                                 : (spread3(m - 1) | spread3(m - 1) << 1 | spread3(m - 1) << 2);
      d.num_cells = (int)last + 2;
    }
  }

  template <typename T>
  int put(T* dst, const std::vector<T>& src) {
    if (src.empty()) return 0;
    HIP_TRY(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
  }

  int upload(const part1* P, const std::vector<int>& si, const std::vector<int>& di) {
    auto colD = [&](const std::vector<int>& ix, double part1::*f) {
      std::vector<double> o(ix.size());
      for (size_t k = 0; k < ix.size(); k++) o[k] = P[ix[k]].*f;
      return o;
    };
    auto colR = [&](const std::vector<int>& ix, double part1::*f) {
      std::vector<R> o(ix.size());
      for (size_t k = 0; k < ix.size(); k++) o[k] = (R)(P[ix[k]].*f);
      return o;
    };
    auto colI = [&](const std::vector<int>& ix, int part1::*f) {
      std::vector<int> o(ix.size());
      for (size_t k = 0; k < ix.size(); k++) o[k] = P[ix[k]].*f;
      return o;
    };
    TRY(put(sx.cur, colD(si, &part1::x)));
    TRY(put(sy.cur, colD(si, &part1::y)));
    TRY(put(sz.cur, colD(si, &part1::z)));
    TRY(put(sh.cur, colD(si, &part1::h)));
    TRY(put(spres.cur, colD(si, &part1::pres)));
    TRY(put(sux.cur, colR(si, &part1::ux)));
    TRY(put(suy.cur, colR(si, &part1::uy)));
    TRY(put(suz.cur, colR(si, &part1::uz)));
    TRY(put(sm.cur, colD(si, &part1::m)));
    TRY(put(srho.cur, colD(si, &part1::rho)));
    TRY(put(stemp.cur, colR(si, &part1::temp)));
    TRY(put(sflt.cur, colR(si, &part1::flt_s)));
    TRY(put(sdpor.cur, colR(si, &part1::dDEMpor)));
    TRY(put(sptype.cur, colI(si, &part1::p_type)));
    TRY(put(sbtype.cur, colI(si, &part1::buffer_type)));
    TRY(put(sitype.cur, colI(si, &part1::i_type)));
    TRY(put(dx_.cur, colD(di, &part1::x)));
    TRY(put(dy_.cur, colD(di, &part1::y)));
    TRY(put(dz_.cur, colD(di, &part1::z)));
    TRY(put(dux.cur, colD(di, &part1::ux)));
    TRY(put(duy.cur, colD(di, &part1::uy)));
    TRY(put(duz.cur, colD(di, &part1::uz)));
    TRY(put(dwx.cur, colD(di, &part1::wx)));
    TRY(put(dwy.cur, colD(di, &part1::wy)));
    TRY(put(dwz.cur, colD(di, &part1::wz)));
    TRY(put(dm.cur, colD(di, &part1::m)));
    TRY(put(dri.cur, colD(di, &part1::ri)));
    TRY(put(drad.cur, colD(di, &part1::rad)));
    TRY(put(dh.cur, colD(di, &part1::h)));
    TRY(put(drho.cur, colD(di, &part1::rho)));
    TRY(put(dtemp.cur, colD(di, &part1::temp)));
    TRY(put(dfdx.cur, colD(di, &part1::Fdx_df)));
    TRY(put(dfdy.cur, colD(di, &part1::Fdy_df)));
    TRY(put(dfdz.cur, colD(di, &part1::Fdz_df)));
    TRY(put(dqsd.cur, colD(di, &part1::Q_sd)));
    TRY(put(dptype.cur, colI(di, &part1::p_type)));
    TRY(put(dbtype.cur, colI(di, &part1::buffer_type)));
    TRY(put(ditype.cur, colI(di, &part1::i_type)));
    TRY(put(didx.cur, colI(di, &part1::dem_idx)));
    if (nd > 0) {
      std::vector<int> hi((size_t)kMaxHist * nd);
      std::vector<double> hd((size_t)kMaxHist * 3 * nd);
      for (int k = 0; k < kMaxHist; k++)
        for (int i = 0; i < nd; i++) {
          const part1& q = P[di[i]];
          hi[(size_t)k * nd + i] = (int)q.ct_idx[k];
          for (int c = 0; c < 3; c++) hd[((size_t)k * 3 + c) * nd + i] = q.del_s[k][c];
        }
      TRY(put(hist_idx[0], hi));
      TRY(put(hist_ds[0], hd));
    }
    return 0;
  }

  // --- NNPS ---------------------------------------------------------------
  int sort_species(int n, double* x, double* y, double* z, int* itype, SortScratch& sc, uint32_t** keys,
                   uint32_t** order, uint32_t* cs, uint32_t* ce) {
    DevParams P = devparams();
    TRY(nnps_keys(stream, P, n, x, y, z, itype, sc.keys_a, sc.vals_a));
    int nbits = std::max(1, ceil_log2((uint32_t)d.num_cells));
    TRY(radix_sort_pairs(stream, n, ((nbits + 7) / 8) * 8, sc, keys, order));
    TRY(cell_ranges(stream, n, d.num_cells, *keys, cs, ce));
    return 0;
  }

  // P3 reset (ISPH_Calc_seperate.cuh:102-105).  pgf and the XSPH velocities
  // live in P1 in the reference and are rewritten each step for the
  // particles their kernels compute; the others (walls, skipped types) keep
  // their input value, which is zero (the DEM drag reads pgf of walls).
  // PPE1..4 (P1) are written for PPE rows only; the rest keep zero.
  int reset_sph_accumulators() {
    std::vector<FillSpan> z;
    for (R* a : {sv.fx, sv.fy, sv.fz, sv.fmag, sv.fpx, sv.fpy, sv.fpz, sv.dtemp, sv.pgx, sv.pgy, sv.pgz, sv.xsx,
                 sv.xsy, sv.xsz, sv.xst})
      z.push_back({(uint32_t*)a, (size_t)ns * sizeof(R) / 4, 0u});
    z.push_back({(uint32_t*)sv.cm, (size_t)6 * ns * sizeof(R) / 4, 0u});
    for (double* a : {sv.ppe1, sv.ppe2, sv.ppe3, sv.ppe4}) z.push_back({(uint32_t*)a, (size_t)ns * 2, 0u});
    return fill_u32_many(stream, z.data(), (int)z.size());
  }

  int nnps() override {
    // a side-stream coarse build left by a PPE stage that returned early
    // finishes before the sort rewrites its inputs
    if (ppe_w.coarse) TRY(coarse_settle(stream, *ppe_w.coarse));
    const bool on = sph_on();
    DevParams P = devparams();
    const bool reuse = sph_sorted && p.solver_type == 1 && p.eulerian_sph;
    if (on && ns > 0 && reuse) {
      // positions unchanged since the last sort: same keys, identity order,
      // same ranges and neighbour list; only the per-step accumulators reset
      sph_order_identity = true;
      TRY(reset_sph_accumulators());
      ppe_assembled = false;
    } else if (on && ns > 0) {
      TimeScope t0(tm(), "nnps_sort_permute_sph");
      sph_order_identity = false;
      TRY(sort_species(ns, sx.cur, sy.cur, sz.cur, sitype.cur, ssort, &skeys, &sorder, scs, sce));
      std::vector<PermuteField> f;
      auto add = [&](auto& fld, int bytes) { f.push_back(PermuteField{fld.cur, fld.spare, bytes}); };
      add(sx, 8); add(sy, 8); add(sz, 8); add(sh, 8); add(spres, 8); add(sm, 8); add(srho, 8);
      if (p_hist > 0) add(sph1, 8);
      if (p_hist > 1) add(sph2, 8);
      const int rb = (int)sizeof(R);
      add(sux, rb); add(suy, rb); add(suz, rb); add(stemp, rb); add(sflt, rb);
      add(sdpor, rb);
      add(sptype, 4); add(sbtype, 4); add(sitype, 4);
      double** lag_dc[5] = {&sv.x0, &sv.y0, &sv.z0, &sv.rho0, &sv.rho_ref};
      R** lag_rc[4] = {&sv.ux0, &sv.uy0, &sv.uz0, &sv.temp0};
      if (wc_lag) {
        for (int k = 0; k < 5; k++) f.push_back(PermuteField{*lag_dc[k], lag_d[k], 8});
        for (int k = 0; k < 4; k++) f.push_back(PermuteField{*lag_rc[k], lag_r[k], rb});
      }
      TRY(permute_fields(stream, ns, sorder, f.data(), (int)f.size()));
      if (wc_lag) {
        for (int k = 0; k < 5; k++) std::swap(*lag_dc[k], lag_d[k]);
        for (int k = 0; k < 4; k++) std::swap(*lag_rc[k], lag_r[k]);
      }
      TRY(cell_offsets(stream, ns, skeys, d.num_cells, scoff));
      for (auto* fl : {&sx, &sy, &sz, &sh, &spres, &sm, &srho}) fl->swap();
      if (p_hist > 0) sph1.swap();
      if (p_hist > 1) sph2.swap();
      for (auto* fl : {&sux, &suy, &suz, &stemp, &sflt, &sdpor}) fl->swap();
      for (auto* fl : {&sptype, &sbtype, &sitype}) fl->swap();
      refresh_views();
      TRY(reset_sph_accumulators());
      // neighbour list with capacity check
      for (int attempt = 0; attempt < 2; attempt++) {
        TRY(fill_u32(stream, (uint32_t*)maxcnt, 0u, 1));
        CellView cv{scs, sce};
        {
          TimeScope t1(tm(), "nbr_build");
          TRY(build_sph_neighbors<R>(stream, P, sv, cv, maxcnt));
        }
        int mc = 0;
        HIP_TRY(hipMemcpyAsync(&mc, maxcnt, 4, hipMemcpyDeviceToHost, stream));
        HIP_TRY(hipStreamSynchronize(stream));
        if (mc <= sv.nbr_cap) break;
        if (attempt == 1) {
          set_error("neighbour list overflow: %d > %d", mc, sv.nbr_cap);
          return SOPHIA_E_CAPACITY;
        }
        TRY(alloc_nbr(((mc + 15) / 16) * 16));
        TRY(mlp_alloc(mlp, ns, sv.nbr_cap));
      }
      // surface labels follow from positions and types alone
      // (ISPH_Calc_seperate.cuh:289-296)
      if (p.free_surface > 0) {
        TimeScope t1(tm(), "surface");
        TRY(sph_surface<R>(stream, P, sv));
      }
      ppe_assembled = false;
      sph_sorted = true;
    }
    TRY(nnps_dem());
    nnps_done = true;
    after_corrector = false;
    return 0;
  }

  // sort the spheres, permute their state and reset their per-step
  // accumulators (the DEM half of NNPS, every step: spheres move)
  int nnps_dem() {
    if (nd > 0) {
      TimeScope t2(tm(), "nnps_sort_permute_dem");
      TRY(sort_species(nd, dx_.cur, dy_.cur, dz_.cur, ditype.cur, dsort, &dkeys, &dorder, dcs, dce));
      std::vector<PermuteField> f;
      auto add = [&](auto& fld, int bytes) { f.push_back(PermuteField{fld.cur, fld.spare, bytes}); };
      for (auto* fl : {&dx_, &dy_, &dz_, &dux, &duy, &duz, &dwx, &dwy, &dwz, &dm, &dri, &drad, &dh, &drho, &dtemp,
                       &dfdx, &dfdy, &dfdz, &dqsd})
        add(*fl, 8);
      for (auto* fl : {&dptype, &dbtype, &ditype, &didx}) add(*fl, 4);
      if (wc_dem)
        for (auto& fl : dws) add(fl, 8);
      for (int k = 0; k < kMaxHist; k++) {
        f.push_back(PermuteField{hist_idx[0] + (size_t)k * nd, hist_idx[1] + (size_t)k * nd, 4});
        for (int c = 0; c < 3; c++)
          f.push_back(PermuteField{hist_ds[0] + ((size_t)k * 3 + c) * nd, hist_ds[1] + ((size_t)k * 3 + c) * nd, 8});
      }
      TRY(permute_fields(stream, nd, dorder, f.data(), (int)f.size()));
      for (auto* fl : {&dx_, &dy_, &dz_, &dux, &duy, &duz, &dwx, &dwy, &dwz, &dm, &dri, &drad, &dh, &drho, &dtemp,
                       &dfdx, &dfdy, &dfdz, &dqsd})
        fl->swap();
      for (auto* fl : {&dptype, &dbtype, &ditype, &didx}) fl->swap();
      if (wc_dem)
        for (auto& fl : dws) fl.swap();
      std::swap(hist_idx[0], hist_idx[1]);
      std::swap(hist_ds[0], hist_ds[1]);
      refresh_views();
      {
        std::vector<FillSpan> z;
        for (double* a : {dv.fx, dv.fy, dv.fz, dv.tqx, dv.tqy, dv.tqz, dv.dtemp})
          z.push_back({(uint32_t*)a, (size_t)nd * 2, 0u});
        TRY(fill_u32_many(stream, z.data(), (int)z.size()));
      }
    } else {
      TRY(fill_u32(stream, dcs, kEmpty, (size_t)d.num_cells));
    }
    return 0;
  }

  int need_nnps() const {
    if (!nnps_done) {
      set_error("stage called before sophia_nnps");
      return SOPHIA_E_STATE;
    }
    if (p.prop_table && !has_table) {
      set_error("property_table is on and no table was installed (sophia_set_table)");
      return SOPHIA_E_STATE;
    }
    return 0;
  }

  // --- stages ---------------------------------------------------------------
  int prep() override {
    TRY(need_nnps());
    DevParams P = devparams();
    CellView sc{scs, sce}, dc{dcs, dce};
    bool on = sph_on();
    if (p.dim != 3) return 0;
    if (on) TRY(sph_mass_init<R>(stream, P, sv));
    if (on && P.prop_table) TRY(sph_props<R>(stream, P, sv));
    // The spheres' porosity reads SPH positions, m, rho and p_type only, which
    // KGC and the SPH porosity leave alone, and writes sphere fields only: it
    // runs beside them on the second stream and joins before k_prep_prep
    // rewrites m and rho.
    const bool side = nd > 0;
    if (side) {
      HIP_TRY(hipEventRecord(ev_dem_in, stream));
      HIP_TRY(hipStreamWaitEvent(dem_stream, ev_dem_in, 0));
      Timers* T = tm();
      hipEvent_t ta = nullptr, tb = nullptr;
      if (T && T->on) {
        ta = T->get();
        tb = T->get();
        HIP_TRY(hipEventRecord(ta, dem_stream));
      }
      TRY(dem_porosity<R>(dem_stream, P, dv, sv, sc, dc));
      if (T && T->on) {
        HIP_TRY(hipEventRecord(tb, dem_stream));
        T->add_span("porosity_dem", ta, tb);
      }
      HIP_TRY(hipEventRecord(ev_dem_out, dem_stream));
    }
    if (on) {
      TimeScope t(tm(), "kgc_porosity_sph");
      TRY(sph_kgc_porosity<R>(stream, P, sv, dv, dc, p.kgc_solve > 0));
    }
    if (side) HIP_TRY(hipStreamWaitEvent(stream, ev_dem_out, 0));
    if (on && P.heat_flux != 0.0) {
      // the flux walls' fluid sums read m and rho after the density reset and
      // write flux_s only, which the prep sums do not read: they run beside
      // the prep sums on the second stream
      TimeScope t(tm(), "prep");
      TRY(sph_prep_prep<R>(stream, P, sv));
      HIP_TRY(hipEventRecord(ev_dem_in, stream));
      HIP_TRY(hipStreamWaitEvent(dem_stream, ev_dem_in, 0));
      TRY(sph_flux_walls<R>(dem_stream, P, sv));
      HIP_TRY(hipEventRecord(ev_dem_out, dem_stream));
      const int rc = sph_prep_sums<R>(stream, P, sv);
      HIP_TRY(hipStreamWaitEvent(stream, ev_dem_out, 0));
      TRY(rc);
    } else if (on) {
      TimeScope t(tm(), "prep");
      TRY(sph_prep<R>(stream, P, sv));
    }
    return 0;
  }

  int interaction() override {
    TRY(need_nnps());
    if (!sph_on() || p.dim != 3) return 0;
    TimeScope t(tm(), "advection");
    return sph_advection<R>(stream, devparams(), sv);
  }

  // The sphere contact reads and writes sphere state only and advection SPH
  // state only: the contact runs on the second stream beside advection, and
  // the main stream waits for it right after advection is queued (also when
  // queueing advection fails), so no later call sees it in flight.
  int interaction_both() override {
    TRY(need_nnps());
    if (p.dim != 3 || nd == 0) {
      TRY(interaction());
      return dem_interaction();
    }
    CellView dc{dcs, dce};
    HIP_TRY(hipEventRecord(ev_dem_in, stream));
    HIP_TRY(hipStreamWaitEvent(dem_stream, ev_dem_in, 0));
    Timers* T = tm();
    hipEvent_t ta = nullptr, tb = nullptr;
    if (T && T->on) {
      ta = T->get();
      tb = T->get();
      HIP_TRY(hipEventRecord(ta, dem_stream));
    }
    TRY(dem_contact(dem_stream, devparams(), dv, dc));
    if (T && T->on) {
      HIP_TRY(hipEventRecord(tb, dem_stream));
      T->add_span("dem_contact", ta, tb);
    }
    HIP_TRY(hipEventRecord(ev_dem_out, dem_stream));
    std::swap(hist_idx[0], hist_idx[1]);
    std::swap(hist_ds[0], hist_ds[1]);
    refresh_views();
    const int rc = interaction();
    HIP_TRY(hipStreamWaitEvent(stream, ev_dem_out, 0));
    return rc;
  }

  int dem_interaction() override {
    TRY(need_nnps());
    if (p.dim != 3 || nd == 0) return 0;
    CellView dc{dcs, dce};
    {
      TimeScope t(tm(), "dem_contact");
      TRY(dem_contact(stream, devparams(), dv, dc));
    }
    std::swap(hist_idx[0], hist_idx[1]);
    std::swap(hist_ds[0], hist_ds[1]);
    refresh_views();
    return 0;
  }

  int coupling() override {
    TRY(need_nnps());
    if (p.dim != 3) return 0;
    DevParams P = devparams();
    CellView sc{scs, sce}, dc{dcs, dce};
    {
      TimeScope t(tm(), "dem_coupling");
      TRY(dem_coupling<R>(stream, P, dv, sv, sc));
    }
    if (sph_on()) {
      TimeScope t(tm(), "sph_reaction");
      TRY(sph_dem_reaction<R>(stream, P, sv, dv, dc));
    }
    return 0;
  }

  // DEM walls after the time update (ISPH_Calc_seperate.cuh:439-441)
  int apply_dem_walls(const DevParams& P) {
    if (!p.dem_walls || nd == 0) return 0;
    DemWallParams W;
    W.mask = p.dem_walls;
    std::memcpy(W.cyl, p.wall_cyl, sizeof(W.cyl));
    std::memcpy(W.box, p.wall_box, sizeof(W.box));
    std::memcpy(W.cone, p.wall_cone, sizeof(W.cone));
    W.rest_n = p.wall_rest[0];
    W.rest_s = p.wall_rest[1];
    W.mu_s = p.wall_mu[0];
    W.mu_rs = p.wall_mu[1];
    TimeScope t(tm(), "dem_walls");
    return dem_walls(stream, P, dv, W);
  }

  int time_stage(int phase) override {
    TRY(need_nnps());
    DevParams P = devparams();
    if (phase == SOPHIA_PREDICT) {
      if (p.time_type != 1) return 0;
      TimeScope t(tm(), "predictor");
      if (sph_on()) TRY(sph_predictor<R>(stream, P, sv));
      TRY(dem_predictor(stream, P, dv));
    } else {
      TimeScope t(tm(), "corrector");
      if (sph_on()) {
        TRY(sph_corrector<R>(stream, P, sv));
        TRY(refresh_pos4<R>(stream, sv));  // the position records follow every move
      }
      TRY(dem_corrector(stream, P, dv));
      if (sph_on() && p.open_boundary == 1) {
        // inlet/outlet buffer state from the interior (ISPH_Calc_seperate.cuh:410-426)
        TimeScope t(tm(), "open_boundary");
        OpenBoundary ob;
        std::memcpy(ob.L, p.ob_band, sizeof(ob.L));
        ob.inlet_u = p.ob_inlet_u;
        ob.inlet_temp = p.ob_inlet_temp;
        TRY(sph_open_boundary<R>(stream, P, sv, CellView{scs, sce}, ob));
      }
      TRY(apply_dem_walls(P));
      after_corrector = true;
    }
    return 0;
  }

  int ppe(const sophia_ppe_opts* o, sophia_ppe_stats* st) override {
    TRY(need_nnps());
    sophia_ppe_stats local{};
    if (!st) st = &local;
    std::memset(st, 0, sizeof(*st));
    if (!sph_on() || p.dim != 3 || ns == 0) return 0;
    sophia_ppe_opts def{};
    def.method = SOPHIA_PPE_JACOBI;
    if (!o) o = &def;
    if (o->method == SOPHIA_PPE_BICGSTAB) {
      bool mixed_ws = ppe_w.mixed;
      bool want_mixed = o->precision == SOPHIA_PPE_MIXED_IR;
      if (mixed_ws != want_mixed) {
        ppe_workspace_free(ppe_w);
        TRY(ppe_workspace_alloc(ppe_w, ns, sv.nbr_cap, want_mixed));
      }
    }
    DevParams P = devparams();
    CellView sc{scs, sce};
    HIP_TRY(hipEventRecord(ev0, stream));
    if (count == 0) {
      TimeScope t(tm(), "neumann");
      TRY(sph_neumann<R>(stream, P, sv, sc));
    }
    {
      TimeScope t(tm(), "ppe_assemble");
      TRY(ppe_assemble<R>(stream, P, sv, ppe_w, skeys, scoff));
      // periodic box without walls: a consistent right-hand side for the
      // singular operator (the oracle's project_mean)
      if (o->zero_mean && o->method == SOPHIA_PPE_BICGSTAB) TRY(ppe_project_mean(stream, ppe_w, ppe_w.rhs));
    }
    ppe_assembled = true;
    if (o->method == SOPHIA_PPE_BICGSTAB && o->precond == SOPHIA_PRECOND_TWO_LEVEL && o->zero_mean == 0 &&
        !(o->guess == SOPHIA_GUESS_LEARNED && mlp.ready)) {
      // the coarse matrix of the two-level preconditioner builds on a side
      // stream while the guess and its residual are computed on this one (not
      // beside the learned forward, which keeps every CU busy by itself)
      if (!ppe_w.coarse) ppe_w.coarse = new CoarseWork();
      TRY(coarse_rows_early(stream, *ppe_w.coarse, ppe_w, tm()));
    }
    if (o->method == SOPHIA_PPE_JACOBI) {
      TRY(rotate_pressure_history());
      TRY(ppe_jacobi<R>(stream, sv, ppe_w));
      st->iterations = 1;
      st->converged = 1;
    } else {
      const double* guess = nullptr;
      st->guess_used = o->guess == SOPHIA_GUESS_ZERO ? SOPHIA_GUESS_ZERO : SOPHIA_GUESS_PREVIOUS;
      const bool learned = o->guess == SOPHIA_GUESS_LEARNED && mlp.ready;
      const bool want_mlp = (learned || o->train) && mlp.ready;
      // the extrapolated guess, also the base of the learned correction
      if (o->guess == SOPHIA_GUESS_EXTRAPOLATE || want_mlp) {
        TRY(ppe_extrapolate(stream, ppe_w, sv.pres, sph1.cur, sph2.cur, p_hist, pbase));
        if (o->guess == SOPHIA_GUESS_EXTRAPOLATE || learned) {
          guess = pbase;
          st->guess_used = SOPHIA_GUESS_EXTRAPOLATE;
        }
      }
      double scale = 0, rr_base = 0, known_rr = -1.0;
      if (want_mlp) {
        TRY(mlp_prepare(pbase, &scale, &rr_base));
        HIP_TRY(hipMemcpyAsync(mlp.pprev, pbase, (size_t)ns * 8, hipMemcpyDeviceToDevice, stream));
      }
      if (learned) {
        // learned correction with residual-based fallback to the extrapolation
        // (the extrapolation's residual is the network's input r0)
        {
          TimeScope t(tm(), "mlp_forward");
          TRY(mlp_forward<R>(stream, P, sv, mlp, ppe_w.row, pbase, scale, mlp.guess));
        }
        double r_l = 0, rr_l = 0;
        TRY(ppe_residual_of(stream, ppe_w, mlp.guess, &r_l, &rr_l));
        st->r0_rel_learned = r_l;
        st->r0_rel_base = rr_l > 0 ? r_l * std::sqrt(rr_base / rr_l) : 0.0;
        if (rr_l < rr_base) {
          guess = mlp.guess;
          known_rr = rr_l;  // w.r64 holds its residual
          st->guess_used = SOPHIA_GUESS_LEARNED;
        } else {
          // back to the extrapolation, whose residual is the network's input
          HIP_TRY(hipMemcpyAsync(ppe_w.r64, mlp.r0, (size_t)ns * 8, hipMemcpyDeviceToDevice, stream));
          known_rr = rr_base;
        }
      }
      TRY(rotate_pressure_history());
      sophia_ppe_opts oo = *o;
      if (oo.max_iter <= 0) oo.max_iter = 10000;
      if (oo.rtol <= 0) oo.rtol = 1e-8;
      TRY(ppe_solve(stream, ppe_w, sv.pres, oo, *st, guess, tm(), known_rr));
      if (o->train && mlp.ready) {
        // one Adam step on (features before the solve, converged pressure)
        TimeScope t(tm(), "mlp_train");
        double loss = 0;
        TRY(mlp_backward<R>(stream, P, sv, mlp, ppe_w.row, mlp.pprev, sv.pres, scale, n_rows, true,
                            o->lr > 0 ? o->lr : 1e-3, &loss));
        st->loss = loss;
      }
    }
    {
      TimeScope t(tm(), "neumann");
      TRY(sph_neumann<R>(stream, P, sv, sc));
    }
    {
      TimeScope t(tm(), "pressure_force");
      TRY(sph_pressure_force<R>(stream, P, sv));
    }
    // the side-stream coarse build of this operator, if any, completes before
    // anything later on the stream rewrites its inputs
    if (ppe_w.coarse) TRY(coarse_settle(stream, *ppe_w.coarse));
    HIP_TRY(hipEventRecord(ev1, stream));
    HIP_TRY(hipEventSynchronize(ev1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    st->ms = ms;
    st->n_rows = n_rows;
    st->nnz = (int32_t)std::min<int64_t>(ppe_w.nnz, INT32_MAX);
    return 0;
  }

  int xsph() override {
    TRY(need_nnps());
    if (!p.xsph_solve || !sph_on() || p.dim != 3) return 0;
    TimeScope t(tm(), "xsph");
    return sph_xsph<R>(stream, devparams(), sv);
  }

  // --- inspection -------------------------------------------------------------
  template <typename T>
  int get(std::vector<T>& h, const T* dptr, size_t n) const {
    h.resize(n);
    if (n == 0) return 0;
    HIP_TRY(hipMemcpyAsync(h.data(), dptr, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
  }

  int get_nnps(int species, int32_t* keys, int32_t* order, uint32_t* cs, uint32_t* ce, int32_t* nc) const override {
    bool S = species == SOPHIA_SPH;
    int n = S ? ns : nd;
    if (!nnps_done) {
      set_error("get_nnps before sophia_nnps");
      return SOPHIA_E_STATE;
    }
    std::vector<uint32_t> t;
    if (keys && n) {
      TRY(get(t, S ? skeys : dkeys, n));
      std::memcpy(keys, t.data(), n * 4);
    }
    if (order && n) {
      if (S && sph_order_identity) {
        for (int i = 0; i < n; i++) order[i] = i;  // the sort of sorted keys that was skipped
      } else {
        TRY(get(t, S ? sorder : dorder, n));
        std::memcpy(order, t.data(), n * 4);
      }
    }
    if (cs) {
      TRY(get(t, S ? scs : dcs, d.num_cells));
      std::memcpy(cs, t.data(), (size_t)d.num_cells * 4);
    }
    if (ce) {
      TRY(get(t, S ? sce : dce, d.num_cells));
      std::memcpy(ce, t.data(), (size_t)d.num_cells * 4);
    }
    if (nc && S && n) {
      std::vector<int> c;
      TRY(get(c, (const int*)sv.nbr_cnt, n));
      std::memcpy(nc, c.data(), n * 4);
    }
    return 0;
  }

  int get_particles(int species, part1* p1, part2* p2, part3* p3) const override {
    if (species == SOPHIA_SPH) {
      int n = ns;
      std::vector<double> X, Y, Z, H, PR, P1v, P2v, P3v, P4v;
      std::vector<R> a[40];
      std::vector<int> pt, bt, it;
      TRY(get(X, sv.x, n)); TRY(get(Y, sv.y, n)); TRY(get(Z, sv.z, n)); TRY(get(H, sv.h, n));
      TRY(get(PR, sv.pres, n));
      TRY(get(P1v, sv.ppe1, n)); TRY(get(P2v, sv.ppe2, n)); TRY(get(P3v, sv.ppe3, n)); TRY(get(P4v, sv.ppe4, n));
      TRY(get(pt, sv.ptype, n)); TRY(get(bt, sv.btype, n)); TRY(get(it, sv.itype, n));
      std::vector<double> Mv, RHv, POv;
      TRY(get(Mv, sv.m, n)); TRY(get(RHv, sv.rho, n)); TRY(get(POv, sv.por, n));
      const R* src[] = {sv.ux, sv.uy, sv.uz, sv.ux, sv.ux, sv.temp, sv.flt_s, sv.dpor, sv.ux, sv.dpor_prev,
                        sv.ux, sv.ux, sv.grx, sv.gry, sv.grz, sv.fx, sv.fy, sv.fz, sv.fmag, sv.fpx,
                        sv.fpy, sv.fpz, sv.pgx, sv.pgy, sv.pgz, sv.dtemp, sv.ux0, sv.uy0, sv.uz0, sv.temp0,
                        sv.xsx, sv.xsy, sv.xsz, sv.xst, sv.drho};
      std::vector<double> RRv, R0v, VOLv, VOL0v, X0v, Y0v, Z0v;
      TRY(get(RRv, sv.rho_ref, n)); TRY(get(R0v, sv.rho0, n));
      TRY(get(VOLv, sv.vol, n)); TRY(get(VOL0v, sv.vol0, n));
      TRY(get(X0v, sv.x0, n)); TRY(get(Y0v, sv.y0, n)); TRY(get(Z0v, sv.z0, n));
      const int NF = sizeof(src) / sizeof(src[0]);
      for (int k = 0; k < NF; k++) TRY(get(a[k], src[k], n));
      std::vector<R> cm;
      TRY(get(cm, sv.cm, (size_t)6 * n));
      std::vector<int> rows, surf;
      if (ppe_assembled) TRY(get(rows, (const int*)ppe_w.row, n));
      TRY(get(surf, (const int*)sv.surf, n));
      for (int i = 0; i < n; i++) {
        auto A = [&](int k) { return (double)a[k][i]; };
        if (p1) {
          part1& q = p1[i];
          std::memset(&q, 0, sizeof(q));
          q.i_type = it[i]; q.buffer_type = bt[i]; q.p_type = pt[i];
          q.x = X[i]; q.y = Y[i]; q.z = Z[i]; q.h = H[i]; q.pres = PR[i];
          q.ux = A(0); q.uy = A(1); q.uz = A(2); q.m = Mv[i]; q.rho = RHv[i]; q.temp = A(5); q.flt_s = A(6);
          q.dDEMpor = A(7); q.DEMpor = POv[i]; q.dDEMpor_prev = A(9);
          q.grad_rhox = A(12); q.grad_rhoy = A(13); q.grad_rhoz = A(14);
          q.pgf_x = A(22); q.pgf_y = A(23); q.pgf_z = A(24);
          q.XSPH_ux = A(30); q.XSPH_uy = A(31); q.XSPH_uz = A(32); q.XSPH_temp = A(33);
          // 3D WCSPH: KERNEL_set_alpha makes every SPH particle Eulerian
          const bool eul = p.solver_type == 0 ? p.dim == 3 : p.eulerian_sph != 0;
          double e = (pt[i] <= 1000 && eul) ? 0.0 : 1.0;
          q.elix = q.eliy = q.eliz = e;
          q.vol = VOLv[i]; q.vol0 = VOL0v[i];
          bool row = !rows.empty() && rows[i];
          q.PPE1 = row ? P1v[i] : 0; q.PPE2 = row ? P2v[i] : 0; q.PPE3 = row ? P3v[i] : 0; q.PPE4 = row ? P4v[i] : 0;
        }
        if (p2) {
          part2& q = p2[i];
          std::memset(&q, 0, sizeof(q));
          q.rho_ref = RRv[i]; q.rho0 = R0v[i];
          q.ux0 = A(26); q.uy0 = A(27); q.uz0 = A(28); q.temp0 = A(29);
          if (p.solver_type == 0) {
            q.x0 = X0v[i]; q.y0 = Y0v[i]; q.z0 = Z0v[i];
          } else {
            q.x0 = X[i]; q.y0 = Y[i]; q.z0 = Z[i];
          }
        }
        if (p3) {
          part3& q = p3[i];
          std::memset(&q, 0, sizeof(q));
          q.ftotalx = A(15); q.ftotaly = A(16); q.ftotalz = A(17); q.ftotal = A(18);
          q.fpx = A(19); q.fpy = A(20); q.fpz = A(21); q.dtemp = A(25); q.drho = A(34);
          q.lbl_surf = surf[i] ? 1.0 : 0.0;
          double c[6];
          for (int k = 0; k < 6; k++) c[k] = (double)cm[(size_t)k * n + i];
          q.Cm[0][0] = c[0]; q.Cm[0][1] = c[1]; q.Cm[0][2] = c[2];
          q.Cm[1][0] = c[1]; q.Cm[1][1] = c[3]; q.Cm[1][2] = c[4];
          q.Cm[2][0] = c[2]; q.Cm[2][1] = c[4]; q.Cm[2][2] = c[5];
        }
      }
      return 0;
    }
    int n = nd;
    const double* src[] = {dv.x, dv.y, dv.z, dv.ux, dv.uy, dv.uz, dv.wx, dv.wy, dv.wz, dv.m, dv.ri, dv.rad,
                           dv.h, dv.rho, dv.temp, dv.fdx, dv.fdy, dv.fdz, dv.qsd, dv.flt_s, dv.flt_sd, dv.flt_sd2,
                           dv.por, dv.fax, dv.fay, dv.faz, dv.fbx, dv.fby, dv.fbz, dv.fx, dv.fy, dv.fz,
                           dv.tqx, dv.tqy, dv.tqz, dv.ux0, dv.uy0, dv.uz0, dv.test6,
                           dv.hrad, dv.hcon, dv.dtemp, dv.temp0};
    const int NF = sizeof(src) / sizeof(src[0]);
    std::vector<double> a[NF];
    for (int k = 0; k < NF; k++) TRY(get(a[k], src[k], n));
    std::vector<int> pt, bt, it, ix, hi;
    std::vector<double> hd, ws[6];
    TRY(get(pt, dv.ptype, n)); TRY(get(bt, dv.btype, n)); TRY(get(it, dv.itype, n)); TRY(get(ix, dv.dem_idx, n));
    if (wc_dem) {
      const double* w0[6] = {dv.x0, dv.y0, dv.z0, dv.wx0, dv.wy0, dv.wz0};
      for (int k = 0; k < 6; k++) TRY(get(ws[k], w0[k], n));
    }
    TRY(get(hi, (const int*)hist_idx[0], (size_t)kMaxHist * n));
    TRY(get(hd, (const double*)hist_ds[0], (size_t)kMaxHist * 3 * n));
    for (int i = 0; i < n; i++) {
      auto A = [&](int k) { return a[k][i]; };
      if (p1) {
        part1& q = p1[i];
        std::memset(&q, 0, sizeof(q));
        q.i_type = it[i]; q.buffer_type = bt[i]; q.p_type = pt[i]; q.dem_idx = ix[i];
        q.x = A(0); q.y = A(1); q.z = A(2); q.ux = A(3); q.uy = A(4); q.uz = A(5);
        q.wx = A(6); q.wy = A(7); q.wz = A(8); q.m = A(9); q.ri = A(10); q.rad = A(11); q.h = A(12);
        q.rho = A(13); q.temp = A(14); q.Fdx_df = A(15); q.Fdy_df = A(16); q.Fdz_df = A(17); q.Q_sd = A(18);
        q.flt_s = A(19); q.flt_sd = A(20); q.flt_sd_2 = A(21); q.DEMpor = A(22);
        q.Fdx_da = A(23); q.Fdy_da = A(24); q.Fdz_da = A(25); q.Fdx_b = A(26); q.Fdy_b = A(27); q.Fdz_b = A(28);
        q.test6 = A(38);
        q.test1 = A(39); q.test2 = A(40);
        q.elix = q.eliy = q.eliz = 1.0;
        for (int k = 0; k < kMaxHist; k++) {
          q.ct_idx[k] = hi[(size_t)k * n + i];
          for (int c = 0; c < 3; c++) q.del_s[k][c] = hd[((size_t)k * 3 + c) * n + i];
        }
      }
      if (p2) {
        part2& q = p2[i];
        std::memset(&q, 0, sizeof(q));
        q.ux0 = A(35); q.uy0 = A(36); q.uz0 = A(37);
        q.temp0 = A(42);
        if (wc_dem) {
          q.x0 = ws[0][i]; q.y0 = ws[1][i]; q.z0 = ws[2][i];
          q.wx0 = ws[3][i]; q.wy0 = ws[4][i]; q.wz0 = ws[5][i];
        }
      }
      if (p3) {
        part3& q = p3[i];
        std::memset(&q, 0, sizeof(q));
        q.ftotalx = A(29); q.ftotaly = A(30); q.ftotalz = A(31);
        q.torqx = A(32); q.torqy = A(33); q.torqz = A(34);
        q.dtemp = A(41);
      }
    }
    return 0;
  }

  int set_pressure(const double* pr, int n) override {
    if (n != ns) {
      set_error("set_pressure: n=%d but %d SPH particles", n, ns);
      return SOPHIA_E_ARG;
    }
    HIP_TRY(hipMemcpyAsync(sv.pres, pr, (size_t)n * 8, hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
  }

  int get_ppe_system(double* diag, double* rhs, int32_t* row) const override {
    if (!ppe_assembled) {
      set_error("PPE not assembled yet");
      return SOPHIA_E_STATE;
    }
    std::vector<double> t;
    if (diag) { TRY(get(t, (const double*)ppe_w.diag, ns)); std::memcpy(diag, t.data(), (size_t)ns * 8); }
    if (rhs) { TRY(get(t, (const double*)ppe_w.rhs, ns)); std::memcpy(rhs, t.data(), (size_t)ns * 8); }
    if (row) {
      std::vector<int> r;
      TRY(get(r, (const int*)ppe_w.row, ns));
      std::memcpy(row, r.data(), (size_t)ns * 4);
    }
    return 0;
  }

  // WCSPH stages in SOPHIA_single's order (WCSPH_Calc.cuh:4-340); see abi.h
  int wcsph_stage(int stage) override {
    if (p.solver_type != 0) {
      set_error("WCSPH stage on a context with solver_type %d", p.solver_type);
      return SOPHIA_E_STATE;
    }
    if (p.rho_type != 1) {
      set_error("WCSPH: supported set is continuity density (rho_type %d)", p.rho_type);
      return SOPHIA_E_ARG;
    }
    if (ns == 0) return 0;
    if (p.dim == 2) return wcsph2d_stage(stage);
    DevParams P = devparams();
    CellView sc{scs, sce};
    switch (stage) {
      case 0:
        // spheres: half step from last step's accelerations, then a fresh
        // sort every step (the reference's WCSPH driver sorts once, which
        // only its Eulerian fluid allows)
        if (nd > 0 && p.time_type == 1) {
          TimeScope t(tm(), "predictor");
          TRY(wc_dem_predictor(stream, P, dv));
        }
        if (count == 0) {
          // NNPS once (Eulerian); P3 starts at zero, so the predictor may
          // follow the sort: it is per particle
          TRY(nnps());
          if (p.time_type == 1) {
            TRY(wc_predictor<R>(stream, P, sv));
            TRY(refresh_pos4<R>(stream, sv));
          }
        } else {
          if (p.time_type == 1) {
            TRY(wc_predictor<R>(stream, P, sv));
            TRY(refresh_pos4<R>(stream, sv));
          }
          TRY(wc_reset_p3<R>(stream, sv));
          TRY(nnps_dem());
        }
        after_corrector = false;
        break;
      case 1:
        TRY(wc_volume_mass<R>(stream, P, sv, sc, (int)std::min<int64_t>(count, INT32_MAX)));
        break;
      case 2:
        if (p.kgc_solve > 0 || p.delSPH_solve == 2) {
          TimeScope t(tm(), "kgc");
          TRY(sph_kgc<R>(stream, P, sv));
        }
        {
          TimeScope t(tm(), "porosity_sph");
          CellView dc{dcs, dce};
          TRY(sph_porosity<R>(stream, P, sv, dv, dc));
        }
        if (nd > 0) {
          // KERNEL_clc_prep3D_coupling (functions_PREP2.cuh:3048), sphere side
          TimeScope t(tm(), "porosity_dem");
          TRY(dem_porosity<R>(stream, P, dv, sv, CellView{scs, sce}, CellView{dcs, dce}));
        }
        {
          TimeScope t(tm(), "prep");
          TRY(sph_prep_sums<R>(stream, P, sv));
        }
        break;
      case 3: {
        TimeScope t(tm(), "continuity");
        TRY(wc_continuity<R>(stream, P, sv));
        break;
      }
      case 4:
        TRY(wc_eos<R>(stream, P, sv));
        break;
      case 5: {
        TimeScope t(tm(), "neumann");
        TRY(sph_neumann<R>(stream, P, sv, sc));
        break;
      }
      case 6: {
        TimeScope t(tm(), "wc_interaction");
        if (P.prop_table) TRY(sph_props<R>(stream, P, sv));
        TRY(wc_interaction<R>(stream, P, sv));
        if (nd > 0) {
          // Hertz-Mindlin contact (commented out in the reference's WCSPH
          // driver, WCSPH_Calc.cuh:206), then KERNEL_DEM_coupling3D and
          // KERNEL_SPH_coupling3D (:225, 236)
          TRY(dem_interaction());
          TRY(dem_coupling<R>(stream, P, dv, sv, CellView{scs, sce}));
          TRY(sph_dem_reaction<R>(stream, P, sv, dv, CellView{dcs, dce}));
        }
        break;
      }
      case 7:
        TRY(wc_corrector<R>(stream, P, sv));
        TRY(refresh_pos4<R>(stream, sv));
        if (nd > 0) {
          TRY(wc_dem_corrector(stream, P, dv));
          TRY(apply_dem_walls(P));
        }
        if (p.open_boundary == 1) {
          // inlet/outlet buffer state from the interior; the reference's
          // WCSPH driver computes it (WCSPH_Calc.cuh:282) but never applies
          // it (no _calc call), so here it is applied as in the ISPH step
          TimeScope t(tm(), "open_boundary");
          OpenBoundary ob;
          std::memcpy(ob.L, p.ob_band, sizeof(ob.L));
          ob.inlet_u = p.ob_inlet_u;
          ob.inlet_temp = p.ob_inlet_temp;
          TRY(sph_open_boundary<R>(stream, P, sv, CellView{scs, sce}, ob));
        }
        after_corrector = true;
        break;
      case 8:
        if (p.xsph_solve) {
          TimeScope t(tm(), "xsph");
          TRY(sph_xsph_only<R>(stream, P, sv));
        }
        time += p.dt;
        count += 1;
        break;
      default:
        set_error("WCSPH stage %d out of range 0..8", stage);
        return SOPHIA_E_ARG;
    }
    return 0;
  }

  // 2D WCSPH: SOPHIA_multi's stage list (WCSPH_Calc.cuh:368-563) in the
  // numbering of the 3D stages; 1 and 5 have no 2D work
  int wcsph2d_stage(int stage) {
    DevParams P = devparams();
    switch (stage) {
      case 0:
        // half step, then a fresh sort: the particles move
        if (p.time_type == 1) {
          TimeScope t(tm(), "predictor");
          TRY(wc_predictor<R>(stream, P, sv));
        }
        TRY(nnps());
        TRY(wc_reset_p3<R>(stream, sv));
        after_corrector = false;
        break;
      case 1:
        break;
      case 2:
        if (p.kgc_solve > 0 || p.delSPH_solve == 2) {
          TimeScope t(tm(), "kgc");
          TRY(wc2_kgc<R>(stream, P, sv));
        }
        {
          TimeScope t(tm(), "prep");
          TRY(wc2_prep<R>(stream, P, sv));
        }
        if (p.noslip_bc == 1) {
          TimeScope t(tm(), "boundary");
          TRY(wc2_boundary<R>(stream, P, sv));
        }
        break;
      case 3: {
        TimeScope t(tm(), "continuity");
        TRY(wc2_continuity<R>(stream, P, sv));
        break;
      }
      case 4:
        TRY(wc_eos<R>(stream, P, sv));
        break;
      case 5:
        break;
      case 6: {
        TimeScope t(tm(), "wc_interaction");
        if (P.prop_table) TRY(sph_props<R>(stream, P, sv));
        TRY(wc2_interaction<R>(stream, P, sv));
        break;
      }
      case 7:
        TRY(wc2_time_update<R>(stream, P, sv));
        if (!p.xsph_solve) TRY(refresh_pos4<R>(stream, sv));
        after_corrector = true;
        break;
      case 8:
        if (p.xsph_solve) {
          TimeScope t(tm(), "xsph");
          TRY(wc2_xsph<R>(stream, P, sv));
          TRY(refresh_pos4<R>(stream, sv));
        }
        time += p.dt;
        count += 1;
        break;
      default:
        set_error("WCSPH stage %d out of range 0..8", stage);
        return SOPHIA_E_ARG;
    }
    return 0;
  }

  int set_table(const sophia_table& t) override {
    if (t.n_tables < 0 || t.n_tables > SOPHIA_TABLE_TYPES || t.n_values < 0 || t.n_values > SOPHIA_TABLE_VALUES) {
      set_error("set_table: %d blocks, %d values (at most %d, %d)", t.n_tables, t.n_values, SOPHIA_TABLE_TYPES,
                SOPHIA_TABLE_VALUES);
      return SOPHIA_E_ARG;
    }
    for (int k = 0; k < SOPHIA_TABLE_TYPES; k++) {
      if (t.size[k] == 0) continue;  // no block for this type: constants
      if (t.size[k] < 2 || t.index[k] < 0 || t.index[k] + t.size[k] > t.n_values) {
        set_error("set_table: block %d (first %d, %d values) outside the %d values", k, t.index[k], t.size[k],
                  t.n_values);
        return SOPHIA_E_ARG;
      }
      for (int j = t.index[k] + 1; j < t.index[k] + t.size[k]; j++)
        if (!(t.T[j] > t.T[j - 1])) {
          set_error("set_table: temperatures of block %d are not ascending", k);
          return SOPHIA_E_ARG;
        }
    }
    if (!dtab) TRY(arena.alloc(&dtab, (size_t)4 * SOPHIA_TABLE_VALUES));
    std::vector<double> h((size_t)4 * SOPHIA_TABLE_VALUES);
    std::memcpy(h.data(), t.T, sizeof(t.T));
    std::memcpy(h.data() + SOPHIA_TABLE_VALUES, t.h, sizeof(t.h));
    std::memcpy(h.data() + 2 * SOPHIA_TABLE_VALUES, t.k, sizeof(t.k));
    std::memcpy(h.data() + 3 * SOPHIA_TABLE_VALUES, t.mu, sizeof(t.mu));
    HIP_TRY(hipMemcpyAsync(dtab, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    table = t;
    has_table = true;
    return 0;
  }

  int monitors(int species, double* out) override {
    if (species == SOPHIA_SPH) {
      if (!sph_on() || ns == 0) {
        out[0] = out[1] = out[2] = 0.0;
        return 0;
      }
      return monitor_max<R>(stream, ns, sv.ptype, sv.itype, sv.ux, sv.uy, sv.uz, sv.rho, sv.fmag, mon_scratch, out);
    }
    // the ISPH-DEM path never writes P3.ftotal of spheres: that monitor is 0
    return monitor_max<double>(stream, nd, dv.ptype, dv.itype, dv.ux, dv.uy, dv.uz, dv.rho, nullptr, mon_scratch,
                               out);
  }
  int ppe_coarse(int32_t* dims, float* E, float* X) const override {
    const CoarseWork* c = ppe_w.coarse;
    if (!c || c->na == 0 || !c->valid) {
      set_error("no two-level coarse operator: run a solve with precond = SOPHIA_PRECOND_TWO_LEVEL first");
      return SOPHIA_E_STATE;
    }
    if (c->e_gen != c->x_gen) {
      // the last solve started from a guess that already met rtol: its E was
      // built ahead of a setup that never ran, so X belongs to an earlier E
      set_error("coarse operator rebuilt for a solve that needed no iterations; its inverse is not refreshed");
      return SOPHIA_E_STATE;
    }
    if (dims) {
      dims[0] = c->AX;
      dims[1] = c->AY;
      dims[2] = c->AZ;
    }
    return coarse_export(stream, *c, E, X);
  }
  int ppe_layout(int64_t* nnz, int32_t* ib, int32_t* cb) const override {
    if (!ppe_assembled) {
      set_error("PPE not assembled yet");
      return SOPHIA_E_STATE;
    }
    if (nnz) *nnz = ppe_w.nnz;
    if (ib) *ib = ppe_w.fmt == kFmtRun8 ? (ppe_w.mixed ? 0 : 1) : 4;
    if (cb) *cb = ppe_w.mixed ? 4 : 8;
    return 0;
  }
  int ppe_matvec_probe_host(int variant, int reps, const double* x, double* y, int n, float* ms) {
    if (!ppe_assembled || n != ns) {
      set_error("matvec probe: operator not assembled or n mismatch");
      return SOPHIA_E_STATE;
    }
    double *dx = nullptr, *dy = nullptr;
    float2* pk = nullptr;
    TRY(arena.alloc(&dx, n));
    TRY(arena.alloc(&dy, n));
    TRY(arena.alloc(&pk, n));
    HIP_TRY(hipMemcpyAsync(dx, x, (size_t)n * 8, hipMemcpyHostToDevice, stream));
    int rc = ppe_matvec_probe<R>(stream, devparams(), sv, ppe_w, variant, reps, dx, dy, pk, ms);
    if (!rc) {
      hipError_t e = hipMemcpyAsync(y, dy, (size_t)n * 8, hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) rc = hip_fail(e, "matvec probe copy", __FILE__, __LINE__);
    }
    arena.drop(dx);
    arena.drop(dy);
    arena.drop(pk);
    return rc;
  }
  int matvec_probe(int variant, int reps, const double* x, double* y, int n, float* ms) override {
    return ppe_matvec_probe_host(variant, reps, x, y, n, ms);
  }

  int ppe_apply(const double* x, double* y, int n) override {
    if (!ppe_assembled || n != ns) {
      set_error("ppe_apply: operator not assembled or n mismatch");
      return SOPHIA_E_STATE;
    }
    double *dx = nullptr, *dy = nullptr;
    TRY(arena.alloc(&dx, n));
    TRY(arena.alloc(&dy, n));
    HIP_TRY(hipMemcpyAsync(dx, x, (size_t)n * 8, hipMemcpyHostToDevice, stream));
    int rc = ppe_apply64(stream, ppe_w, dx, dy, false, nullptr);
    if (!rc) {
      hipError_t e = hipMemcpyAsync(y, dy, (size_t)n * 8, hipMemcpyDeviceToHost, stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) rc = hip_fail(e, "ppe_apply copy", __FILE__, __LINE__);
    }
    arena.drop(dx);
    arena.drop(dy);
    return rc;
  }

  // the pressure before this solve becomes history level 1, level 1 level 2
  int rotate_pressure_history() {
    std::swap(sph1.cur, sph2.cur);
    HIP_TRY(hipMemcpyAsync(sph1.cur, sv.pres, (size_t)ns * 8, hipMemcpyDeviceToDevice, stream));
    if (p_hist < 2) p_hist++;
    return 0;
  }

  // --- learned warm start -------------------------------------------------------
  // residual of the base guess and the network's node inputs
  int mlp_prepare(const double* base, double* scale, double* rr_base = nullptr) {
    TRY(ppe_apply64(stream, ppe_w, base, mlp.r0, true, nullptr));
    return mlp_features<R>(stream, devparams(), sv, mlp, ppe_w.row, ppe_w.diag, scale, rr_base);
  }
  int mlp_check(int n) const {
    if (!ppe_assembled || n != ns) {
      set_error("learned warm start: PPE not assembled this step or n=%d != %d", n, ns);
      return SOPHIA_E_STATE;
    }
    if (!mlp.ready) {
      set_error("learned warm start: parameters not set");
      return SOPHIA_E_STATE;
    }
    return 0;
  }
  int mlp_predict(double* out, int n) override {
    TRY(mlp_check(n));
    double scale = 0;
    TRY(mlp_prepare(sv.pres, &scale));
    TRY(mlp_forward<R>(stream, devparams(), sv, mlp, ppe_w.row, sv.pres, scale, mlp.guess));
    HIP_TRY(hipMemcpyAsync(out, mlp.guess, (size_t)n * 8, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
  }
  int mlp_train(const double* target, int n, double lr, double* loss, float* grad, int n_params, bool step) override {
    TRY(mlp_check(n));
    if (grad && n_params != kNParam) {
      set_error("mlp_loss_grad: buffer of %d, network has %d", n_params, kNParam);
      return SOPHIA_E_ARG;
    }
    double scale = 0;
    TRY(mlp_prepare(sv.pres, &scale));
    HIP_TRY(hipMemcpyAsync(mlp.target, target, (size_t)n * 8, hipMemcpyHostToDevice, stream));
    TRY(mlp_backward<R>(stream, devparams(), sv, mlp, ppe_w.row, sv.pres, mlp.target, scale, n_rows, step, lr, loss));
    if (grad) {
      HIP_TRY(hipMemcpyAsync(grad, mlp.grad, (size_t)kNParam * 4, hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    return 0;
  }
  int mlp_feat(float* nf, int n, double* scale) override {
    if (!ppe_assembled || n != ns) {
      set_error("mlp_features: PPE not assembled this step or n mismatch");
      return SOPHIA_E_STATE;
    }
    double sc = 0;
    TRY(mlp_prepare(sv.pres, &sc));
    if (scale) *scale = sc;
    if (nf) {
      HIP_TRY(hipMemcpyAsync(nf, mlp.nf, (size_t)n * sizeof(float2), hipMemcpyDeviceToHost, stream));
      HIP_TRY(hipStreamSynchronize(stream));
    }
    return 0;
  }
  int mlp_set(const float* w, int n) override { return mlp_set_params(stream, mlp, w, n); }
  int mlp_get(float* w, int n) const override { return mlp_get_params(stream, mlp, w, n); }
  int mlp_reset() override { return mlp_reset_adam(stream, mlp); }
};

}  // namespace

// ---------------------------------------------------------------------------
// parameters and files
// ---------------------------------------------------------------------------
extern "C" {

const char* sophia_last_error(void) { return sophia::g_err.c_str(); }

void sophia_params_default(sophia_params* p) {
  std::memset(p, 0, sizeof(*p));
  p->solver_type = 1;
  p->dim = 3;
  p->kernel_type = SOPHIA_KNL_WENDLAND2;
  p->rho_type = 1;
  p->time_type = 1;
  p->open_boundary = 0;
  p->flag_z_index = 0;
  p->freq_filt = 40;
  p->freq_output = 10000;
  p->fp_solve = 1;
  p->fv_solve = 1;
  p->fva_solve = 0;
  p->fg_solve = 1;
  p->con_solve = 0;
  p->xsph_solve = 1;
  p->kgc_solve = 1;
  p->delSPH_solve = 2;
  p->decouple_stride = 1;
  p->min_iteration = 3;
  p->max_iteration = 10;
  p->kappa = 2.0;
  p->p_ref = 0.0;
  p->gamma = 1.0;
  p->dt = 5.0e-7;
  p->time = 0.0;
  p->time_end = 5.0;
  p->u_limit = 40.0;
  p->c_xsph = 0.1;
  p->soundspeed = 200.0;
  p->rho0_eos = 1.784;
  p->drho_th = 0.005;
  p->dp_th = 10;
  p->p_relaxation = 1;
  // reference constants (Parameters.cuh, functions_PROP.cuh, functions_DEM_INTERACTION2.cuh)
  p->rho_fluid = 1.784;
  p->rho_water = 1000.0;
  p->h_over_dx = 1.6;
  p->mu_fluid = 2.0e-5;
  p->mu_water = 1.0e-3;
  p->k_fluid = 0.0518;
  p->k_dem = 2.5;
  p->cp_fluid = 1006.4;
  p->cp_dem = 1068.0;
  p->gravity = 9.8;
  p->eos_b_factor = 1.186;
  p->delta_sph = 0.1;
  p->dem_shear_modulus = 8.3e8;
  p->dem_poisson = 0.23;
  p->dem_friction = 0.5;
  p->dem_damping = std::sqrt(3.333333333) * 0.0535184;
  p->dem_hertz_factor = 1.33333;
  p->drag_switch = 0.65;
  p->eulerian_sph = 1;
  p->fix_xsph_position = 0;
  p->dem_max_contacts = 16;
  p->sph_max_neighbors = 0;
  // DEM walls off; geometry and contact constants of functions_DEM_BC.cuh
  p->dem_walls = 0;
  const double cyl[3] = {0.0, 0.0, 0.0125};                                      // :254-257
  const double box[6] = {0.0, 0.0, -0.00176647182, 1.15, 1.015, 0.03353294364};  // :1503-1509
  const double cone[5] = {0.0, 0.0, 0.0125, 0.0018, 0.01853294364};              // :723-724, 4240-4242
  std::memcpy(p->wall_cyl, cyl, sizeof(cyl));
  std::memcpy(p->wall_box, box, sizeof(box));
  std::memcpy(p->wall_cone, cone, sizeof(cone));
  p->wall_rest[0] = 0.9;   // rest_n (:1)
  p->wall_rest[1] = -0.9;  // rest_s (:2)
  p->wall_mu[0] = 0.10;    // mu_s (:4)
  p->wall_mu[1] = 0.2;     // mu_rs (:5)
  // open-boundary bands and inlet state (functions_NNPS.cuh:2-5; functions_BC.cuh:961-964)
  const double band[4] = {-10.505, -0.021, 0.006, 10.25};
  std::memcpy(p->ob_band, band, sizeof(band));
  p->ob_inlet_u = 24.045;
  p->ob_inlet_temp = 823.0;
  p->reference_compat = 0;
  // heat constants (functions_SPH_DEM_COUPLING2.cuh:708-718, 997, 1041, 1081; functions_PROP.cuh:279-340)
  p->pr_fluid = 0.7;
  p->pr_air = 1.0;
  p->dem_k_mat[0] = 0.84;
  p->dem_k_mat[1] = 55.0;
  p->dem_alpha_mat[0] = 2.5e-6;
  p->dem_alpha_mat[1] = 1.77e-5;
  p->dem_mat_radius = 0.0012;
  p->heat_flux = 0.0;
  p->free_surface = 0;
  p->fs_threshold = 0.3;
  p->prop_table = 0;
  p->noslip_bc = 0;
}

// read_solv_input (function_init.cuh:216-615): whitespace tokens; a key is
// followed by a fixed number of tokens before its value
int sophia_read_solv(const char* path, sophia_params* p) {
  FILE* f = std::fopen(path, "r");
  if (!f) {
    set_error("cannot open %s", path);
    return SOPHIA_E_IO;
  }
  std::vector<std::string> tok;
  char buf[1024];
  while (std::fscanf(f, "%1023s", buf) == 1) tok.push_back(buf);
  std::fclose(f);
  auto at = [&](size_t i) -> std::string { return i < tok.size() ? tok[i] : std::string(); };
  auto yes = [](const std::string& s) { return s == "YES" ? 1 : 0; };
  for (size_t i = 0; i < tok.size(); i++) {
    const std::string& k = tok[i];
    if (k == "solver_type") {
      std::string v = at(i + 2);
      p->solver_type = v == "ISPH" ? 1 : (v == "DFSPH" ? 2 : 0);
    } else if (k == "dimension(1/2/3):") {
      p->dim = std::atoi(at(i + 1).c_str());
    } else if (k == "kernel" && at(i + 1) == "type") {
      std::string v = at(i + 3);
      if (v == "Quartic") p->kernel_type = SOPHIA_KNL_QUARTIC;
      if (v == "Gaussian") p->kernel_type = SOPHIA_KNL_GAUSSIAN;
      if (v == "Quintic") p->kernel_type = SOPHIA_KNL_QUINTIC;
      if (v == "Wendland2") p->kernel_type = SOPHIA_KNL_WENDLAND2;
      if (v == "Wendland4") p->kernel_type = SOPHIA_KNL_WENDLAND4;
      if (v == "Wendland6") p->kernel_type = SOPHIA_KNL_WENDLAND6;
    } else if (k == "density" && at(i + 1) == "calculation") {
      std::string v = at(i + 3);
      if (v == "Direct") p->rho_type = 0;
      if (v == "Continuity") p->rho_type = 1;
    } else if (k == "time" && at(i + 1) == "integration") {
      std::string v = at(i + 3);
      if (v == "Euler") p->time_type = 0;
      if (v == "Predictor_Corrector") p->time_type = 1;
    } else if (k == "open" && at(i + 1) == "boundary") {
      std::string v = at(i + 3);
      if (v == "OPEN") p->open_boundary = 1;
      if (v == "PERIODIC") p->open_boundary = 2;
      if (v == "NO") p->open_boundary = 0;
    } else if (k == "pressure-force") {
      p->fp_solve = yes(at(i + 2));
    } else if (k == "viscous-force") {
      p->fv_solve = yes(at(i + 2));
    } else if (k == "artificial-viscous-force") {
      p->fva_solve = yes(at(i + 2));
    } else if (k == "gravitational-force") {
      p->fg_solve = yes(at(i + 2));
    } else if (k == "property_table(YES/NO):") {
      p->prop_table = yes(at(i + 1));
    } else if (k == "Conduction(YES/NO):") {
      p->con_solve = yes(at(i + 1));
    } else if (k == "XSPH") {
      p->xsph_solve = yes(at(i + 2));
    } else if (k == "kernel-gradient-correction") {
      std::string v = at(i + 2);
      p->kgc_solve = v == "KGC" ? 1 : (v == "FPM" ? 2 : (v == "DFPM" ? 3 : (v == "KGF" ? 4 : 0)));
    } else if (k == "delta-SPH") {
      std::string v = at(i + 2);
      p->delSPH_solve = v == "Molteni" ? 1 : (v == "Antuono" ? 2 : 0);
    } else if (k == "noslip_boundary") {
      // function_init.cuh:446-450: the value two words on
      std::string v = at(i + 2);
      if (v == "YES") p->noslip_bc = 1;
      if (v == "NO") p->noslip_bc = 0;
    } else if (k == "reference-pressure") {
      p->p_ref = std::atof(at(i + 2).c_str());
    } else if (k == "sound-speed:") {
      p->soundspeed = std::atof(at(i + 1).c_str());
    } else if (k == "reference-density-eos:") {
      p->rho0_eos = std::atof(at(i + 1).c_str());
    } else if (k == "gamma") {
      p->gamma = std::atof(at(i + 3).c_str());
    } else if (k == "kappa:") {
      p->kappa = std::atof(at(i + 1).c_str());
    } else if (k == "XSPH-coefficient:") {
      p->c_xsph = std::atof(at(i + 1).c_str());
    } else if (k == "minimum-iteration:") {
      p->min_iteration = std::atoi(at(i + 1).c_str());
    } else if (k == "maximum-iteration:") {
      p->max_iteration = std::atoi(at(i + 1).c_str());
    } else if (k == "pressure-convergence") {
      p->dp_th = std::atof(at(i + 2).c_str());
    } else if (k == "density-convergence") {
      p->drho_th = std::atof(at(i + 2).c_str());
    } else if (k == "pressure-relaxation") {
      p->p_relaxation = std::atof(at(i + 2).c_str());
    } else if (k == "time-step") {
      p->dt = std::atof(at(i + 2).c_str());
    } else if (k == "start-time") {
      p->time = std::atof(at(i + 2).c_str());
    } else if (k == "end-time") {
      p->time_end = std::atof(at(i + 2).c_str());
    } else if (k == "z-indexing") {
      p->flag_z_index = yes(at(i + 2));
    } else if (k == "filtering" && at(i + 1) == "frequency:") {
      p->freq_filt = std::atoi(at(i + 2).c_str());
    } else if (k == "plot-output") {
      p->freq_output = std::atoi(at(i + 2).c_str());
    } else if (k == "velocity-limit") {
      p->u_limit = std::atof(at(i + 2).c_str());
    } else if (k == "Xmargin(-):") {
      p->xmargin_m = std::atof(at(i + 1).c_str());
    } else if (k == "Xmargin(+):") {
      p->xmargin_p = std::atof(at(i + 1).c_str());
    } else if (k == "Ymargin(-):") {
      p->ymargin_m = std::atof(at(i + 1).c_str());
    } else if (k == "Ymargin(+):") {
      p->ymargin_p = std::atof(at(i + 1).c_str());
    } else if (k == "Zmargin(-):") {
      p->zmargin_m = std::atof(at(i + 1).c_str());
    } else if (k == "Zmargin(+):") {
      p->zmargin_p = std::atof(at(i + 1).c_str());
    } else if (k == "decouple-stride:") {
      p->decouple_stride = std::atoi(at(i + 1).c_str());
    }
  }
  return 0;
}

// read_input (function_init.cuh:743-891): first line = tab-separated column
// codes, then values row by row; i_type = 1 for every particle read
int sophia_read_input(const char* path, part1* out, int32_t capacity, int32_t* n) {
  FILE* f = std::fopen(path, "r");
  if (!f) {
    set_error("cannot open %s", path);
    return SOPHIA_E_IO;
  }
  char line[1 << 16];
  if (!std::fgets(line, sizeof(line), f)) {
    std::fclose(f);
    set_error("%s: empty file", path);
    return SOPHIA_E_IO;
  }
  std::vector<int> cols;
  for (char* t = std::strtok(line, "\t\r\n "); t; t = std::strtok(nullptr, "\t\r\n ")) cols.push_back(std::atoi(t));
  if (cols.empty()) {
    std::fclose(f);
    set_error("%s: no column codes", path);
    return SOPHIA_E_IO;
  }
  int count = 0;
  char tokb[256];
  bool done = false;
  while (!done) {
    part1 q;
    std::memset(&q, 0, sizeof(q));
    for (size_t c = 0; c < cols.size(); c++) {
      if (std::fscanf(f, "%255s", tokb) != 1) {
        done = true;
        break;
      }
      double v = std::atof(tokb);
      int iv = std::atoi(tokb);
      switch (cols[c]) {
        case 1: q.x = v; break;
        case 2: q.y = v; break;
        case 3: q.z = v; break;
        case 4: q.ux = v; break;
        case 5: q.uy = v; break;
        case 6: q.uz = v; break;
        case 7: q.m = v; break;
        case 8: q.p_type = iv; break;
        case 9: q.h = v; break;
        case 10: q.temp = v; break;
        case 11: q.pres = v; break;
        case 12: q.rho = v; break;
        case 15: q.concn = v; break;
        case 18: q.ct_boundary = iv; break;
        case 26: q.flt_s = v; break;
        case 27: q.rad = v; break;
        case 28: q.ri = v; break;
        case 29: q.dem_idx = iv; break;
        case 30: q.wx = v; break;
        case 31: q.wy = v; break;
        case 32: q.wz = v; break;
        case 33: q.temp1 = v; break;
        case 34: q.temp2 = v; break;
        case 35: q.temp3 = v; break;
        case 36: q.vol_power = v; break;
        default: break;  // codes the reader ignores (13, 14, 16, 17, 19-25)
      }
    }
    if (done) break;
    q.i_type = 1;
    if (out) {
      if (count >= capacity) {
        std::fclose(f);
        set_error("%s: more than %d particles", path, capacity);
        return SOPHIA_E_CAPACITY;
      }
      out[count] = q;
    }
    count++;
  }
  std::fclose(f);
  *n = count;
  return 0;
}

int sophia_read_ptype(const char* path, int32_t* out, int32_t capacity, int32_t* n) {
  FILE* f = std::fopen(path, "r");
  if (!f) {
    set_error("cannot open %s", path);
    return SOPHIA_E_IO;
  }
  int v, c = 0;
  while (std::fscanf(f, "%d", &v) == 1) {
    if (out) {
      if (c >= capacity) {
        std::fclose(f);
        set_error("%s: more than %d entries", path, capacity);
        return SOPHIA_E_CAPACITY;
      }
      out[c] = v;
    }
    c++;
  }
  std::fclose(f);
  *n = c;
  return 0;
}

int sophia_read_table(const char* path, sophia_table* out) {
  if (!path || !out) return SOPHIA_E_ARG;
  FILE* f = std::fopen(path, "r");
  if (!f) {
    set_error("cannot open %s", path);
    return SOPHIA_E_IO;
  }
  std::memset(out, 0, sizeof(*out));
  char line[1000];
  int block = -1, start = 0, rows = 0, len = 0;
  int err = 0;
  auto fail = [&](const char* fmt, int a) {
    set_error(fmt, path, a);
    err = SOPHIA_E_IO;
  };
  // one row: comma-separated numbers, written from the block's first slot
  auto parse_row = [&](char* text, double* to) -> int {
    int c = 0;
    for (char* tok = std::strtok(text, ","); tok; tok = std::strtok(nullptr, ",")) {
      char* e = nullptr;
      double v = std::strtod(tok, &e);
      if (e == tok) continue;  // the newline after the last comma
      if (start + c >= SOPHIA_TABLE_VALUES) return -1;
      to[start + c] = v;
      c++;
    }
    return c;
  };
  while (!err && std::fgets(line, sizeof(line), f)) {
    char* t = line;
    while (*t == ' ' || *t == '\t') t++;
    if (block < 0) {
      if (!std::strncmp(t, "#p_type", 7)) {
        int num = std::atoi(t + 7);
        if (num < 0 || num >= SOPHIA_TABLE_TYPES) {
          fail("%s: block #p_type %d outside 0..9", num);
          break;
        }
        if (out->size[num] != 0) {
          fail("%s: two blocks for p_type %d", num);
          break;
        }
        block = num;
        rows = 0;
        len = SOPHIA_TABLE_VALUES;
      }
      continue;
    }
    if (!std::strncmp(t, "#end", 4)) {
      if (rows != 15 || len < 2) {
        fail("%s: block %d needs the rows //T, //h, //k and //mu with at least 2 values each", block);
        break;
      }
      out->index[block] = start;
      out->size[block] = len;
      start += len;
      out->n_tables++;
      out->n_values = start;
      block = -1;
      continue;
    }
    double* dst = nullptr;
    int bit = 0;
    if (!std::strncmp(t, "//T", 3)) dst = out->T, bit = 1;
    else if (!std::strncmp(t, "//h", 3)) dst = out->h, bit = 2;
    else if (!std::strncmp(t, "//k", 3)) dst = out->k, bit = 4;
    else if (!std::strncmp(t, "//mu", 4)) dst = out->mu, bit = 8;
    else continue;
    if (!std::fgets(line, sizeof(line), f)) {
      fail("%s: block %d ends inside a row", block);
      break;
    }
    int c = parse_row(line, dst);
    if (c < 0) {
      fail("%s: more than %d table values", SOPHIA_TABLE_VALUES);
      break;
    }
    rows |= bit;
    len = std::min(len, c);
  }
  std::fclose(f);
  if (!err && block >= 0) {
    set_error("%s: block %d has no #end", path, block);
    err = SOPHIA_E_IO;
  }
  return err;
}

int sophia_table_eval(const sophia_table* t, int32_t p_type, int32_t what, double x, double* y) {
  if (!t || !y) return SOPHIA_E_ARG;
  const int pt = std::abs(p_type);
  if (pt >= SOPHIA_TABLE_TYPES || t->size[pt] < 2) {
    set_error("table: no block for p_type %d", p_type);
    return SOPHIA_E_ARG;
  }
  const int i0 = t->index[pt], n = t->size[pt];
  const double *xs = t->T + i0, *ys = nullptr;
  switch (what) {
    case SOPHIA_PROP_MU: ys = t->mu + i0; break;
    case SOPHIA_PROP_K: ys = t->k + i0; break;
    case SOPHIA_PROP_H: ys = t->h + i0; break;
    case SOPHIA_PROP_T: xs = t->h + i0; ys = t->T + i0; break;
    default:
      set_error("table: unknown property %d", what);
      return SOPHIA_E_ARG;
  }
  *y = sophia::interp_table(xs, ys, n, x);
  return 0;
}

// ---------------------------------------------------------------------------
// context lifetime
// ---------------------------------------------------------------------------
int sophia_create_from_arrays(const sophia_params* p, const part1* particles, int32_t n, int32_t device,
                              int32_t precision, sophia_ctx** out) {
  if (!p || !particles || n <= 0 || !out) {
    set_error("create: null argument or n <= 0");
    return SOPHIA_E_ARG;
  }
  if (p->dim != 2 && p->dim != 3) {
    set_error("create: dim must be 2 or 3");
    return SOPHIA_E_ARG;
  }
  if (p->solver_type != 0 && p->dim != 3) {
    // the ISPH-DEM stages exist in 3D only (KERNEL_*3D_sph / _dem)
    set_error("create: the ISPH-DEM step is 3D only (dim = %d)", p->dim);
    return SOPHIA_E_ARG;
  }
  if (p->periodic_mask & 7) {
    // periodic boxes: ISPH fluid only, in a fixed domain whose periodic axes
    // share one period of at least three cells; the wall, sphere and WCSPH
    // kernels do not wrap
    const int m = p->periodic_mask & (p->dim == 2 ? 3 : 7);
    double L = 0;
    bool ok = p->domain_fixed && p->solver_type == 1 && p->dim == 3;
    for (int a = 0; ok && a < 3; a++) {
      const double s = p->domain_max[a] - p->domain_min[a];
      if (!(s > 0)) ok = false;
      if (m >> a & 1) {
        if (L == 0) L = s;
        if (std::fabs(s - L) > 1e-12 * L || s < 3.0 * 1.1 * p->kappa * particles[0].h) ok = false;
      }
    }
    for (int a = 0; ok && a < 3; a++)
      if (!(m >> a & 1) && p->domain_max[a] - p->domain_min[a] > L * (1000.0 / 1024.0)) ok = false;
    if (!ok) {
      set_error("create: periodic_mask needs the 3D ISPH step, a fixed domain, equal periods of >= 3 cells on "
                "the periodic axes and no longer span on the others");
      return SOPHIA_E_ARG;
    }
    for (int i = 0; i < n; i++)
      if (particles[i].p_type <= 0 || particles[i].p_type > 1000) {
        set_error("create: periodic boxes take fluid only (particle %d has p_type %d)", i, particles[i].p_type);
        return SOPHIA_E_ARG;
      }
  }
  if (p->solver_type == 0 && p->dim == 2) {
    // every 2D DEM call of the WCSPH drivers is commented out
    // (WCSPH_Calc.cuh:205, 224, 235, 503-535)
    for (int i = 0; i < n; i++)
      if (particles[i].p_type > 1000) {
        set_error("create: 2D WCSPH takes no spheres (particle %d has p_type %d)", i, particles[i].p_type);
        return SOPHIA_E_ARG;
      }
  }
  if (p->kappa <= 0 || p->dt <= 0) {
    set_error("create: kappa and dt must be positive");
    return SOPHIA_E_ARG;
  }
  for (int i = 0; i < n; i++)
    if (!(particles[i].h > 0)) {
      set_error("create: particle %d has h <= 0", i);
      return SOPHIA_E_ARG;
    }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    set_error("create: no HIP device available (%s)", hipGetErrorString(e));
    return SOPHIA_E_HIP;
  }
  if (device < 0 || device >= ndev) {
    set_error("create: device %d out of range (%d devices)", device, ndev);
    return SOPHIA_E_ARG;
  }
  sophia_ctx* c = nullptr;
  int rc;
  if (precision == SOPHIA_PREC_FP64) {
    auto* x = new Ctx<double>();
    x->device = device;
    x->precision = precision;
    rc = x->init(*p, particles, n);
    c = x;
  } else if (precision == SOPHIA_PREC_MIXED) {
    auto* x = new Ctx<float>();
    x->device = device;
    x->precision = precision;
    rc = x->init(*p, particles, n);
    c = x;
  } else {
    set_error("create: unknown precision %d", precision);
    return SOPHIA_E_ARG;
  }
  if (rc) {
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

int sophia_create(const char* solv_path, const char* input_path, const char* ptype_path, const char* data_path,
                  int32_t device, int32_t precision, sophia_ctx** out) {
  if (!solv_path || !input_path || !out) {
    set_error("create: null path or output");
    return SOPHIA_E_ARG;
  }
  sophia_params p;
  sophia_params_default(&p);
  int rc = sophia_read_solv(solv_path, &p);
  if (rc) return rc;
  int32_t n = 0;
  rc = sophia_read_input(input_path, nullptr, 0, &n);
  if (rc) return rc;
  std::vector<part1> v((size_t)n);
  rc = sophia_read_input(input_path, v.data(), n, &n);
  if (rc) return rc;
  if (ptype_path) {
    int32_t m = 0;
    rc = sophia_read_ptype(ptype_path, nullptr, 0, &m);
    if (rc) return rc;
    std::vector<int32_t> pt((size_t)m);
    rc = sophia_read_ptype(ptype_path, pt.data(), m, &m);
    if (rc) return rc;
    if (m != n) {
      set_error("create: %s lists %d types for %d particles of %s", ptype_path, m, n, input_path);
      return SOPHIA_E_ARG;
    }
    for (int i = 0; i < n; i++)
      if (pt[i] != v[i].p_type) {
        set_error("create: particle %d has p_type %d in %s and %d in %s", i, pt[i], ptype_path, v[i].p_type,
                  input_path);
        return SOPHIA_E_ARG;
      }
  }
  sophia_table tab;
  if (data_path) {
    rc = sophia_read_table(data_path, &tab);
    if (rc) return rc;
  } else if (p.prop_table) {
    set_error("create: %s asks for property tables and no data.txt was given", solv_path);
    return SOPHIA_E_ARG;
  }
  rc = sophia_create_from_arrays(&p, v.data(), n, device, precision, out);
  if (rc) return rc;
  if (data_path) {
    rc = sophia_set_table(*out, &tab);
    if (rc) {
      sophia_destroy(*out);
      *out = nullptr;
    }
  }
  return rc;
}

int sophia_destroy(sophia_ctx* ctx) {
  delete ctx;
  return 0;
}

#define CHECK_CTX(c)                  \
  do {                                \
    if (!(c)) {                       \
      set_error("null context");      \
      return SOPHIA_E_ARG;            \
    }                                 \
    hipError_t e_ = hipSetDevice((c)->device); \
    if (e_ != hipSuccess) return hip_fail(e_, "hipSetDevice", __FILE__, __LINE__); \
  } while (0)

int sophia_get_domain(const sophia_ctx* c, sophia_domain* d) {
  if (!c || !d) return SOPHIA_E_ARG;
  *d = c->d;
  return 0;
}
int sophia_get_params(const sophia_ctx* c, sophia_params* p) {
  if (!c || !p) return SOPHIA_E_ARG;
  *p = c->p;
  return 0;
}
int sophia_get_clock(const sophia_ctx* c, int64_t* count, double* time) {
  if (!c) return SOPHIA_E_ARG;
  if (count) *count = c->count;
  if (time) *time = c->time;
  return 0;
}
int sophia_set_table(sophia_ctx* c, const sophia_table* t) {
  if (!c || !t) return SOPHIA_E_ARG;
  return c->set_table(*t);
}

int sophia_sync(sophia_ctx* c) {
  CHECK_CTX(c);
  HIP_TRY(hipStreamSynchronize(c->stream));
  return 0;
}
int sophia_count(const sophia_ctx* c, int32_t species, int32_t* n) {
  if (!c || !n) return SOPHIA_E_ARG;
  *n = species == SOPHIA_SPH ? c->d.n_sph : c->d.n_dem;
  return 0;
}

extern "C++" {
namespace {

// FloatSwap / IntSwap (functions_OUTPUT.cuh:68-121): each value as a
// big-endian 4-byte word; doubles are rounded to float first
void put_be32(FILE* f, uint32_t u) {
  const unsigned char b[4] = {(unsigned char)(u >> 24), (unsigned char)(u >> 16), (unsigned char)(u >> 8),
                              (unsigned char)u};
  std::fwrite(b, 1, 4, f);
}
void put_float(FILE* f, double v) {
  const float x = (float)v;
  uint32_t u;
  std::memcpy(&u, &x, 4);
  put_be32(f, u);
}

struct Records {
  std::vector<part1> p1;
  std::vector<part2> p2;
  std::vector<part3> p3;
};
int fetch(const sophia_ctx* c, int species, Records& r) {
  const int n = species == SOPHIA_SPH ? c->d.n_sph : c->d.n_dem;
  r.p1.assign(n, part1{});
  r.p2.assign(n, part2{});
  r.p3.assign(n, part3{});
  if (n == 0) return 0;
  return c->get_particles(species, r.p1.data(), r.p2.data(), r.p3.data());
}

// one array "name ncomp count type" over the selected particles; float
// arrays are rounded to float (FloatSwap of a double or an int), int arrays
// are written as they are (IntSwap)
using Sel = bool (*)(const part1&);
template <class G>
void vtk_field(FILE* f, const char* name, int count, const Records& r, Sel sel, G get) {
  std::fprintf(f, "%s\t1\t%d\tfloat\n", name, count);
  for (size_t i = 0; i < r.p1.size(); i++)
    if (sel(r.p1[i])) put_float(f, get(i));
}
template <class G>
void vtk_int(FILE* f, const char* name, int count, const Records& r, Sel sel, G get) {
  std::fprintf(f, "%s\t1\t%d\tint\n", name, count);
  for (size_t i = 0; i < r.p1.size(); i++)
    if (sel(r.p1[i])) put_be32(f, (uint32_t)get(i));
}
// three components per particle, interleaved (the fluid2 writer's vectors)
template <class G>
void vtk_vec3(FILE* f, const char* name, int count, const Records& r, Sel sel, G get) {
  std::fprintf(f, "%s\t3\t%d\tfloat\n", name, count);
  for (size_t i = 0; i < r.p1.size(); i++)
    if (sel(r.p1[i]))
      for (int c = 0; c < 3; c++) put_float(f, get(i, c));
}
bool sel_fluid(const part1& q) { return q.p_type == 1; }
bool sel_fluid2(const part1& q) { return q.p_type == 2; }
bool sel_air(const part1& q) { return q.p_type == 3; }
bool sel_wall(const part1& q) { return q.p_type == 0; }
bool sel_moving(const part1& q) { return q.p_type == 9; }
bool sel_active(const part1& q) { return q.i_type == 1 || q.i_type == 2; }
bool sel_dem(const part1& q) { return q.p_type > 1000; }
bool cnt_dem(const part1& q) { return q.dem_idx > 0; }

// one writer of functions_OUTPUT.cuh: the species it reads, the particles
// it writes, the predicate its point count uses (the sphere writers count
// dem_idx > 0 but write p_type > 1000) and its number of arrays
struct VtkKind {
  int species;
  Sel sel, cnt;
  int nfields;
};
const VtkKind kVtkKinds[] = {
    {SOPHIA_SPH, sel_fluid, sel_fluid, 9},     // SOPHIA_VTK_FLUID
    {SOPHIA_DEM, sel_dem, cnt_dem, 10},        // SOPHIA_VTK_DEM
    {SOPHIA_SPH, sel_air, sel_air, 10},        // SOPHIA_VTK_AIR
    {SOPHIA_DEM, sel_dem, cnt_dem, 10},        // SOPHIA_VTK_SOLID
    {SOPHIA_SPH, sel_moving, sel_moving, 2},   // SOPHIA_VTK_MOVING
    {SOPHIA_SPH, sel_fluid2, sel_fluid2, 7},   // SOPHIA_VTK_FLUID2
    {SOPHIA_SPH, sel_wall, sel_wall, 6},       // SOPHIA_VTK_BOUNDARY
    {SOPHIA_SPH, sel_fluid, sel_fluid, 12},    // SOPHIA_VTK_FLUID_PGF
    {SOPHIA_SPH, sel_active, sel_active, 3},   // SOPHIA_VTK_ACTIVE
    {SOPHIA_SPH, sel_moving, sel_moving, 2},   // SOPHIA_VTK_MOVING_UY
};
constexpr int kNumVtkKinds = (int)(sizeof(kVtkKinds) / sizeof(kVtkKinds[0]));

void vtk_arrays(FILE* f, int kind, int n, const Records& r, Sel sel) {
  const auto& P1 = r.p1;
  const auto& P2 = r.p2;
  const auto& P3 = r.p3;
  switch (kind) {
    case SOPHIA_VTK_FLUID:  // save_plot_fluid_vtk_bin_fluid_sph (:445-659)
      vtk_field(f, "porosity", n, r, sel, [&](size_t i) { return P1[i].DEMpor; });
      vtk_field(f, "conv_rate", n, r, sel, [&](size_t i) { return P1[i].Q_sdf; });
      vtk_field(f, "pressure", n, r, sel, [&](size_t i) { return P1[i].pres; });
      vtk_field(f, "cond_rate", n, r, sel, [&](size_t i) { return P1[i].Q_f; });
      vtk_field(f, "dtemp", n, r, sel, [&](size_t i) { return P3[i].dtemp; });
      vtk_int(f, "buffer_type", n, r, sel, [&](size_t i) { return P1[i].buffer_type; });
      vtk_field(f, "density", n, r, sel, [&](size_t i) { return P1[i].rho; });
      vtk_field(f, "temperature", n, r, sel, [&](size_t i) { return P1[i].temp; });
      vtk_field(f, "uz", n, r, sel, [&](size_t i) { return P1[i].uz; });
      break;
    case SOPHIA_VTK_DEM:  // save_plot_fluid_vtk_bin_solid_dem (:1135-1446)
      vtk_field(f, "itype", n, r, sel, [&](size_t i) { return (double)P1[i].i_type; });
      vtk_field(f, "mag_uijf", n, r, sel, [&](size_t i) { return P1[i].test6; });
      vtk_field(f, "Euler", n, r, sel, [&](size_t i) { return P1[i].eliz; });
      vtk_field(f, "porosity", n, r, sel, [&](size_t i) { return P1[i].DEMpor; });
      vtk_field(f, "uz", n, r, sel, [&](size_t i) { return P1[i].uz; });
      vtk_field(f, "Q_sd", n, r, sel, [&](size_t i) { return P1[i].Q_sd; });
      vtk_field(f, "ftotal_z", n, r, sel, [&](size_t i) { return P3[i].ftotalz; });
      vtk_field(f, "temperature_surface", n, r, sel, [&](size_t i) { return P1[i].temp; });
      vtk_field(f, "rad", n, r, sel, [&](size_t i) { return P1[i].rad; });
      vtk_field(f, "dtemp_dt", n, r, sel, [&](size_t i) { return P3[i].dtemp; });
      break;
    case SOPHIA_VTK_AIR:  // save_plot_fluid_vtk_bin_air (:660-822)
      vtk_field(f, "rho_ref", n, r, sel, [&](size_t i) { return P2[i].rho_ref; });
      vtk_field(f, "buffer_type", n, r, sel, [&](size_t i) { return (double)P1[i].buffer_type; });
      vtk_field(f, "uz", n, r, sel, [&](size_t i) { return P1[i].uz; });
      vtk_field(f, "temperature", n, r, sel, [&](size_t i) { return P1[i].temp; });
      vtk_field(f, "density", n, r, sel, [&](size_t i) { return P1[i].rho; });
      vtk_field(f, "porosity", n, r, sel, [&](size_t i) { return P1[i].DEMpor; });
      vtk_field(f, "pressure", n, r, sel, [&](size_t i) { return P1[i].pres; });
      vtk_field(f, "drho_dt", n, r, sel, [&](size_t i) { return P3[i].drho; });
      vtk_field(f, "dtemp_dt", n, r, sel, [&](size_t i) { return P3[i].dtemp; });
      vtk_field(f, "ptype", n, r, sel, [&](size_t i) { return (double)P1[i].p_type; });
      break;
    case SOPHIA_VTK_SOLID:  // save_plot_fluid_vtk_bin_solid (:824-1134)
      vtk_field(f, "dem_idx", n, r, sel, [&](size_t i) { return (double)P1[i].dem_idx; });
      vtk_field(f, "mag_uijf", n, r, sel, [&](size_t i) { return P1[i].test6; });
      vtk_field(f, "drag_r", n, r, sel, [&](size_t i) { return P1[i].test7; });
      vtk_field(f, "porosity", n, r, sel, [&](size_t i) { return P1[i].DEMpor; });
      vtk_field(f, "uz", n, r, sel, [&](size_t i) { return P1[i].uz; });
      vtk_field(f, "Q_sd", n, r, sel, [&](size_t i) { return P1[i].Q_sd; });
      vtk_field(f, "ftotal_z", n, r, sel, [&](size_t i) { return P3[i].ftotalz; });
      vtk_field(f, "temperature_surface", n, r, sel, [&](size_t i) { return P1[i].temp; });
      vtk_field(f, "rad", n, r, sel, [&](size_t i) { return P1[i].rad; });
      vtk_field(f, "dtemp_dt", n, r, sel, [&](size_t i) { return P3[i].dtemp; });
      break;
    case SOPHIA_VTK_MOVING:  // save_plot_fluid_vtk_bin_moving (:1447-1578)
      vtk_field(f, "density", n, r, sel, [&](size_t i) { return P1[i].rho; });
      vtk_field(f, "pressure", n, r, sel, [&](size_t i) { return P1[i].pres; });
      break;
    case SOPHIA_VTK_FLUID2:  // save_plot_fluid_vtk_bin_fluid2 (:1579-1702)
      vtk_field(f, "rho", n, r, sel, [&](size_t i) { return P1[i].rho; });
      vtk_field(f, "pres", n, r, sel, [&](size_t i) { return P1[i].pres; });
      vtk_field(f, "curvature", n, r, sel, [&](size_t i) { return P3[i].curv; });
      vtk_field(f, "detection", n, r, sel, [&](size_t i) { return P3[i].lbl_surf; });
      vtk_vec3(f, "surface_tension", n, r, sel, [&](size_t i, int c) {
        return c == 0 ? P3[i].fsx : c == 1 ? P3[i].fsy : P3[i].fsz;
      });
      vtk_vec3(f, "normal_vector_c", n, r, sel, [&](size_t i, int c) {
        return c == 0 ? P3[i].nx_c : c == 1 ? P3[i].ny_c : P3[i].nz_c;
      });
      vtk_vec3(f, "normal_vector", n, r, sel, [&](size_t i, int c) {
        return c == 0 ? P3[i].nx : c == 1 ? P3[i].ny : P3[i].nz;
      });
      break;
    case SOPHIA_VTK_BOUNDARY:  // save_plot_fluid_vtk_bin_boundary (:1703-1841)
      vtk_field(f, "uy", n, r, sel, [&](size_t i) { return P1[i].uy; });
      vtk_field(f, "ux", n, r, sel, [&](size_t i) { return P1[i].ux; });
      vtk_field(f, "uz", n, r, sel, [&](size_t i) { return P1[i].uz; });
      vtk_field(f, "temperature", n, r, sel, [&](size_t i) { return P1[i].temp; });
      vtk_field(f, "pressure", n, r, sel, [&](size_t i) { return P1[i].pres; });
      vtk_field(f, "density", n, r, sel, [&](size_t i) { return P1[i].rho; });
      break;
    case SOPHIA_VTK_FLUID_PGF:  // save_plot_fluid_vtk_bin_fluid (:202-444)
      vtk_field(f, "porosity", n, r, sel, [&](size_t i) { return P1[i].DEMpor; });
      vtk_field(f, "pressuregrad_x", n, r, sel, [&](size_t i) { return P1[i].pgf_x; });
      vtk_field(f, "pressuregrad_y", n, r, sel, [&](size_t i) { return P1[i].pgf_y; });
      vtk_field(f, "pressuregrad_z", n, r, sel, [&](size_t i) { return P1[i].pgf_z; });
      vtk_field(f, "ux", n, r, sel, [&](size_t i) { return P1[i].ux; });
      vtk_field(f, "pressure", n, r, sel, [&](size_t i) { return P1[i].pres; });
      vtk_field(f, "uy", n, r, sel, [&](size_t i) { return P1[i].uy; });
      vtk_field(f, "dtemp", n, r, sel, [&](size_t i) { return P3[i].dtemp; });
      vtk_int(f, "buffer_type", n, r, sel, [&](size_t i) { return P1[i].buffer_type; });
      vtk_field(f, "density", n, r, sel, [&](size_t i) { return P1[i].rho; });
      vtk_field(f, "temperature", n, r, sel, [&](size_t i) { return P1[i].temp; });
      vtk_field(f, "uz", n, r, sel, [&](size_t i) { return P1[i].uz; });
      break;
    case SOPHIA_VTK_ACTIVE:  // save_plot_fluid_vtk_bin (:126-201)
      vtk_field(f, "rho", n, r, sel, [&](size_t i) { return P1[i].rho; });
      vtk_int(f, "ptype", n, r, sel, [&](size_t i) { return P1[i].p_type; });
      vtk_field(f, "pres", n, r, sel, [&](size_t i) { return P1[i].pres; });
      break;
    case SOPHIA_VTK_MOVING_UY:  // save_plot_moving_vtk_bin (:1842-1925)
      vtk_field(f, "rho", n, r, sel, [&](size_t i) { return P1[i].rho; });
      vtk_field(f, "uy", n, r, sel, [&](size_t i) { return P1[i].uy; });
      break;
  }
}

}  // namespace
}  // extern "C++"

int sophia_write_vtk(const sophia_ctx* c, const char* path, int32_t kind) {
  if (!c || !path || kind < 0 || kind >= kNumVtkKinds) return SOPHIA_E_ARG;
  const VtkKind& k = kVtkKinds[kind];
  Records r;
  TRY(fetch(c, k.species, r));
  int count = 0;
  for (const part1& q : r.p1) count += k.cnt(q);
  FILE* f = std::fopen(path, "wb");
  if (!f) {
    set_error("cannot open %s for writing", path);
    return SOPHIA_E_IO;
  }
  std::fprintf(f, "# vtk DataFile Version 3.0\n");
  std::fprintf(f, "Print out results in vtk format\n");
  std::fprintf(f, "BINARY\n");
  std::fprintf(f, "DATASET POLYDATA\n");
  std::fprintf(f, "POINTS\t%d\tfloat\n", count);
  for (const part1& q : r.p1)
    if (k.sel(q)) {
      put_float(f, q.x);
      put_float(f, q.y);
      put_float(f, q.z);
    }
  std::fprintf(f, "POINT_DATA\t%d\n", count);
  std::fprintf(f, "FIELD FieldData\t%d\n", k.nfields);
  vtk_arrays(f, kind, count, r, k.sel);
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  if (!ok) {
    set_error("write error on %s", path);
    return SOPHIA_E_IO;
  }
  return 0;
}

int sophia_write_restart(const sophia_ctx* c, const char* path) {
  if (!c || !path) return SOPHIA_E_ARG;
  Records sph, dem;
  TRY(fetch(c, SOPHIA_SPH, sph));
  TRY(fetch(c, SOPHIA_DEM, dem));
  FILE* f = std::fopen(path, "w");
  if (!f) {
    set_error("cannot open %s for writing", path);
    return SOPHIA_E_IO;
  }
  std::fprintf(f, "1 2 3 4 5 6 7 8 9 10 11 12 13 26\n");
  for (const Records* r : {&sph, &dem})
    for (size_t i = 0; i < r->p1.size(); i++) {
      const part1& q = r->p1[i];
      std::fprintf(f, "%f\t%f\t%f\t", q.x, q.y, q.z);
      std::fprintf(f, "%f\t%f\t%f\t", q.ux, q.uy, q.uz);
      std::fprintf(f, "%e\t%d\t%e\t", q.m, q.p_type, q.h);
      std::fprintf(f, "%f\t%f\t%f\t", q.temp, q.pres, q.rho);
      std::fprintf(f, "%f\t%f\n", r->p2[i].rho_ref, q.flt_s);
    }
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  if (!ok) {
    set_error("write error on %s", path);
    return SOPHIA_E_IO;
  }
  return 0;
}

int sophia_nnps(sophia_ctx* c) { CHECK_CTX(c); return c->nnps(); }
int sophia_prep(sophia_ctx* c) { CHECK_CTX(c); return c->prep(); }
int sophia_interaction(sophia_ctx* c) { CHECK_CTX(c); return c->interaction(); }
int sophia_dem_interaction(sophia_ctx* c) { CHECK_CTX(c); return c->dem_interaction(); }
int sophia_sph_dem_coupling(sophia_ctx* c) { CHECK_CTX(c); return c->coupling(); }
int sophia_time(sophia_ctx* c, int32_t phase) {
  CHECK_CTX(c);
  if (phase != SOPHIA_PREDICT && phase != SOPHIA_CORRECT) {
    set_error("time: phase must be SOPHIA_PREDICT or SOPHIA_CORRECT");
    return SOPHIA_E_ARG;
  }
  return c->time_stage(phase);
}
int sophia_ppe(sophia_ctx* c, const sophia_ppe_opts* o, sophia_ppe_stats* st) { CHECK_CTX(c); return c->ppe(o, st); }
int sophia_xsph(sophia_ctx* c) { CHECK_CTX(c); return c->xsph(); }

int sophia_wcsph_stage(sophia_ctx* c, int32_t stage) {
  CHECK_CTX(c);
  return c->wcsph_stage(stage);
}

int sophia_step(sophia_ctx* c, const sophia_ppe_opts* o, sophia_ppe_stats* st) {
  CHECK_CTX(c);
  if (c->p.solver_type == 0) {
    if (st) std::memset(st, 0, sizeof(*st));
    for (int k = 0; k <= 8; k++) TRY(c->wcsph_stage(k));
    return 0;
  }
  TRY(c->nnps());
  TRY(c->prep());
  TRY(c->interaction_both());
  TRY(c->coupling());
  TRY(c->time_stage(SOPHIA_PREDICT));
  TRY(c->ppe(o, st));
  TRY(c->time_stage(SOPHIA_CORRECT));
  TRY(c->xsph());
  c->time += c->p.dt;
  c->count += 1;
  return 0;
}

int sophia_run(sophia_ctx* c, int32_t n, const sophia_ppe_opts* o, sophia_ppe_stats* st) {
  for (int i = 0; i < n; i++) TRY(sophia_step(c, o, st));
  return 0;
}

int sophia_get_nnps(const sophia_ctx* c, int32_t species, int32_t* keys, int32_t* order, uint32_t* cs,
                    uint32_t* ce, int32_t* nc) {
  if (!c) return SOPHIA_E_ARG;
  return c->get_nnps(species, keys, order, cs, ce, nc);
}
int sophia_get_particles(const sophia_ctx* c, int32_t species, part1* p1, part2* p2, part3* p3) {
  if (!c) return SOPHIA_E_ARG;
  return c->get_particles(species, p1, p2, p3);
}
int sophia_set_pressure(sophia_ctx* c, const double* p, int32_t n) { CHECK_CTX(c); return c->set_pressure(p, n); }
int sophia_get_ppe_system(const sophia_ctx* c, double* diag, double* rhs, int32_t* row) {
  if (!c) return SOPHIA_E_ARG;
  return c->get_ppe_system(diag, rhs, row);
}
int sophia_ppe_apply(sophia_ctx* c, const double* x, double* y, int32_t n) { CHECK_CTX(c); return c->ppe_apply(x, y, n); }
int sophia_ppe_matvec_probe(sophia_ctx* c, int32_t variant, int32_t reps, const double* x, double* y, int32_t n,
                            double* us_per_launch) {
  CHECK_CTX(c);
  if (!x || !y || reps < 0 || (variant != 0 && variant != 1)) return SOPHIA_E_ARG;
  float ms = 0.f;
  int rc = c->matvec_probe(variant, reps, x, y, n, &ms);
  if (!rc && us_per_launch) *us_per_launch = 1e3 * (double)ms;
  return rc;
}
int sophia_monitors(sophia_ctx* c, int32_t species, double* out) {
  CHECK_CTX(c);
  if (!out || (species != SOPHIA_SPH && species != SOPHIA_DEM)) return SOPHIA_E_ARG;
  return c->monitors(species, out);
}
int sophia_ppe_layout(const sophia_ctx* c, int64_t* nnz, int32_t* ib, int32_t* cb) {
  if (!c) return SOPHIA_E_ARG;
  return c->ppe_layout(nnz, ib, cb);
}
int sophia_ppe_coarse(const sophia_ctx* c, int32_t dims[3], float* E, float* X) {
  if (!c || !dims) return SOPHIA_E_ARG;
  return c->ppe_coarse(dims, E, X);
}

int sophia_knl_eval(int32_t kt, int32_t dim, double h, const double* r, int32_t n, double* w, double* dw) {
  if (!r || !w || !dw || n < 0) return SOPHIA_E_ARG;
  if (n == 0) return 0;
  double *dr = nullptr, *dwv = nullptr, *ddw = nullptr;
  HIP_TRY(hipMalloc(&dr, (size_t)n * 8));
  HIP_TRY(hipMalloc(&dwv, (size_t)n * 8));
  HIP_TRY(hipMalloc(&ddw, (size_t)n * 8));
  int rc = 0;
  hipError_t e = hipMemcpy(dr, r, (size_t)n * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    rc = knl_eval_device(nullptr, kt, dim, h, dr, n, dwv, ddw);
    if (!rc) {
      e = hipMemcpy(w, dwv, (size_t)n * 8, hipMemcpyDeviceToHost);
      if (e == hipSuccess) e = hipMemcpy(dw, ddw, (size_t)n * 8, hipMemcpyDeviceToHost);
    }
  }
  if (e != hipSuccess) rc = hip_fail(e, "knl_eval copy", __FILE__, __LINE__);
  (void)hipFree(dr);
  (void)hipFree(dwv);
  (void)hipFree(ddw);
  return rc;
}

int sophia_knl_eval_noalloc(int32_t kt, int32_t dim, double h, const double* r, int32_t n, double* w, double* dw) {
  if (!r || !w || !dw || n < 0) return SOPHIA_E_ARG;
  return sophia::knl_eval_noalloc(kt, dim, h, r, n, w, dw);
}

// learned warm start
int sophia_mlp_param_count(int32_t* n) {
  if (!n) return SOPHIA_E_ARG;
  *n = mlp_param_count();
  return 0;
}
int sophia_mlp_set_params(sophia_ctx* c, const float* w, int32_t n) { CHECK_CTX(c); return c->mlp_set(w, n); }
int sophia_mlp_get_params(const sophia_ctx* c, float* w, int32_t n) {
  if (!c) return SOPHIA_E_ARG;
  return c->mlp_get(w, n);
}
int sophia_mlp_predict(sophia_ctx* c, double* out, int32_t n) { CHECK_CTX(c); return c->mlp_predict(out, n); }
int sophia_mlp_train_step(sophia_ctx* c, const double* target, int32_t n, double lr, double* loss) {
  CHECK_CTX(c);
  return c->mlp_train(target, n, lr, loss, nullptr, 0, true);
}
int sophia_mlp_loss_grad(sophia_ctx* c, const double* target, int32_t n, double* loss, float* grad, int32_t np) {
  CHECK_CTX(c);
  return c->mlp_train(target, n, 0.0, loss, grad, np, false);
}
int sophia_mlp_reset_adam(sophia_ctx* c) { CHECK_CTX(c); return c->mlp_reset(); }
int sophia_mlp_features(sophia_ctx* c, float* nf, int32_t n, double* scale) {
  CHECK_CTX(c);
  return c->mlp_feat(nf, n, scale);
}

int sophia_timers_enable(sophia_ctx* c, int32_t on) {
  CHECK_CTX(c);
  c->timers.off();
  c->timers.reset();
  c->timers.on = on != 0;
  return 0;
}

int sophia_timers_report(sophia_ctx* c, char* buf, int32_t len) {
  CHECK_CTX(c);
  if (!buf || len <= 0) return SOPHIA_E_ARG;
  c->timers.collect();
  std::string out;
  char line[256];
  for (auto& kv : c->timers.acc) {
    std::snprintf(line, sizeof(line), "%s %ld %.6f\n", kv.first.c_str(), kv.second.first, kv.second.second);
    out += line;
  }
  if ((int)out.size() + 1 > len) {
    set_error("timers_report: buffer of %d bytes too small (%zu needed)", len, out.size() + 1);
    return SOPHIA_E_CAPACITY;
  }
  std::memcpy(buf, out.c_str(), out.size() + 1);
  return 0;
}

}  // extern "C"
