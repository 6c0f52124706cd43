/*
 * sophia-mi355x C ABI: the drop-in boundary for the SOPHIA SPH-DEM step.
 *
 * The reference drives one time step through ~33 raw kernel launches from
 * SOPHIA_single_ISPH (ISPH_Calc_seperate.cuh:13-503) or SOPHIA_single
 * (WCSPH_Calc.cuh:4-362), over AoS part1/part2/part3 buffers and the
 * __constant__ parameter vectors k_vii/k_vif (SOPHIA_gpu.cu:43-48).  This
 * header replaces that surface with an opaque context that owns all device
 * memory and exposes the same stages by name:
 *
 *   NNPS               sophia_nnps                (ISPH_Calc_seperate.cuh:44-113)
 *   KNL                sophia_knl_eval            (functions_KNL.cuh:2-173)
 *   PREP               sophia_prep                (ISPH_Calc_seperate.cuh:123-178)
 *   INTERACTION        sophia_interaction         (ISPH_Calc_seperate.cuh:203-206)
 *   DEM_INTERACTION    sophia_dem_interaction     (ISPH_Calc_seperate.cuh:217-218)
 *   SPH_DEM_COUPLING   sophia_sph_dem_coupling    (ISPH_Calc_seperate.cuh:232-249)
 *   TIME               sophia_time                (ISPH_Calc_seperate.cuh:259-269, 391-396)
 *   PPE                sophia_ppe                 (ISPH_Calc_seperate.cuh:305-371)
 *   XSPH               sophia_xsph                (ISPH_Calc_seperate.cuh:448-466)
 *   whole step         sophia_step
 *
 * Every call returns 0 on success or a negative SOPHIA_E* code; it never
 * exits the process (the reference's CUDA_CHECK calls exit, ISPH.cuh:1-8).
 * sophia_last_error() returns a thread-local message for the last failure.
 * One context = one device + one HIP stream; calls on a context are
 * stream-ordered and asynchronous unless documented otherwise; a context
 * must not be used from two threads at once.  No pointer crosses the ABI
 * except caller-owned host arrays.
 */
#ifndef SOPHIA_ABI_H
#define SOPHIA_ABI_H

#include <stddef.h>
#include <stdint.h>
#include "particles.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SOPHIA_ABI_VERSION 1

/* error codes */
#define SOPHIA_OK 0
#define SOPHIA_E_ARG -1       /* bad argument / size */
#define SOPHIA_E_HIP -2       /* HIP runtime failure */
#define SOPHIA_E_IO -3        /* file open / parse failure */
#define SOPHIA_E_STATE -4     /* call out of order (e.g. stage before nnps) */
#define SOPHIA_E_NOMEM -5     /* device allocation failed */
#define SOPHIA_E_CAPACITY -6  /* neighbour/contact slot capacity exceeded */
#define SOPHIA_E_NUMERIC -7   /* NaN/Inf detected or solver breakdown */

/* species */
#define SOPHIA_SPH 0
#define SOPHIA_DEM 1

/* precision of the device state */
#define SOPHIA_PREC_FP64 0   /* every field and all arithmetic in FP64 */
#define SOPHIA_PREC_MIXED 1  /* FP32 fields; FP64 positions, pressure and reductions */

/* kernel ids (Parameters.cuh:38-44) */
#define SOPHIA_KNL_GAUSSIAN 0
#define SOPHIA_KNL_QUINTIC 1
#define SOPHIA_KNL_QUARTIC 2
#define SOPHIA_KNL_WENDLAND2 3
#define SOPHIA_KNL_WENDLAND4 4
#define SOPHIA_KNL_WENDLAND6 5

/* particle types (Parameters.cuh:10-24) */
#define SOPHIA_PT_BOUNDARY 0
#define SOPHIA_PT_FLUID 1
#define SOPHIA_PT_MOVING 9
#define SOPHIA_PT_DEM 1001

/*
 * Run parameters.  The first block mirrors the solv.txt keys parsed by
 * read_solv_input (function_init.cuh:216-615) into vii/vif
 * (function_init.cuh:2-105).  The second block holds model constants that
 * the reference hard-codes; sophia_params_default() sets them to the
 * reference values so that parity runs reproduce it, and configs override
 * them.
 */
typedef struct sophia_params {
  /* --- solv.txt --- */
  int32_t solver_type;        /* 0 WCSPH, 1 ISPH */
  int32_t dim;                /* 2 or 3 */
  int32_t kernel_type;        /* SOPHIA_KNL_* */
  int32_t rho_type;           /* 0 Direct (mass sum), 1 Continuity */
  int32_t time_type;          /* 0 Euler, 1 Predictor_Corrector */
  int32_t open_boundary;      /* 0 NO, 1 OPEN, 2 PERIODIC */
  int32_t flag_z_index;       /* 1: Morton cell keys */
  int32_t freq_filt;          /* Shepard filter frequency */
  int32_t freq_output;
  int32_t fp_solve, fv_solve, fva_solve, fg_solve;
  int32_t con_solve;          /* conduction */
  int32_t xsph_solve;
  int32_t kgc_solve;          /* 0 NO, 1 KGC */
  int32_t delSPH_solve;       /* 0 NO, 1 Molteni, 2 Antuono */
  int32_t decouple_stride;
  int32_t min_iteration, max_iteration;
  double kappa, p_ref, gamma, dt, time, time_end;
  double xmargin_m, xmargin_p, ymargin_m, ymargin_p, zmargin_m, zmargin_p;
  double u_limit, c_xsph, soundspeed, rho0_eos;
  double drho_th, dp_th, p_relaxation;

  /* --- model constants hard-coded in the reference --- */
  double rho_fluid;      /* DENSITY_AIR (Parameters.cuh:53): m_ref and b_i scale, p_type 0/1/3/-3 */
  double rho_water;      /* DENSITY_WATER (Parameters.cuh:52): p_type 9 and others */
  double h_over_dx;      /* 1.6 in vol = (h/1.6)^dim (functions_MASS.cuh:33) */
  double mu_fluid;       /* viscosity() -> VISCOSITY_AIR (functions_PROP.cuh:240-242) */
  double mu_water;       /* VISCOSITY_WATER (mixture drag branch) */
  double k_fluid;        /* conductivity() SPH value 0.0518 (functions_PROP.cuh:272) */
  double k_dem;          /* conductivity() DEM value 2.5 */
  double cp_fluid;       /* heat_capacity() SPH value 1006.4 */
  double cp_dem;         /* heat_capacity() DEM value 1068 */
  double gravity;        /* 9.8 (Parameters.cuh:81) */
  double eos_b_factor;   /* 1.186 in B = c0^2*1.186/gamma (functions_PROP.cuh:22) */
  double delta_sph;      /* delta = 0.1 (Parameters.cuh:93) */
  double dem_shear_modulus;  /* G = 8.3e8 (functions_PROP.cuh:507) */
  double dem_poisson;        /* nu = 0.23 (functions_PROP.cuh:492) */
  double dem_friction;       /* Coulomb cap 0.5*F_n (functions_DEM_INTERACTION2.cuh:1528) */
  double dem_damping;        /* sqrt(3.333333333)*0.0535184 (e = 0.9) */
  double dem_hertz_factor;   /* 1.33333 (functions_DEM_INTERACTION2.cuh:1465) */
  double drag_switch;        /* 0.65 fluid-fraction threshold (functions_SPH_DEM_COUPLING2.cuh:979) */
  int32_t eulerian_sph;      /* 1: SPH particles fixed (elix=0, functions_ALE.cuh:35-39) */
  int32_t fix_xsph_position; /* 0: reference (XSPH never writes x); 1: x from XSPH velocity */
  int32_t dem_max_contacts;  /* history slots per sphere, <= 16 */
  int32_t sph_max_neighbors; /* ELL width of the SPH neighbour list */
  /* optional domain override (periodic boxes); used when domain_fixed != 0 */
  int32_t domain_fixed;
  int32_t periodic_mask;     /* bit0 x, bit1 y, bit2 z: periodic wrap in that axis */
  double domain_min[3], domain_max[3];
  /* pressure-gradient pair term of the pressure force and of pgf:
   * 0: sum_j m_j (p_i + p_j)/(rho_i rho_j) grad~W_ij (reference,
   *    functions_PPE2.cuh:2303-2315, :1445-1450);
   * 1: sum_j m_j (p_j - p_i)/(rho_i rho_j) grad~W_ij (zero for a constant
   *    pressure; keeps a converged PPE solve stable, see the design notes) */
  int32_t pgrad_form;
  /* DEM wall boundaries after the DEM time update (ISPH_Calc_seperate.cuh:
   * 439-441), applied in this order: bit 0 cylinder about z
   * (KERNEL_treat_DEM_cylinder_z_dem, functions_DEM_BC.cuh:491-729), bit 1
   * box (KERNEL_treat_DEM_box_z_dem, :1770-2027), bit 2 truncated cone
   * about z (KERNEL_treat_DEM_truncated_cone_z_dem, :4485-4725).  The
   * reference calls all three every step with hard-coded geometry; here they
   * are off by default (0) and the geometry defaults to those values. */
  int32_t dem_walls;
  double wall_cyl[3];   /* centre x, y; radius (xz_cylinder, yz_cylinder, rz_cylinder: :254-257) */
  double wall_box[6];   /* centre x, y, z; side lengths x, y, z (xz_box .. Lzz_box: :1503-1509) */
  double wall_cone[5];  /* centre x, y; top radius r1, bottom radius r2, height h (:723-724, 4240-4242) */
  double wall_rest[2];  /* restitution: normal rest_n, tangential rest_s (:1-2) */
  double wall_mu[2];    /* friction: sliding mu_s, rolling mu_rs (:4-5) */
  /* open boundaries (open_boundary = 1): z bands [L1, L2) inlet and
   * [L3, L4) outlet (functions_NNPS.cuh:2-5, 26-37), inlet velocity and
   * temperature that the reference hard-codes (functions_BC.cuh:961-964) */
  double ob_band[4];
  double ob_inlet_u;
  double ob_inlet_temp;
  /* Reference quirks reproduced on request (bit set = as the reference;
   * 0 = corrected, the default):
   *   bit 0 SOPHIA_COMPAT_NO_SPIN: the ISPH step never integrates DEM spin
   *         from the contact torque (KERNEL_time_update_projection_dem,
   *         functions_TIME_ISPH2.cuh:624-776, has no omega update; the WCSPH
   *         corrector does, functions_TIME.cuh:753-755);
   *   bit 1 SOPHIA_COMPAT_NUSSELT: pow(Pr, 1/3) with integer 1/3 = Pr^0
   *         (functions_SPH_DEM_COUPLING2.cuh:998, 1042, 1082);
   *   bit 2 SOPHIA_COMPAT_HEAT_OVERWRITE: the coupling's convective rate
   *         replaces the DEM contact/radiative rate instead of adding to it
   *         (:1103 writes P3_dem.dtemp after KERNEL_DEM_interaction3D_dem
   *         set it, functions_DEM_INTERACTION2.cuh:1725-1729);
   *   bit 3 SOPHIA_COMPAT_CP_ARGS: heat_capacity(p_type, temp) with swapped
   *         arguments in the SPH reaction heat (:1860). */
  int32_t reference_compat;
  /* heat transfer (con_solve = 1) */
  double pr_fluid;          /* Pr of the fluid-regime Nusselt number (0.7, :997) */
  double pr_air;            /* Pr of the air and mixture regimes (1.0, :1041, :1081) */
  double dem_k_mat[2];      /* sphere conductivity of material 0, 1 (0.84, 55.0: functions_PROP.cuh:279-308) */
  double dem_alpha_mat[2];  /* sphere diffusivity of material 0, 1 (2.5e-6, 1.77e-5: :312-340) */
  double dem_mat_radius;    /* spheres with rad < this are material 1 (0.0012, functions_SPH_DEM_COUPLING2.cuh:708) */
  /* wall heat flux into the fluid, W/m^2 (0 = off; no reference
   * counterpart in the compiled kernels, HEAT_FLUX = 0 at
   * functions_DEM_INTERACTION2.cuh:1).  Wall particles (p_type 0) inside
   * heat_flux_box [xmin ymin zmin xmax ymax zmax] are flux walls of face
   * area (h / h_over_dx)^2 each; their power is spread over the fluid by
   * the kernel, normalised by each wall's fluid Shepard sum, so the fluid
   * receives exactly heat_flux times the total face area. */
  double heat_flux;
  double heat_flux_box[6];
  /* free surface: fluid rows detected as surface (KERNEL_clc_surface_detect3D,
   * functions_PST.cuh:263-320: |mean unit vector to neighbours| > fs_threshold
   * or fewer than 10 neighbours) take the Dirichlet row p = p_ref, the
   * commented-out lines of functions_PPE2.cuh:533-534.  Neighbours counted
   * are every SPH particle within kappa*h (walls included), where the
   * reference counts only its own p_type (which marks fluid next to walls
   * as surface).  0 = off (reference: no surface rows). */
  int32_t free_surface;
  double fs_threshold;      /* 0.3 (functions_PST.cuh:319) */
  /* property_table(YES/NO) of solv.txt (function_init.cuh:261-262, vii[2]):
   * 1 = viscosity and conductivity of SPH particles whose |p_type| has a
   * table in data.txt are interpolated in temperature (viscosity() and
   * conductivity(), functions_PROP.cuh:228-264); needs the table, passed
   * to sophia_create or sophia_set_table.  Types without a table (spheres,
   * |p_type| >= 10) keep the constants. */
  int32_t prop_table;
  /* noslip_boundary(YES/NO) of solv.txt (function_init.cuh:446-450,
   * vii[51]): 2D walls take the mirrored fluid velocity before the
   * interaction (KERNEL_boundary2D, functions_BC.cuh:1234-1328) */
  int32_t noslip_bc;
  double reserved[7];
} sophia_params;

#define SOPHIA_COMPAT_NO_SPIN 1
#define SOPHIA_COMPAT_NUSSELT 2
#define SOPHIA_COMPAT_HEAT_OVERWRITE 4
#define SOPHIA_COMPAT_CP_ARGS 8
#define SOPHIA_COMPAT_ALL 15

/* Domain geometry after find_minmax (function_init.cuh:913-950) and the
 * cell sizing of ISPH.cuh:93-113. */
typedef struct sophia_domain {
  double x_min, x_max, y_min, y_max, z_min, z_max;
  double dcell;
  int32_t NI, NJ, NK;
  int32_t num_cells;      /* cells addressable by a key (+1 dummy slot) */
  int32_t n_sph, n_dem;
} sophia_domain;

/* PPE solve selection */
#define SOPHIA_PPE_JACOBI 0     /* reference: one Jacobi sweep (functions_PPE2.cuh:868-878) */
#define SOPHIA_PPE_BICGSTAB 1   /* converged Krylov solve */

#define SOPHIA_PPE_FP64 0       /* FP64 operator, FP64 vectors */
#define SOPHIA_PPE_MIXED_IR 1   /* FP32 vectors and operator, folded into an FP64 solution
                                   with FP64 residuals (reliable updates) */

#define SOPHIA_GUESS_ZERO 0
#define SOPHIA_GUESS_PREVIOUS 1
#define SOPHIA_GUESS_LEARNED 2  /* extrapolated guess plus a learned correction; falls back
                                   to the extrapolated guess if its residual is larger */
#define SOPHIA_GUESS_EXTRAPOLATE 3 /* 3 p^n - 3 p^(n-1) + p^(n-2) from the last three solutions
                                      (2 p^n - p^(n-1) with two, p^n with one) */

#define SOPHIA_PRECOND_JACOBI 0    /* D^-1 (the reference's diagonal) */
#define SOPHIA_PRECOND_TWO_LEVEL 1 /* D^-1 + Z E^-1 Z^T: additive coarse correction over
                                      aggregates of 3x3x3 cells, E = Z^T A Z (row-major
                                      cell keys; Jacobi otherwise) */

typedef struct sophia_ppe_opts {
  int32_t method;       /* SOPHIA_PPE_JACOBI / SOPHIA_PPE_BICGSTAB */
  int32_t precision;    /* SOPHIA_PPE_FP64 / SOPHIA_PPE_MIXED_IR */
  int32_t guess;        /* SOPHIA_GUESS_* */
  int32_t max_iter;     /* total inner iterations cap */
  int32_t max_outer;    /* restarts after a Krylov breakdown */
  int32_t zero_mean;    /* project out the constant null space (periodic, no walls) */
  double rtol;          /* ||b - A p|| <= rtol * ||b|| */
  double inner_rtol;    /* mixed: fold the FP32 iterate into the FP64 solution and
                           recompute the FP64 residual each time the FP32 residual
                           has dropped by this factor (0 -> 0.01) */
  int32_t train;        /* 1: after the solve, one Adam step of the warm-start
                           network on (features before the solve, solution) */
  int32_t precond;      /* SOPHIA_PRECOND_* (BiCGStab) */
  double lr;            /* Adam learning rate when train != 0 */
} sophia_ppe_opts;

typedef struct sophia_ppe_stats {
  int32_t iterations;       /* inner Krylov iterations */
  int32_t outer;            /* FP64 residual recomputations (folds) */
  int32_t converged;
  int32_t guess_used;       /* SOPHIA_GUESS_* actually used (after fallback) */
  double r0_rel;            /* ||b - A p0|| / ||b|| of the guess used */
  double r0_rel_learned;    /* same for the learned guess (if tried) */
  double r_rel;             /* final relative residual (FP64) */
  double ms;                /* wall time of the PPE stage, device events */
  int32_t n_rows;
  int32_t nnz;              /* stored operator entries (row columns) */
  double loss;              /* training loss of the in-solve Adam step */
  int32_t precond_used;     /* SOPHIA_PRECOND_* actually used */
  int32_t n_coarse;         /* coarse unknowns (two-level) */
  double coarse_res;        /* ||I - E X||_F / sqrt(n_coarse) of the coarse inverse X used */
  double r0_rel_base;       /* same for the extrapolated guess the learned one corrects (if tried) */
} sophia_ppe_stats;

typedef struct sophia_ctx sophia_ctx;

/* --- parameters and files ------------------------------------------------ */
void sophia_params_default(sophia_params* p);
/* solv.txt parser (function_init.cuh:216-615); unknown keys are ignored */
int sophia_read_solv(const char* path, sophia_params* p);
/* input.txt: header row of column codes 1-36, tab separated
 * (function_init.cuh:704-891).  With out == NULL only counts rows into *n. */
int sophia_read_input(const char* path, part1* out, int32_t capacity, int32_t* n);
/* p_type.txt: one integer per line (ISPH.cuh via gpu_count_boundary_numbers) */
int sophia_read_ptype(const char* path, int32_t* out, int32_t capacity, int32_t* n);

/* data.txt property tables (read_table, function_init.cuh:966-1182): blocks
 * "#p_type N" ... "#end", each with comma-separated rows after the lines
 * "//T", "//h", "//k" and "//mu" (temperature K, enthalpy J/kg,
 * conductivity W/m/K, viscosity Pa s), stored one after the other.  A
 * block's length is its shortest row: the reference takes the last row's
 * count (count2, :1108), and its own data.txt has a 6-value T row over
 * 5-value rows, whose sixth value the next block overwrites.  Blocks are
 * filed under their own p_type (0..9); the reference files them by reading
 * order (:1117-1124), which for its data.txt leaves the "#p_type 9" block
 * with an unset size.  At most SOPHIA_TABLE_VALUES values over all blocks
 * (table_size, Parameters.cuh:131); types with no block have size 0. */
#define SOPHIA_TABLE_VALUES 50
#define SOPHIA_TABLE_TYPES 10
typedef struct sophia_table {
  int32_t n_tables;                   /* number of blocks read */
  int32_t n_values;                   /* values over all blocks */
  int32_t index[SOPHIA_TABLE_TYPES];  /* first value of the block of p_type t (host_table_index) */
  int32_t size[SOPHIA_TABLE_TYPES];   /* values in that block (host_table_size) */
  double T[SOPHIA_TABLE_VALUES];      /* host_Tab_T */
  double h[SOPHIA_TABLE_VALUES];      /* host_Tab_h */
  double k[SOPHIA_TABLE_VALUES];      /* host_Tab_k */
  double mu[SOPHIA_TABLE_VALUES];     /* host_Tab_vis */
} sophia_table;
int sophia_read_table(const char* path, sophia_table* out);
/* one property of |p_type| at x by the reference's interp2
 * (functions_PROP.cuh:143-175: piecewise linear, extended linearly beyond
 * both ends): what = SOPHIA_PROP_MU, _K (of temperature), _H (enthalpy of
 * temperature, Ttoh) or _T (temperature of enthalpy, htoT).  A type with
 * no block is SOPHIA_E_ARG. */
#define SOPHIA_PROP_MU 0
#define SOPHIA_PROP_K 1
#define SOPHIA_PROP_H 2
#define SOPHIA_PROP_T 3
int sophia_table_eval(const sophia_table* t, int32_t p_type, int32_t what, double x, double* y);

/* --- context ------------------------------------------------------------- */
/* particles in reference order, SPH and DEM mixed; split by p_type <= 1000
 * (read_sph_dem, function_init.cuh:893-911). */
int sophia_create_from_arrays(const sophia_params* p, const part1* particles,
                              int32_t n, int32_t device, int32_t precision,
                              sophia_ctx** out);
/* The reference's start-up from its input directory (main,
 * SOPHIA_gpu.cu:145-151; ISPH.cuh): solv.txt, input.txt, and optionally
 * p_type.txt (NULL to skip; when given it must list the p_type column of
 * input.txt, which the reference uses to count boundary particles,
 * gpu_count_boundary_numbers) and data.txt (NULL to skip; required when
 * solv.txt asks for property_table YES). */
int sophia_create(const char* solv_path, const char* input_path, const char* ptype_path,
                  const char* data_path, int32_t device, int32_t precision, sophia_ctx** out);
/* install property tables on a context (used when prop_table = 1) */
int sophia_set_table(sophia_ctx* ctx, const sophia_table* t);
int sophia_destroy(sophia_ctx* ctx);
const char* sophia_last_error(void);
int sophia_get_domain(const sophia_ctx* ctx, sophia_domain* d);
int sophia_get_params(const sophia_ctx* ctx, sophia_params* p);
int sophia_sync(sophia_ctx* ctx);
/* step counter and simulated time (count, time in the reference vii/vif) */
int sophia_get_clock(const sophia_ctx* ctx, int64_t* count, double* time);

/* --- stages (ISPH-DEM order of SOPHIA_single_ISPH) ------------------------ */
int sophia_nnps(sophia_ctx* ctx);
int sophia_prep(sophia_ctx* ctx);
int sophia_interaction(sophia_ctx* ctx);
int sophia_dem_interaction(sophia_ctx* ctx);
int sophia_sph_dem_coupling(sophia_ctx* ctx);
#define SOPHIA_PREDICT 0
#define SOPHIA_CORRECT 1
int sophia_time(sophia_ctx* ctx, int32_t phase);
int sophia_ppe(sophia_ctx* ctx, const sophia_ppe_opts* opts, sophia_ppe_stats* stats);
int sophia_xsph(sophia_ctx* ctx);
/* one full time step; advances count and time */
int sophia_step(sophia_ctx* ctx, const sophia_ppe_opts* opts, sophia_ppe_stats* stats);
/* n steps back to back (stats of the last step) */
int sophia_run(sophia_ctx* ctx, int32_t n, const sophia_ppe_opts* opts, sophia_ppe_stats* stats);

/* --- inspection / parity hooks (synchronous) ----------------------------- */
/* sorted-order cell keys and the permutation applied this step (reference
 * g_idx / p_idx), cell ranges over num_cells (g_str / g_end; empty =
 * 0xffffffff) and, for SPH, per-particle neighbour counts within kappa*h_i.
 * Any pointer may be NULL. */
int sophia_get_nnps(const sophia_ctx* ctx, int32_t species, int32_t* keys,
                    int32_t* order, uint32_t* cell_start, uint32_t* cell_end,
                    int32_t* nbr_count);
/* export the state in sorted order as reference records (any may be NULL) */
int sophia_get_particles(const sophia_ctx* ctx, int32_t species, part1* p1,
                         part2* p2, part3* p3);
/* One stage of the WCSPH step (solver_type 0; SOPHIA_single, WCSPH_Calc.cuh:
 * 4-340), fluid only, 3D, continuity density: 0 set_alpha + predictor
 * (KERNEL_clc_predictor, functions_TIME.cuh:77-292) + NNPS at step 0 / P3
 * reset later, 1 KERNEL_clc_volume_update3D + KERNEL_clc_mass_update
 * (functions_MASS.cuh:148-222, 45-60), 2 KGC + KERNEL_clc_prep3D_coupling +
 * KERNEL_clc_prep3D, 3 KERNEL_clc_continuity_norm3D (functions_MASS.cuh:
 * 771-914), 4 KERNEL_EOS (functions_PROP.cuh:7-28), 5
 * KERNEL_Neumann_boundary3D (functions_BC.cuh:2-137), 6 KERNEL_interaction3D
 * (functions_INTERACTION.cuh:413-832), 7 KERNEL_time_update_single
 * (functions_TIME.cuh:676-877), 8 KERNEL_xsph3D + clock.  sophia_step runs
 * stages 0-8 when solver_type is 0. */
int sophia_wcsph_stage(sophia_ctx* ctx, int32_t stage);
/* run monitors of one species (kernel_copy_max_sph/_dem + cub::DeviceReduce::Max,
 * functions_OUTPUT.cuh:35-65, ISPH_Calc_seperate.cuh:928-969): out[0] = max |u|,
 * out[1] = max rho, out[2] = max P3.ftotal, over particles with p_type != 0 and
 * i_type < 2 (others count as 0) */
int sophia_monitors(sophia_ctx* ctx, int32_t species, double out[3]);
/* number of particles of a species */
int sophia_count(const sophia_ctx* ctx, int32_t species, int32_t* n);
/* Output files of the current state, in sorted order as the reference writes
 * its arrays.  Legacy VTK 3.0 binary POLYDATA, big-endian 4-byte words
 * (FloatSwap / IntSwap, functions_OUTPUT.cuh:68-121), one kind per writer of
 * that file (species read, particles written, arrays):
 *   SOPHIA_VTK_FLUID      save_plot_fluid_vtk_bin_fluid_sph (:445-659): SPH,
 *                         p_type 1, 9 arrays;
 *   SOPHIA_VTK_DEM        save_plot_fluid_vtk_bin_solid_dem (:1135-1446):
 *                         spheres (p_type > 1000; the point count is the
 *                         spheres with dem_idx > 0, as the reference counts),
 *                         10 arrays;
 *   SOPHIA_VTK_AIR        save_plot_fluid_vtk_bin_air (:660-822): SPH,
 *                         p_type 3, 10 arrays;
 *   SOPHIA_VTK_SOLID      save_plot_fluid_vtk_bin_solid (:824-1134): spheres,
 *                         counted as SOPHIA_VTK_DEM, 10 arrays (dem_idx and
 *                         drag_r where the other has itype and Euler);
 *   SOPHIA_VTK_MOVING     save_plot_fluid_vtk_bin_moving (:1447-1578): SPH,
 *                         p_type 9 (MOVING), density and pressure;
 *   SOPHIA_VTK_FLUID2     save_plot_fluid_vtk_bin_fluid2 (:1579-1702): SPH,
 *                         p_type 2, 7 arrays, three of them 3-component;
 *   SOPHIA_VTK_BOUNDARY   save_plot_fluid_vtk_bin_boundary (:1703-1841): SPH,
 *                         p_type 0, 6 arrays;
 *   SOPHIA_VTK_FLUID_PGF  save_plot_fluid_vtk_bin_fluid (:202-444): SPH,
 *                         p_type 1, 12 arrays with the pressure gradient;
 *   SOPHIA_VTK_ACTIVE     save_plot_fluid_vtk_bin (:126-201): SPH, i_type 1
 *                         or 2, rho, ptype (int) and pres;
 *   SOPHIA_VTK_MOVING_UY  save_plot_moving_vtk_bin (:1842-1925): SPH,
 *                         p_type 9, rho and uy.
 * Unknown kinds return SOPHIA_E_ARG. */
#define SOPHIA_VTK_FLUID 0
#define SOPHIA_VTK_DEM 1
#define SOPHIA_VTK_AIR 2
#define SOPHIA_VTK_SOLID 3
#define SOPHIA_VTK_MOVING 4
#define SOPHIA_VTK_FLUID2 5
#define SOPHIA_VTK_BOUNDARY 6
#define SOPHIA_VTK_FLUID_PGF 7
#define SOPHIA_VTK_ACTIVE 8
#define SOPHIA_VTK_MOVING_UY 9
int sophia_write_vtk(const sophia_ctx* ctx, const char* path, int32_t kind);
/* restart.txt of save_restart (functions_OUTPUT.cuh:1926-1955): a header of
 * input column codes "1 2 3 4 5 6 7 8 9 10 11 12 13 26" and one line per
 * particle (SPH, then spheres): x y z ux uy uz (%f), m (%e), p_type (%d),
 * h (%e), temp pres rho rho_ref flt_s (%f), tab-separated.  sophia_create
 * reads it back as an input file (column 13 is skipped, as the reference's
 * reader does). */
int sophia_write_restart(const sophia_ctx* ctx, const char* path);
/* evaluate W and dW/dr on the device (calc_kernel_wij / calc_kernel_dwij) */
int sophia_knl_eval(int32_t kernel_type, int32_t dim, double h, const double* r,
                    int32_t n, double* w, double* dw);
/* The same evaluation with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Staging
 * for r, W and dW/dr is a file-scope device array of 3 x 65536 doubles
 * (1.5 MiB per device, held until the process ends); n > 65536 returns
 * SOPHIA_E_CAPACITY.  Every call on one device shares that array, so calls
 * that overlap on one device (from several host threads) race. */
#define SOPHIA_KNL_NOALLOC_MAX 65536
int sophia_knl_eval_noalloc(int32_t kernel_type, int32_t dim, double h, const double* r,
                            int32_t n, double* w, double* dw);
/* overwrite the SPH pressure (sorted order) e.g. to pose a PPE problem */
int sophia_set_pressure(sophia_ctx* ctx, const double* p, int32_t n);
/* PPE system of the current step in sorted order: diag = sum_j a_ij,
 * rhs b_i (PPE3 + wall part), row mask (1 = unknown).  Any may be NULL. */
int sophia_get_ppe_system(const sophia_ctx* ctx, double* diag, double* rhs,
                          int32_t* row_mask);
/* y = A x with the assembled operator (FP64), sorted order, SPH count */
int sophia_ppe_apply(sophia_ctx* ctx, const double* x, double* y, int32_t n);
/* Timing probe of the inner FP32 matvec of the mixed-precision solve (no
 * reference counterpart): variant 0 runs the solver's kernel on the stored
 * operator, variant 1 a matrix-free kernel that recomputes every a_ij from
 * the neighbour list and the position records.  y = A x (FP32 arithmetic,
 * sorted order, SPH count) and the mean device time of `reps` launches after
 * one untimed one.  Needs a mixed-precision context with an assembled
 * operator in run-offset format (row-major cell keys). */
int sophia_ppe_matvec_probe(sophia_ctx* ctx, int32_t variant, int32_t reps, const double* x, double* y,
                            int32_t n, double* us_per_launch);
/* storage of the assembled operator: stored entries, bytes per column index
 * (1: u8 offset inside a cell run, 4: u32 index, 0: the u8 offset rides in
 * the low byte of the FP32 inner coefficient, whose significand is rounded
 * to 16 bits) and per coefficient as read by the inner iterations (4: FP32,
 * 8: FP64).  Any may be NULL. */
int sophia_ppe_layout(const sophia_ctx* ctx, int64_t* nnz, int32_t* index_bytes, int32_t* coef_bytes);
/* Coarse operator of the last two-level solve (no reference counterpart;
 * inspection): aggregates per axis into dims[3] (n_coarse = product), and,
 * where non-NULL, E = Z^T A Z and its approximate inverse X, row-major
 * n_coarse x n_coarse FP32 (aggregates without rows have E_II = 1). */
int sophia_ppe_coarse(const sophia_ctx* ctx, int32_t dims[3], float* E, float* X);

/* --- learned PPE warm start (message-passing MLP, BF16 MFMA) ------------- */
/* The network predicts the pressure change of the PPE rows from the
 * residual of the previous pressure: p0 = p_prev + s * y, where y is the
 * network output, s = rms(r)/mean(diag A), r = b - A p_prev.  Edge inputs
 * over the SPH neighbour list: (x_i - x_j)/h_i (3), |x_i - x_j|/h_i,
 * rho_j/rho_fluid, r_j/rms(r).  Edge MLP 6 -> 64 -> 64 (ReLU, BF16 weights
 * on MFMA, FP32 accumulation), sum aggregation scaled by 1/64, node MLP
 * 64 -> 64 -> 1.  Parameters are an FP32 master vector (layout in
 * learned.hpp); the kernels use BF16 copies of the weight matrices.  All
 * calls below need an assembled PPE system of the current step (they run
 * after sophia_ppe, or inside it through the opts). */
#define SOPHIA_MLP_HIDDEN 64
#define SOPHIA_MLP_EDGE_IN 6
int sophia_mlp_param_count(int32_t* n);
int sophia_mlp_set_params(sophia_ctx* ctx, const float* params, int32_t n);
int sophia_mlp_get_params(const sophia_ctx* ctx, float* params, int32_t n);
/* prediction p0 for the current PPE rows with the current pressure as
 * p_prev (sorted order, SPH count; non-rows return p_prev) */
int sophia_mlp_predict(sophia_ctx* ctx, double* out, int32_t n);
/* one Adam step against target pressure (sorted order); loss before step */
int sophia_mlp_train_step(sophia_ctx* ctx, const double* target, int32_t n, double lr, double* loss);
/* loss and gradient w.r.t. the FP32 parameters, no update */
int sophia_mlp_loss_grad(sophia_ctx* ctx, const double* target, int32_t n, double* loss, float* grad,
                         int32_t n_params);
int sophia_mlp_reset_adam(sophia_ctx* ctx);
/* node inputs the network sees (n x 2: rho_j/rho_fluid, r_j/rms(r)) and the
 * output scale s */
int sophia_mlp_features(sophia_ctx* ctx, float* node_feat, int32_t n, double* scale);

/* --- device timers (HIP events around stage kernels) ---------------------- */
/* enable (1) or disable (0); enabling clears the accumulated totals */
int sophia_timers_enable(sophia_ctx* ctx, int32_t on);
/* "name launches total_ms" lines, NUL terminated */
int sophia_timers_report(sophia_ctx* ctx, char* buf, int32_t len);

#ifdef __cplusplus
}
#endif

#endif /* SOPHIA_ABI_H */
