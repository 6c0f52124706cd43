/*
 * FP64 CPU oracle of the SOPHIA SPH-DEM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C++/OpenMP restatement of the
 * reference kernels on the reference's own AoS records (part1/part2/part3),
 * in FP64 with no FMA contraction.  Only the test suite, the smoke check and
 * the CPU-baseline leg of the benchmark may load it, and only as the checker
 * or the reported CPU baseline; the product library never links it.
 *
 * Every stage cites the reference file:line it restates.  Parity pins for
 * each function are listed in the header of oracle.cpp.
 */
#ifndef SOPHIA_ORACLE_H
#define SOPHIA_ORACLE_H

#include "../include/sophia/abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

orc_state* orc_create(const sophia_params* p, const part1* particles, int32_t n);
void orc_destroy(orc_state* s);
/* data.txt property tables, used when p->prop_table = 1 */
void orc_set_table(orc_state* s, const sophia_table* t);
void orc_get_domain(const orc_state* s, sophia_domain* d);
void orc_get_clock(const orc_state* s, int64_t* count, double* time);
int orc_count(const orc_state* s, int32_t species);

/* stages, ISPH-DEM order (ISPH_Calc_seperate.cuh:13-503) */
void orc_nnps(orc_state* s);
void orc_prep(orc_state* s);
void orc_interaction(orc_state* s);
void orc_dem_interaction(orc_state* s);
void orc_sph_dem_coupling(orc_state* s);
void orc_time(orc_state* s, int32_t phase);
int orc_ppe(orc_state* s, const sophia_ppe_opts* o, sophia_ppe_stats* st);
void orc_xsph(orc_state* s);
int orc_step(orc_state* s, const sophia_ppe_opts* o, sophia_ppe_stats* st);
/* WCSPH (solver_type 0) stages of SOPHIA_single (WCSPH_Calc.cuh:4-340):
 * 0 set_alpha + predictor + NNPS (count 0) or P1 -> SP1 copy, 1 volume and
 * mass update, 2 KGC + porosity + prep, 3 continuity, 4 EOS, 5 wall
 * Neumann, 6 interaction, 7 corrector, 8 XSPH + clock.  -1 when the state
 * is outside the supported set (DEM present, 2D, mass summation). */
int orc_wcsph_stage(orc_state* s, int32_t stage);

/* inspection */
void orc_get_nnps(const orc_state* s, int32_t species, int32_t* keys, int32_t* order,
                  uint32_t* cell_start, uint32_t* cell_end, int32_t* nbr_count);
void orc_get_particles(const orc_state* s, int32_t species, part1* p1, part2* p2, part3* p3);
void orc_set_pressure(orc_state* s, const double* p, int32_t n);
/* PPE system of the last assembly (sorted order) */
void orc_get_ppe_system(const orc_state* s, double* diag, double* rhs, int32_t* row_mask);
void orc_ppe_apply(const orc_state* s, const double* x, double* y);

/* kernel functions (functions_KNL.cuh:2-173) */
double orc_knl_tmpA(int32_t kernel_type, int32_t dim, double h);
void orc_knl_eval(int32_t kernel_type, int32_t dim, double h, const double* r, int32_t n,
                  double* w, double* dw);
/* cell key of one position (functions_NNPS.cuh:776-790, 378-390) */
int32_t orc_cell_key(const orc_state* s, double x, double y, double z);
/* O(N^2) neighbour counts of SPH particles (sorted order) with the
 * reference membership test sqrt(r2)+1e-20 < kappa*h_i over all SPH j,
 * without the cell restriction */
void orc_brute_neighbor_counts(const orc_state* s, int32_t* counts);

/* threads used by the OpenMP loops (0 = runtime default) */
void orc_set_threads(int32_t n);
int32_t orc_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
