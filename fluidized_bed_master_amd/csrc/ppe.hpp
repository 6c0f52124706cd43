// PPE operator storage, workspace and entry points (ppe.hip).
#pragma once

#include "common.hpp"
#include "timers.hpp"

namespace sophia {

// Column encodings of the compact operator.
//   kFmtRun8 : row-major cell keys.  The 27-cell stencil of a particle is 9
//              runs of 3 x-adjacent cells, each a contiguous range of the
//              cell-sorted particle order; a column is stored as a u8 offset
//              from its run's first particle, and a u64 per row holds the 9
//              run lengths (7 bits each).  Needs every run <= 256 particles.
//   kFmtIdx32: plain u32 column indices (Morton keys or overfull runs).
enum { kFmtRun8 = 0, kFmtIdx32 = 1 };
constexpr int kRunBits = 7;  // kFmtRun8: bits per run length in the row's u64
constexpr int kRunMax = (1 << kRunBits) - 1;

// Compact operator: per row only the columns that are PPE rows themselves
// (wall/buffer columns live in the right-hand side), slot-quad major:
// entry k of row i sits at ((k / 4) * n + i) * 4 + k % 4, so a wave reads
// 64 consecutive 16 B (FP32) or 32 B (FP64) quads per four slots.
// Mixed path: the inner coefficient is fl32(a) rounded to nearest at a
// 16-bit significand (the low 8 bits of the FP32 word cleared), and with
// kFmtRun8 the entry's u8 column offset rides in those 8 bits, so an entry
// is one 4-byte word.  The inner diagonal takes up the rounding of its row,
// dgi_i = d_i - sum_j (a_ij - hi16(a_ij)): the inner operator then maps a
// constant vector exactly as the FP64 one does, and differs from it by
// sum_j (a_ij - hi16(a_ij)) (x_i - x_j) only (relative 2^-16 of a
// difference).  hi + lo keeps a to about 2^-40 for the FP64 residual.
__device__ inline uint32_t hi16_bits(double a) { return (__float_as_uint((float)a) + 0x80u) & 0xFFFFFF00u; }
__device__ inline float hi_val(uint32_t e) { return __uint_as_float(e & 0xFFFFFF00u); }

struct PpeOp {
  int n = 0, nq = 0, fmt = kFmtRun8;
  int NI = 1, NJ = 1, ncell = 1;
  const uint32_t* keys = nullptr;      // sorted cell key per particle
  const uint32_t* cell_off = nullptr;  // first particle of each cell (dense, num_cells + 1)
  const int* row = nullptr;
  const int* ccnt = nullptr;           // compact entries per row
  const uint64_t* runc = nullptr;      // kFmtRun8 run lengths
  const uint32_t* idx = nullptr;       // kFmtRun8 FP64 path: u32 [q][i] (4 offsets); kFmtIdx32: u32 [q][i][4]
  const float* hi = nullptr;           // mixed: hi16(a) | offset    [q][i][4] (hi_bits / hi_val)
  const float* lo = nullptr;           // mixed: fl32(a - hi16(a))   [q][i][4]
  const double* c64 = nullptr;         // FP64 path                  [q][i][4]
  const double* diag = nullptr;
  const float* dgi = nullptr;          // mixed: inner diagonal d_i - sum_j (a_ij - hi16(a_ij))
  const double* dinv = nullptr;
  const double* rhs = nullptr;
};

struct PpeWork {
  int n = 0, cap4 = 0, nbg = 0;
  int64_t nnz = 0;  // stored entries of the last assembly
  bool mixed = false;
  int fmt = kFmtRun8;
  int NI = 1, NJ = 1, ncell = 1;
  bool rowmajor = true;  // cell keys are row-major (not Morton)
  const uint32_t* keys = nullptr;
  const uint32_t* cell_off = nullptr;
  int* row = nullptr;
  int* ccnt = nullptr;
  uint64_t* runc = nullptr;
  uint32_t* idx = nullptr;
  float *hi = nullptr, *lo = nullptr;
  float* dgi = nullptr;  // mixed: inner diagonal
  double* c64 = nullptr;
  double *diag = nullptr, *dinv = nullptr, *rhs = nullptr, *x64 = nullptr, *r64 = nullptr;
  double* partials = nullptr;
  void* state = nullptr;       // device solver state
  void* host_state = nullptr;  // pinned mirror
  unsigned* flags = nullptr;   // assembly overflow flags
  unsigned* host_flags = nullptr;
  void* vec[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int vec_bytes = 0;
  struct CoarseWork* coarse = nullptr;  // two-level preconditioner, created on first use
  unsigned mv_seq = 0;                  // inner matvec launches seen by the timers
  double bn = -1;                       // ||rhs|| over the rows of this assembly (< 0: not yet taken)
  uint64_t gen = 0;                     // assemblies so far: tags data derived from one operator
};

constexpr unsigned kMvSample = 4;  // timers time one inner matvec launch in this many

int ppe_workspace_alloc(PpeWork& w, int n, int cap, bool mixed);
void ppe_workspace_free(PpeWork& w);
PpeOp ppe_op(const PpeWork& w);

// dense first-particle-of-cell table from sorted keys (num_cells + 1 entries)
int cell_offsets(hipStream_t s, int n, const uint32_t* keys, int num_cells, uint32_t* off);

template <typename R>
int ppe_assemble(hipStream_t s, const DevParams& P, const SphView<R>& v, PpeWork& w, const uint32_t* keys,
                 const uint32_t* cell_off);
template <typename R>
int ppe_jacobi(hipStream_t s, const SphView<R>& v, PpeWork& w);
// y = A x with the full-precision operator (FP64, or hi + lo); residual form
// y = rhs - A x when residual; ||y||^2 over rows into *norm2 (synchronises)
int ppe_apply64(hipStream_t s, PpeWork& w, const double* x, double* y, bool residual, double* norm2);
// timing probe of the inner FP32 matvec: variant 0 the stored operator
// (the solver's kernel), 1 matrix-free from the neighbour list; reps timed
// launches after one untimed; y = A x for the device vector x; pk is n
// float2 of scratch for variant 1
template <typename R>
int ppe_matvec_probe(hipStream_t s, const DevParams& P, const SphView<R>& v, PpeWork& w, int variant, int reps,
                     const double* x64, double* y64, float2* pk, float* ms);
// known_rr >= 0: w.r64 already holds b - A guess over the rows, with this
// squared norm (the caller has just evaluated the guess)
int ppe_solve(hipStream_t s, PpeWork& w, double* pres, const sophia_ppe_opts& o, sophia_ppe_stats& st,
              const double* guess_or_null, Timers* tm, double known_rr = -1.0);
// ||b - A x|| / ||b|| into *rel; leaves b - A x in w.r64 and its squared norm in *rr_out
int ppe_residual_of(hipStream_t s, PpeWork& w, const double* x, double* rel, double* rr_out = nullptr);
// out = 3 p - 3 h1 + h2 (depth 2), 2 p - h1 (depth 1) or p on rows; p elsewhere
int ppe_extrapolate(hipStream_t s, const PpeWork& w, const double* p, const double* h1, const double* h2, int depth,
                    double* out);
int ppe_rhs_norm(hipStream_t s, PpeWork& w, double* bn);
// v -= mean of v over the rows (zero-mean PPE of a periodic box)
int ppe_project_mean(hipStream_t s, PpeWork& w, double* v);

}  // namespace sophia
