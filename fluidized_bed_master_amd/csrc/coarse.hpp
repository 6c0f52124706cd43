// Two-level additive preconditioner of the PPE Krylov solve (coarse.hip).
//
//   M^-1 = D^-1 + Z E^-1 Z^T,   E = Z^T A Z,
//
// Z the indicator matrix of aggregates of kAgg^3 cells (a particle belongs to
// the aggregate of its cell).  D^-1 is the Jacobi part the solver already
// applies; the coarse part restricts a vector to aggregate sums, multiplies by
// a dense approximate inverse X ~ E^-1 and adds the result back to every row of
// the aggregate.  X is kept from step to step and refreshed by Newton-Schulz
// steps X <- X (2 I - E X) (two GEMMs) whenever ||I - E X||_F / sqrt(na) of the
// new E exceeds kCoarseTol; a cold start begins from X = diag(E)^-1.
#pragma once

#include "ppe.hpp"

namespace sophia {

constexpr int kAgg = 3;              // cells per aggregate edge
constexpr double kCoarseTol = 0.02;  // accepted ||I - E X||_F / sqrt(na)
constexpr int kCoarseMaxNa = 16384;  // dense E and X (4 x na^2 x 4 B) above this are refused

struct CoarseWork {
  int NI = 0, NJ = 0, NK = 0;
  int AX = 0, AY = 0, AZ = 0, na = 0;  // aggregates per axis, total
  int ld = 0;            // leading dimension of the dense matrices (na rounded up to 64)
  float* E = nullptr;    // Z^T A Z, row-major na x na
  float* X = nullptr;    // approximate inverse, row-major
  float* X2 = nullptr;   // Newton-Schulz product
  float* U = nullptr;    // 2 I - E X
  double* c = nullptr;   // restricted vector (na)
  double* y = nullptr;   // coarse correction (na)
  double* partials = nullptr;
  double* slots = nullptr; // 8 aggregate-slot sums per particle (coarse build)
  int slots_n = 0;
  double* host = nullptr;  // pinned copy of the partials
  void* blas = nullptr;    // rocblas_handle bound to the solver stream
  bool valid = false;      // X is an accepted inverse of some earlier E
  double res = 0;          // ||I - E X||_F / sqrt(na) of X against the current E
  int ns_steps = 0;        // Newton-Schulz steps of the last setup
  // E built ahead on a second stream (default priority) while the solver
  // stream computes the initial guess and its residual
  hipStream_t side = nullptr;
  hipEvent_t ev_in = nullptr, ev_rows = nullptr;
  bool rows_pending = false;
  uint64_t rows_gen = 0;  // operator generation (PpeWork::gen) of the pending side-stream build
  uint64_t e_gen = 0;     // operator generation E was built from
  uint64_t x_gen = 0;     // generation of the E that X was last refreshed against
};

// Start building E from the operator just assembled on stream s, on the side
// stream; the next coarse_setup on s for the same operator generation waits
// for it instead of building E.
int coarse_rows_early(hipStream_t s, CoarseWork& c, const PpeWork& w, Timers* tm);
// Make stream s wait for a pending side-stream build and clear it.  Called at
// the end of every PPE stage, so that no later kernel on s (NNPS, the next
// assembly) rewrites what the side stream still reads, whether or not the
// solve reached coarse_setup.
int coarse_settle(hipStream_t s, CoarseWork& c);

// Build E from the assembled operator (row-major cell keys) and bring X to
// ||I - E X|| <= kCoarseTol.  Returns 0 with *usable = false when X cannot be
// brought there (the solve then runs with Jacobi alone).
int coarse_setup(hipStream_t s, CoarseWork& c, const PpeWork& w, bool* usable, Timers* tm);
// The Krylov vector update that precedes a coarse application (kind 1: the
// BiCGStab direction p = r + beta (p - omega v), or p = r when *first; kind 2:
// s = r - alpha v, in place in r), with out = D^-1 of the updated vector; the
// scalars are read on the device.
template <typename V>
struct CoarseUpd {
  const double* dinv;
  V* r;
  V* p;
  const V* v;
  V* out;
  const double *alpha, *beta, *omega;
  const int* first;
};
// The update of kind `kind`, then out += Z X Z^T of the updated vector over
// rows; the update and the restriction Z^T are one pass (skipped when *done or
// *brk).
template <typename V>
int coarse_apply_update(hipStream_t s, const CoarseWork& c, const PpeWork& w, const CoarseUpd<V>& u, int kind,
                        const int* done, const int* brk);
void coarse_free(CoarseWork& c);
// copy E and X (row-major na x na, FP32) to the host where non-null (synchronises)
int coarse_export(hipStream_t s, const CoarseWork& c, float* E, float* X);

}  // namespace sophia
