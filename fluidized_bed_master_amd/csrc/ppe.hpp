// PPE workspace and entry points (see ppe.hip).
#pragma once

#include "common.hpp"
#include "timers.hpp"

namespace sophia {

struct PpeView {
  int n, cap;
  const uint32_t* nbr;
  const int* cnt;
  int* row;
  double *diag, *dinv, *rhs;
  double* coef64;
  float *coef_hi, *coef_lo;
};

struct PpeWork {
  int n = 0, nb = 0, cap = 0;
  bool mixed = false;
  int* row = nullptr;
  double *diag = nullptr, *dinv = nullptr, *rhs = nullptr, *x64 = nullptr, *r64 = nullptr;
  double *partials = nullptr, *sc = nullptr, *host_sc = nullptr;
  void* vec[8] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  void* d32 = nullptr;
  double* coef64 = nullptr;
  float *coef_hi = nullptr, *coef_lo = nullptr;
};

int ppe_workspace_alloc(PpeWork& w, int n, int cap, bool mixed);
void ppe_workspace_free(PpeWork& w);
PpeView ppe_view(const PpeWork& w, const uint32_t* nbr, const int* cnt, int cap);
template <typename R>
int ppe_assemble(hipStream_t s, const DevParams& P, const SphView<R>& v, PpeWork& w);
template <typename R>
int ppe_jacobi(hipStream_t s, const SphView<R>& v, PpeWork& w);
int ppe_apply64(hipStream_t s, PpeWork& w, const uint32_t* nbr, const int* cnt, int cap, const double* x, double* y,
                bool residual, double* norm2);
int ppe_solve(hipStream_t s, PpeWork& w, const uint32_t* nbr, const int* cnt, int cap, double* pres,
              const sophia_ppe_opts& o, sophia_ppe_stats& st, const double* guess_or_null, Timers* tm);
int ppe_residual_of(hipStream_t s, PpeWork& w, const uint32_t* nbr, const int* cnt, int cap, const double* x,
                    double* rel);
int ppe_rhs_norm(hipStream_t s, PpeWork& w, double* bn);

}  // namespace sophia
