// Learned PPE warm start (placeholder until the MFMA kernels land).
#include "learned.hpp"

namespace sophia {

namespace {
__global__ void k_prev(int n, const int* row, const double* pres, double* x) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = row[i] ? pres[i] : 0.0;
}
}  // namespace

int mlp_param_count() { return 0; }
int mlp_alloc(Mlp& m, int n, int cap) {
  m.n = n;
  m.cap = cap;
  if (!m.guess) HIP_TRY(hipMalloc(&m.guess, (size_t)(n > 0 ? n : 1) * 8));
  return 0;
}
void mlp_free(Mlp& m) {
  if (m.guess) (void)hipFree(m.guess);
  m = Mlp{};
}
int mlp_set_params(hipStream_t, Mlp&, const float*, int) {
  set_error("learned warm start not built");
  return SOPHIA_E_STATE;
}
int mlp_get_params(hipStream_t, const Mlp&, float*, int) {
  set_error("learned warm start not built");
  return SOPHIA_E_STATE;
}
int mlp_reset_adam(hipStream_t, Mlp&) { return 0; }
template <typename R>
int mlp_predict_rows(hipStream_t, const DevParams&, const SphView<R>&, Mlp&, const int*, double*) {
  set_error("learned warm start not built");
  return SOPHIA_E_STATE;
}
template <typename R>
int mlp_train_rows(hipStream_t, const DevParams&, const SphView<R>&, Mlp&, const int*, const double*, double, double*,
                   float*, int, bool) {
  set_error("learned warm start not built");
  return SOPHIA_E_STATE;
}
int mlp_gather_previous(hipStream_t s, int n, const int* row, const double* pres, double* x) {
  if (n <= 0) return 0;
  k_prev<<<grid_for(n), kBlock, 0, s>>>(n, row, pres, x);
  HIP_TRY(hipGetLastError());
  return 0;
}
template int mlp_predict_rows<float>(hipStream_t, const DevParams&, const SphView<float>&, Mlp&, const int*, double*);
template int mlp_predict_rows<double>(hipStream_t, const DevParams&, const SphView<double>&, Mlp&, const int*, double*);
template int mlp_train_rows<float>(hipStream_t, const DevParams&, const SphView<float>&, Mlp&, const int*,
                                   const double*, double, double*, float*, int, bool);
template int mlp_train_rows<double>(hipStream_t, const DevParams&, const SphView<double>&, Mlp&, const int*,
                                    const double*, double, double*, float*, int, bool);

}  // namespace sophia
