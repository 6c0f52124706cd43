// Learned PPE warm start: one message-passing layer over the SPH neighbour
// list, BF16 weights on MFMA with FP32 accumulation (see learned.hip).
#pragma once

#include "common.hpp"

namespace sophia {

struct Mlp {
  bool ready = false;      // parameters set
  int n = 0, cap = 0;
  float* master = nullptr;   // FP32 master parameters
  float* grad = nullptr;     // FP32 gradient
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  uint16_t* wbf = nullptr;   // BF16 copy used by the MFMA kernels
  int adam_t = 0;
  // per-row scratch
  double* guess = nullptr;   // predicted pressure (FP64, sorted order)
  float* node_feat = nullptr;  // per particle: rho_j/rho0, normalised source
  float* agg = nullptr;        // 64 per particle: aggregated edge messages
  float* hid = nullptr;        // 64 per particle: node hidden layer
  float* out = nullptr;        // per particle network output
  float* dout = nullptr;       // dLoss/dout
  float* dagg = nullptr;       // 64 per particle
  double* r0 = nullptr;        // residual of the previous pressure
  double* scal = nullptr;      // device scalars (rms of source, mean diag, loss)
  double* host_scal = nullptr;
};

int mlp_param_count();
int mlp_alloc(Mlp& m, int n, int cap);
void mlp_free(Mlp& m);
int mlp_set_params(hipStream_t s, Mlp& m, const float* host, int n);
int mlp_get_params(hipStream_t s, const Mlp& m, float* host, int n);
int mlp_reset_adam(hipStream_t s, Mlp& m);
template <typename R>
int mlp_predict_rows(hipStream_t s, const DevParams& P, const SphView<R>& v, Mlp& m, const int* row, double* guess);
template <typename R>
int mlp_train_rows(hipStream_t s, const DevParams& P, const SphView<R>& v, Mlp& m, const int* row,
                   const double* target_host, double lr, double* loss, float* grad_host, int n_params, bool step);
int mlp_gather_previous(hipStream_t s, int n, const int* row, const double* pres, double* x);

}  // namespace sophia
