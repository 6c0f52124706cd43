// Learned PPE warm start: one message-passing layer over the SPH neighbour
// list, BF16 weights on MFMA with FP32 accumulation (see learned.hip).
#pragma once

#include "common.hpp"

namespace sophia {

// parameter layout of the FP32 master vector (row-major W[out][in])
constexpr int kH = 64;
constexpr int kEin = 6;
constexpr int OFF_W1 = 0;                    // 64 x 6
constexpr int OFF_B1 = OFF_W1 + kH * kEin;   // 64
constexpr int OFF_W2 = OFF_B1 + kH;          // 64 x 64
constexpr int OFF_B2 = OFF_W2 + kH * kH;     // 64
constexpr int OFF_W3 = OFF_B2 + kH;          // 64 x 64
constexpr int OFF_B3 = OFF_W3 + kH * kH;     // 64
constexpr int OFF_W4 = OFF_B3 + kH;          // 64
constexpr int OFF_B4 = OFF_W4 + kH;          // 1
constexpr int kNParam = OFF_B4 + 1;          // 8833
constexpr int kMlpBwdBlocks = 1024;          // persistent blocks of the backward pass

struct Mlp {
  bool ready = false;      // parameters set
  int n = 0, cap = 0;
  float* master = nullptr;   // FP32 master parameters
  float* grad = nullptr;     // FP32 gradient
  float* adam_m = nullptr;
  float* adam_v = nullptr;
  __bf16* wbf = nullptr;     // BF16 copy read by the kernels
  int adam_t = 0;
  float2* nf = nullptr;      // per particle: rho_j / rho_fluid, r_j / rms(r)
  double* r0 = nullptr;      // residual b - A p_prev of the PPE rows
  double* guess = nullptr;   // predicted pressure (FP64, sorted order)
  double* target = nullptr;  // training target (FP64, sorted order)
  double* pprev = nullptr;   // pressure before the solve (FP64, sorted order)
  float* partials = nullptr; // per-block gradient partials (kMlpBwdBlocks x (kNParam + 1))
  double* scal = nullptr;    // device scalars: sum r^2, sum diag, rows, scale, loss
  double* fpart = nullptr;   // per-block partials of the feature sums
  double* host_scal = nullptr;
};

int mlp_param_count();
int mlp_alloc(Mlp& m, int n, int cap);
void mlp_free(Mlp& m);
int mlp_set_params(hipStream_t s, Mlp& m, const float* host, int n);
int mlp_get_params(hipStream_t s, const Mlp& m, float* host, int n);
int mlp_reset_adam(hipStream_t s, Mlp& m);
// node features from the residual r0 (rows) and the operator diagonal; the
// output scale is returned in *scale (synchronises)
template <typename R>
int mlp_features(hipStream_t s, const DevParams& P, const SphView<R>& v, Mlp& m, const int* row, const double* diag,
                 double* scale, double* rr_base = nullptr);
// guess_i = pprev_i + scale * y_i on rows, pprev_i elsewhere
template <typename R>
int mlp_forward(hipStream_t s, const DevParams& P, const SphView<R>& v, Mlp& m, const int* row, const double* pprev,
                double scale, double* guess);
// loss of y against (target - pprev)/scale on rows, gradient into m.grad;
// with step, one Adam update (lr) and a refresh of the BF16 copy
template <typename R>
int mlp_backward(hipStream_t s, const DevParams& P, const SphView<R>& v, Mlp& m, const int* row, const double* pprev,
                 const double* target, double scale, int n_rows, bool step, double lr, double* loss);
int mlp_gather_previous(hipStream_t s, int n, const int* row, const double* pres, double* x);

}  // namespace sophia
