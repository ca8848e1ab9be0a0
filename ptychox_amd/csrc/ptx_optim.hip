// Optimiser and batching helpers for device-resident reconstruction loops:
// the Adam update of torch.optim.Adam as the reference drivers configure it
// (python/train_ptychography.py:500-507: betas (0.9, 0.999), eps 1e-16, complex
// parameters updated per real component), and the gather of a permuted
// mini-batch's scan positions.
//
// Step-dependent scalars (step size lr/(1-beta1^t), sqrt(1-beta2^t), eps) are
// read from device memory, so a captured hipGraph of a whole epoch can be
// replayed with the next epoch's values written into that buffer.
#include "common_kernels.hpp"
#include "ptx_internal.hpp"

namespace ptx {
namespace {

// torch's foreach Adam (torch/optim/adam.py, _multi_tensor_adam) in one pass:
//   m = lerp(m, g, 1 - b1);  v = b2 v + (1 - b2) g^2
//   denom = sqrt(v) / sqrt(1 - b2^t) + eps;  p += -(lr / (1 - b1^t)) * (m / denom)
// PROJECT: after the update, interleaved complex pairs become (max(re, 0), 0)
// with clamp_min's NaN-propagating comparison (x < 0 ? 0 : x).
template <bool PROJECT>
static __global__ void __launch_bounds__(256)
    k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, int64_t n,
           float beta1, float beta2, const float* __restrict__ hyper) {
  const float step_size = hyper[0];
  const float bc2_sqrt = hyper[1];
  const float eps = hyper[2];
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    float mi = m[i], vi = v[i];
    float pi = adam_component(p[i], g[i], mi, vi, beta1, beta2, step_size, bc2_sqrt, eps);
    m[i] = mi;
    v[i] = vi;
    if (PROJECT) pi = (i & 1) ? 0.0f : (pi < 0.0f ? 0.0f : pi);
    p[i] = pi;
  }
}

static __global__ void __launch_bounds__(256)
    k_gather_positions(const int64_t* __restrict__ order, const int64_t* __restrict__ pos_all, int64_t count,
                       int64_t B, int64_t* __restrict__ out, int* __restrict__ bad) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b >= B) return;
  int64_t k = order[b];
  if (k < 0 || k >= count) {  // never read outside the position table
    if (bad) atomicOr(bad, 1);
    k = 0;
  }
  out[2 * b] = pos_all[2 * k];
  out[2 * b + 1] = pos_all[2 * k + 1];
}

}  // namespace

ptx_status launch_adam(const AdamArgs& adam, const float2* grad, int64_t n, hipStream_t s) {
  if (n <= 0) return PTX_OK;
  int64_t blocks = (2 * n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  k_adam<false><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(reinterpret_cast<float*>(adam.param),
                                                             reinterpret_cast<const float*>(grad), adam.m, adam.v,
                                                             2 * n, adam.beta1, adam.beta2, adam.hyper);
  return check_launch("adam", s);
}

ptx_status launch_gather_positions(const int64_t* order, const int64_t* positions_all, int64_t count, int64_t batch,
                                   int64_t* positions_out, int* bad, hipStream_t s) {
  if (batch <= 0) return PTX_OK;
  k_gather_positions<<<dim3((unsigned)((batch + 255) / 256)), dim3(256), 0, s>>>(order, positions_all, count, batch,
                                                                                 positions_out, bad);
  return check_launch("gather positions", s);
}

}  // namespace ptx

using namespace ptx;

extern "C" {

ptx_status ptx_adam_step(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n, float beta1,
                         float beta2, const float* hyper, unsigned flags, void* stream) {
  if (n < 0 || (n > 0 && (!param || !grad || !exp_avg || !exp_avg_sq || !hyper))) {
    set_error("ptx_adam_step: invalid arguments");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (!(beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f)) {
    set_error("ptx_adam_step: betas must lie in [0, 1)");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (flags & ~PTX_ADAM_PROJECT_REAL_NONNEG) {
    set_error("ptx_adam_step: unknown flags");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  const bool project = (flags & PTX_ADAM_PROJECT_REAL_NONNEG) != 0;
  if (project && (n & 1)) {
    set_error("ptx_adam_step: the real non-negative projection needs complex pairs (n even)");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (n == 0) return PTX_OK;
  hipStream_t s = (hipStream_t)stream;
  timing_begin(s);
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (project)
    k_adam<true><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(param, grad, exp_avg, exp_avg_sq, n, beta1, beta2, hyper);
  else
    k_adam<false><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(param, grad, exp_avg, exp_avg_sq, n, beta1, beta2,
                                                               hyper);
  return check_launch("adam", s);
}

ptx_status ptx_gather_positions(const int64_t* order, const int64_t* positions_all, int64_t count, int64_t batch,
                                int64_t* positions_out, int* bad_index_flag, void* stream) {
  if (batch < 0 || count < 0 || (batch > 0 && (!order || !positions_all || !positions_out || !bad_index_flag))) {
    set_error("ptx_gather_positions: invalid arguments");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (batch > 0 && count == 0) {  // no entry 0 to fall back on
    set_error("ptx_gather_positions: empty position table for a batch of %lld", (long long)batch);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  return launch_gather_positions(order, positions_all, count, batch, positions_out, bad_index_flag,
                                 (hipStream_t)stream);
}

}  // extern "C"
