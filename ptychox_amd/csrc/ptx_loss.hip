// Alternative losses and regularisers with their gradients, on stacks of
// frames [batch][h][w]:
//   * residual losses of predicted vs measured amplitudes (real): the finite-
//     difference losses of python/train_coded.py:101-107 (fd_loss, isotropic
//     and anisotropic) and of matlab/Examples/Ptychography/+loss_fun/fd_loss.m,
//     the squared error of +loss_fun/l2_loss.m;
//   * regularisers of complex fields: +loss_fun/tv_loss.m and sparse_loss.m.
// With d = pred - target (or the field itself), dx(i,j) = d(i,j+1) - d(i,j),
// dy(i,j) = d(i+1,j) - d(i,j) with wrap-around (torch.roll(.., -1) /
// MATLAB [2:end,1]), the per-site penalty phi(dx, dy) and
//   L_b = scale * sum_{i,j} phi(i,j),
//   grad(i,j) = scale * (a_x(i,j-1) + a_y(i-1,j) - a_x(i,j) - a_y(i,j)),
// where (a_x, a_y) = (dphi/d dx, dphi/d dy) in PyTorch's convention
// (dL/dRe + i dL/dIm for complex values).  One thread per site reads the
// 7-point neighbourhood it needs (L1/L2 serve the overlap); per-frame sums go
// through fixed-order work-group and frame reductions (no atomics).
#include "common_kernels.hpp"
#include "ptx_internal.hpp"

namespace ptx {
namespace {

constexpr int kLossBlock = 256;

// real and complex values through one interface
__device__ __forceinline__ float v_sub(float a, float b) { return a - b; }
__device__ __forceinline__ float2 v_sub(float2 a, float2 b) { return c_sub(a, b); }
__device__ __forceinline__ float v_abs2(float a) { return a * a; }
__device__ __forceinline__ float v_abs2(float2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ float v_scale(float a, float s) { return a * s; }
__device__ __forceinline__ float2 v_scale(float2 a, float s) { return c_scale(a, s); }
__device__ __forceinline__ float v_add(float a, float b) { return a + b; }
__device__ __forceinline__ float2 v_add(float2 a, float2 b) { return c_add(a, b); }
template <class T>
__device__ __forceinline__ T v_zero();
template <>
__device__ __forceinline__ float v_zero<float>() { return 0.0f; }
template <>
__device__ __forceinline__ float2 v_zero<float2>() { return c_zero(); }
// sgn(z) = z/|z|, 0 at 0 (torch.abs backward)
__device__ __forceinline__ float v_sgn(float a) { return (float)((a > 0.0f) - (a < 0.0f)); }
__device__ __forceinline__ float2 v_sgn(float2 a) {
  const float m = sqrtf(v_abs2(a));
  return m > 0.0f ? c_scale(a, 1.0f / m) : c_zero();
}

// phi and its partials at one site from (dx, dy)
template <int KIND, class T>
__device__ __forceinline__ float site(T dx, T dy, float eps, T& ax, T& ay) {
  if constexpr (KIND == PTX_LOSS_FD_ISO || KIND == PTX_LOSS_SPARSE_ISO) {
    const float s = sqrtf(v_abs2(dx) + v_abs2(dy) + eps);
    float k = s > 0.0f ? 1.0f / s : 0.0f;  // eps = 0 and a flat site: the zero subgradient
    float phi = s;
    if constexpr (KIND == PTX_LOSS_SPARSE_ISO) {
      const float e = expf(-8.0f * s);
      phi = 1.0f - e;
      k *= 8.0f * e;
    }
    ax = v_scale(dx, k);
    ay = v_scale(dy, k);
    return phi;
  } else if constexpr (KIND == PTX_LOSS_FD_ANISO || KIND == PTX_LOSS_SPARSE_ANISO) {
    const float m = sqrtf(v_abs2(dx)) + sqrtf(v_abs2(dy));
    ax = v_sgn(dx);
    ay = v_sgn(dy);
    if constexpr (KIND == PTX_LOSS_SPARSE_ANISO) {
      const float e = expf(-8.0f * m);
      ax = v_scale(ax, 8.0f * e);
      ay = v_scale(ay, 8.0f * e);
      return 1.0f - e;
    }
    return m;
  } else {  // PTX_LOSS_GRAD_L2
    ax = v_scale(dx, 2.0f);
    ay = v_scale(dy, 2.0f);
    return v_abs2(dx) + v_abs2(dy);
  }
}

template <int KIND, class T>
__global__ void __launch_bounds__(kLossBlock)
    k_stencil_loss(const T* __restrict__ pred, const T* __restrict__ target, int64_t h, int64_t w,
                   int64_t blocks_per_frame, float scale, float eps, float* __restrict__ part, T* __restrict__ grad) {
  __shared__ float red[16];
  const int64_t frame = blockIdx.x / blocks_per_frame;
  const int64_t idx = (blockIdx.x % blocks_per_frame) * kLossBlock + threadIdx.x;
  const bool valid = idx < h * w;
  float phi = 0.0f;
  if (valid) {
    const int64_t i = idx / w, j = idx % w;
    const int64_t ip = i + 1 == h ? 0 : i + 1, im = i == 0 ? h - 1 : i - 1;
    const int64_t jp = j + 1 == w ? 0 : j + 1, jm = j == 0 ? w - 1 : j - 1;
    const int64_t base = frame * h * w;
    auto D = [&](int64_t r, int64_t c) -> T {
      const int64_t o = base + r * w + c;
      return target ? v_sub(pred[o], target[o]) : pred[o];
    };
    if constexpr (KIND == PTX_LOSS_L2) {
      const T d = D(i, j);
      phi = v_abs2(d);
      grad[base + idx] = v_scale(d, 2.0f * scale);
    } else {
      const T d = D(i, j), dr = D(i, jp), dl = D(i, jm), dd = D(ip, j), du = D(im, j);
      const T ddl = D(ip, jm), dur = D(im, jp);
      T ax, ay, axl, ayl, axu, ayu;
      phi = site<KIND>(v_sub(dr, d), v_sub(dd, d), eps, ax, ay);   // site (i, j)
      site<KIND>(v_sub(d, dl), v_sub(ddl, dl), eps, axl, ayl);      // site (i, j-1)
      site<KIND>(v_sub(dur, du), v_sub(d, du), eps, axu, ayu);      // site (i-1, j)
      const T g = v_sub(v_add(axl, ayu), v_add(ax, ay));
      grad[base + idx] = v_scale(g, scale);
    }
  }
  const float tot = block_sum(phi, red);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

template <class T>
ptx_status stencil_loss(const char* name, int kind, const T* pred, const T* target, int64_t B, int64_t h, int64_t w,
                        float scale, float eps, float* loss, T* grad, void* workspace, size_t ws_bytes,
                        hipStream_t s) {
  if (B < 0 || h < 1 || w < 1 || (B > 0 && (!pred || !loss || !grad))) {
    set_error("%s: invalid arguments", name);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (kind < PTX_LOSS_L2 || kind > PTX_LOSS_SPARSE_ANISO) {
    set_error("%s: unknown loss kind %d", name, kind);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (!(eps >= 0.0f)) {
    set_error("%s: eps must be >= 0", name);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (B == 0) return PTX_OK;
  const int64_t per = (h * w + kLossBlock - 1) / kLossBlock;
  if ((uint64_t)(B * per) >= (1ull << 31)) {
    set_error("%s: %lld frames of %lld x %lld exceed the grid", name, (long long)B, (long long)h, (long long)w);
    return PTX_ERR_UNSUPPORTED;
  }
  const size_t need = (size_t)(B * per) * sizeof(float);
  if (!workspace || ws_bytes < need) {
    set_error("%s: workspace of %zu bytes needed, %zu given", name, need, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  timing_begin(s);
  float* part = (float*)workspace;
  const dim3 grid((unsigned)(B * per)), block(kLossBlock);
  switch (kind) {
    case PTX_LOSS_L2:
      k_stencil_loss<PTX_LOSS_L2, T><<<grid, block, 0, s>>>(pred, target, h, w, per, scale, eps, part, grad);
      break;
    case PTX_LOSS_FD_ISO:
      k_stencil_loss<PTX_LOSS_FD_ISO, T><<<grid, block, 0, s>>>(pred, target, h, w, per, scale, eps, part, grad);
      break;
    case PTX_LOSS_FD_ANISO:
      k_stencil_loss<PTX_LOSS_FD_ANISO, T><<<grid, block, 0, s>>>(pred, target, h, w, per, scale, eps, part, grad);
      break;
    case PTX_LOSS_GRAD_L2:
      k_stencil_loss<PTX_LOSS_GRAD_L2, T><<<grid, block, 0, s>>>(pred, target, h, w, per, scale, eps, part, grad);
      break;
    case PTX_LOSS_SPARSE_ISO:
      k_stencil_loss<PTX_LOSS_SPARSE_ISO, T><<<grid, block, 0, s>>>(pred, target, h, w, per, scale, eps, part, grad);
      break;
    default:
      k_stencil_loss<PTX_LOSS_SPARSE_ANISO, T><<<grid, block, 0, s>>>(pred, target, h, w, per, scale, eps, part,
                                                                       grad);
      break;
  }
  ptx_status st = check_launch("stencil loss", s);
  if (st != PTX_OK) return st;
  loss_rows(part, per, B, scale, loss, s);
  return check_launch("loss frame sums", s);
}

}  // namespace
}  // namespace ptx

using namespace ptx;

extern "C" {

size_t ptx_loss_workspace(int64_t batch, int64_t h, int64_t w) {
  TuneScope tune_scope;
  if (batch < 0 || h < 1 || w < 1) return 0;
  return (size_t)(batch * ((h * w + kLossBlock - 1) / kLossBlock)) * sizeof(float);
}

ptx_status ptx_amplitude_loss(int kind, const float* pred, const float* target, int64_t batch, int64_t h, int64_t w,
                              float scale, float eps, float* loss_per_frame, float* grad, void* workspace,
                              size_t workspace_bytes, void* stream) {
  TuneScope tune_scope;
  return stencil_loss<float>("ptx_amplitude_loss", kind, pred, target, batch, h, w, scale, eps, loss_per_frame, grad,
                             workspace, workspace_bytes, (hipStream_t)stream);
}

ptx_status ptx_field_regularizer(int kind, const void* field, int64_t batch, int64_t h, int64_t w, float scale,
                                 float eps, float* loss_per_frame, void* grad, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  TuneScope tune_scope;
  return stencil_loss<float2>("ptx_field_regularizer", kind, (const float2*)field, nullptr, batch, h, w, scale, eps,
                              loss_per_frame, (float2*)grad, workspace, workspace_bytes, (hipStream_t)stream);
}

}  // extern "C"
