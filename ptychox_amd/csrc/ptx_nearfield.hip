// Near-field ptychography (ptychoX_nearfield): forward, adjoint and the fused
// amplitude-loss training step.
//
// The reference implements near-field only in MATLAB, as a full-frame
// propagation (matlab/PtychoX_updated/scripts/forward/fwd_adPtychoX_nearf.cu:58-114:
// FFT, pixel-wise transfer multiply, inverse FFT, detector magnitude).  This
// operator applies that angular-spectrum propagation to each probe-illuminated
// scan window, as the conventional operator does with its far-field FFT:
//   Psi_b = ifft2(T fft2(P_b p)),   A_b = |Psi_b|          (T has DC at [0, 0])
//   chi_b = ifft2(conj(T) fft2(G Psi/max(|Psi|, eps)))     (adjoint of F^-1 D_T F)
//   grad_sample[win_b] += chi_b conj(p),  grad_probe += sum_b chi_b conj(P_b)
//
// Passes: R1 gather * probe + row FFT | C_T column FFT, * T, column iFFT |
// R3 row iFFT / N -> field, amplitude.  Backward: gamma + row FFT | C_T with
// conj(T) | row iFFT / N + scatter.  The step joins R3 and the gamma row pass
// through the loss (five passes in all).
#include "scan_common.hpp"

namespace ptx {
namespace {

// row iFFT / N -> field and amplitude
struct NfFieldRows {
  const float2* in;
  float2* field;
  float* amp;
  float inv_n;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const float2* p = in + row * N;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = valid ? p[t + q * T] : c_zero();
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int64_t k = row * N + t + q * T;
      const float2 e = c_scale(v[q], inv_n);
      field[k] = e;
      amp[k] = c_abs(e);
    }
  }
};

// Gamma = G Psi / max(|Psi|, eps), row FFT
struct NfGammaRows {
  const float2* field;
  const float* g;
  float2* out;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int64_t k = row * N + t + q * T;
      if (valid) {
        const float2 e = field[k];
        v[q] = c_scale(e, g[k] * fast_rcp(fmaxf(c_abs(e), kFltEps)));
      } else {
        v[q] = c_zero();
      }
    }
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    if (!valid) return;
    float2* p = out + row * N;
#pragma unroll
    for (int q = 0; q < E; ++q) p[t + q * T] = v[q];
  }
};

// step: row iFFT / N -> A, loss partial per row, Gamma, row FFT (in place)
template <int N, int E>
__global__ void __launch_bounds__(256)
    k_nf_step_rows(float2* __restrict__ tb, const float* __restrict__ measured, float* __restrict__ amp_out,
                   float* __restrict__ loss_part, int64_t nrows, float gscale, float inv_n) {
  using Cf = RowCfg<N, E>;
  constexpr int T = Cf::T, S = Cf::S;
  __shared__ float2 lds[S * lds_len<N>()];
  __shared__ float red[4];  // per-wave loss sums
  const int seq = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int64_t r0 = (int64_t)blockIdx.x * S;
  const int64_t row = r0 + seq;
  const bool valid = row < nrows;
  // the work-group's rows as one buffer of each array: rows past nrows lie
  // beyond it, so their loads return zero and their stores are dropped -- no
  // branches around the loads, and the measured amplitudes are requested with
  // the frame row, their latency hidden behind the forward transform
  const uint32_t live = (uint32_t)((nrows - r0 < S ? nrows - r0 : S) * N);
  const uint32_t el = (uint32_t)(seq * N + t);
  const auto rt = buffer_rsrc(tb + r0 * N, live * 8u);
  float2 v[E];
  float m[E];
  {
    const auto rm = buffer_rsrc(measured + r0 * N, live * 4u);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = buf_load_f2(rt, el * 8u, (uint32_t)(q * T * 8));
#pragma unroll
    for (int q = 0; q < E; ++q) m[q] = buf_load_f1(rm, el * 4u, (uint32_t)(q * T * 4));
  }
  fft_row<N, E, +1>(v, t, &lds[seq * lds_len<N>()]);
  float lsum = 0.0f;
  // (an invalid row transforms zeros: amplitude 0 against 0, no loss, no gradient)
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const float2 e = c_scale(v[q], inv_n);
    const AbsInv ai = c_abs_inv(e, 1.0f / kFltEps);
    const float d = ai.a - m[q];
    lsum += d * d;
    m[q] = ai.a;
    v[q] = c_scale(e, gscale * d * ai.inv);
  }
  if (amp_out) {
    const auto ra = buffer_rsrc(amp_out + r0 * N, live * 4u);
#pragma unroll
    for (int q = 0; q < E; ++q) buf_store_f1(m[q], ra, el * 4u, (uint32_t)(q * T * 4));
  }
  // row sum in a fixed order: butterfly within each wave, then the row's
  // waves in order (a row is T >= 64 threads, whole waves; shorter rows share
  // a wave as T-lane groups).  The barrier also orders the LDS reuse of the
  // second transform.
#pragma unroll
  for (int off = 1; off < 64; off <<= 1)
    if (off < T) lsum += __shfl_xor(lsum, off, 64);
  if constexpr (T > 64) {
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lsum;
    __syncthreads();
    if (valid && t == 0) {
      float sum = 0.0f;
      for (int k = 0; k < T / 64; ++k) sum += red[(seq * T >> 6) + k];
      loss_part[row] = sum;
    }
  } else {
    // the row lies within one wave, whose lanes all hold its sum
    if (valid && t == 0) loss_part[row] = lsum;
    seq_sync<true>();  // the second transform rewrites the row's LDS
  }
  fft_row<N, E, -1>(v, t, &lds[seq * lds_len<N>()]);
#pragma unroll
  for (int q = 0; q < E; ++q) buf_store_f2(v[q], rt, el * 8u, (uint32_t)(q * T * 8));
}

template <int N, int E>
struct TransferColsRun {
  static ptx_status run(float2* tb, const float2* transfer, int64_t frames, int64_t width, bool conj, hipStream_t s) {
    launch_transfer_cols<N, E>(tb, transfer, frames, width, conj, s);
    return check_launch(conj ? "nf transfer* cols" : "nf transfer cols", s);
  }
};
template <int N, int E>
struct FieldRowsRun {
  static ptx_status run(const NfFieldRows& op, int64_t nrows, hipStream_t s) {
    launch_rows<N, E, +1>(op, nrows, s);
    return check_launch("nf field rows", s);
  }
};
template <int N, int E>
struct GammaRowsRun {
  static ptx_status run(const NfGammaRows& op, int64_t nrows, hipStream_t s) {
    launch_rows<N, E, -1>(op, nrows, s);
    return check_launch("nf gamma rows", s);
  }
};
template <int N, int E>
struct StepRowsRun {
  static ptx_status run(float2* tb, const float* measured, float* amp, float* loss_part, int64_t nrows, float gscale,
                        float inv_n, hipStream_t s) {
    const int64_t blocks = (nrows + RowCfg<N, E>::S - 1) / RowCfg<N, E>::S;
    k_nf_step_rows<N, E><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(tb, measured, amp, loss_part, nrows, gscale, inv_n);
    return check_launch("nf step rows", s);
  }
};

ptx_status nf_check(const char* fn, int64_t H, int64_t W, int64_t h, int64_t w, int64_t B, const void* transfer) {
  ptx_status st = check_shapes(fn, H, W, h, w, B);
  if (st != PTX_OK) return st;
  if (B > 0 && !transfer) {
    set_error("%s: null transfer function", fn);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  return PTX_OK;
}

}  // namespace
}  // namespace ptx

using namespace ptx;

extern "C" {

ptx_status ptx_nearfield_forward(const void* sample, int64_t H, int64_t W, const void* probe, int64_t h, int64_t w,
                                 const void* transfer, const int64_t* pos, int64_t B, void* field, float* amplitude,
                                 void* stream) {
  TuneScope tune_scope;
  ptx_status st = nf_check("ptx_nearfield_forward", H, W, h, w, B, transfer);
  if (st != PTX_OK || B == 0) return st;
  if (!sample || !probe || !pos || !field || !amplitude) {
    set_error("ptx_nearfield_forward: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  timing_begin(s);
  ScanGatherRows<false> r1{(const float2*)sample, H, W, (const float2*)probe, (int)h, pos, (float2*)field};
  if ((st = dispatch_ne<GatherRowsRun>(PASS_ROW, w, r1, B * h, s)) != PTX_OK) return st;
  if ((st = dispatch_ne<TransferColsRun>(PASS_COL, h, (float2*)field, (const float2*)transfer, B, w, false, s)) != PTX_OK)
    return st;
  NfFieldRows r3{(const float2*)field, (float2*)field, amplitude, 1.0f / (float)(h * w)};
  return dispatch_ne<FieldRowsRun>(PASS_ROW, w, r3, B * h, s);
}

size_t ptx_nearfield_backward_workspace(int64_t H, int64_t W, int64_t h, int64_t w, int64_t B, unsigned flags) {
  TuneScope tune_scope;
  if (check_shapes("ptx_nearfield_backward_workspace", H, W, h, w, B) != PTX_OK) return 0;
  ScatterPlan sp{};
  dispatch_ne<ScatterPlanQuery>(PASS_ROW_LOOP, w, h, B, &sp);
  return carve_scan(nullptr, H, W, h, w, B, flags, false, sp, 0).bytes;
}

ptx_status ptx_nearfield_backward(const float* grad_output, const void* field, const void* sample, int64_t H,
                                  int64_t W, const void* probe, int64_t h, int64_t w, const void* transfer,
                                  const int64_t* pos, int64_t B, void* grad_sample, void* grad_probe, void* workspace,
                                  size_t ws_bytes, unsigned flags, void* stream) {
  TuneScope tune_scope;
  ptx_status st = nf_check("ptx_nearfield_backward", H, W, h, w, B, transfer);
  if (st != PTX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  if (!grad_sample || !grad_probe) {
    set_error("ptx_nearfield_backward: null gradient output");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  timing_begin(s);
  if (B == 0) {
    if (!(flags & PTX_FLAG_ACCUMULATE)) {
      fill_zero2((float2*)grad_sample, H * W, s);
      fill_zero2((float2*)grad_probe, h * w, s);
    }
    return check_launch("zero fill", s);
  }
  if (!grad_output || !field || !sample || !probe || !pos) {
    set_error("ptx_nearfield_backward: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  ScatterPlan sp{};
  dispatch_ne<ScatterPlanQuery>(PASS_ROW_LOOP, w, h, B, &sp);
  const ScanWorkspace ws = carve_scan(workspace, H, W, h, w, B, flags, false, sp, 0);
  if (!workspace || ws_bytes < ws.bytes) {
    set_error("ptx_nearfield_backward: workspace of %zu bytes needed, %zu given", ws.bytes, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  NfGammaRows r1{(const float2*)field, grad_output, ws.tb};
  if ((st = dispatch_ne<GammaRowsRun>(PASS_ROW, w, r1, B * h, s)) != PTX_OK) return st;
  if ((st = dispatch_ne<TransferColsRun>(PASS_COL, h, ws.tb, (const float2*)transfer, B, w, true, s)) != PTX_OK) return st;
  return scan_scatter<false>(ws, sp, (const float2*)sample, H, W, (const float2*)probe, h, w, pos, B,
                             (float2*)grad_sample, (float2*)grad_probe, flags, 1.0f / (float)(h * w), s);
}

size_t ptx_nearfield_step_workspace(int64_t H, int64_t W, int64_t h, int64_t w, int64_t B, unsigned flags) {
  TuneScope tune_scope;
  if (check_shapes("ptx_nearfield_step_workspace", H, W, h, w, B) != PTX_OK) return 0;
  ScatterPlan sp{};
  dispatch_ne<ScatterPlanQuery>(PASS_ROW_LOOP, w, h, B, &sp);
  return carve_scan(nullptr, H, W, h, w, B, flags, true, sp, h).bytes;
}

ptx_status ptx_nearfield_step(const void* sample, int64_t H, int64_t W, const void* probe, int64_t h, int64_t w,
                              const void* transfer, const int64_t* pos, int64_t B, const float* measured,
                              float loss_scale, float* loss_per_position, float* amplitude, void* grad_sample,
                              void* grad_probe, void* workspace, size_t ws_bytes, unsigned flags, void* stream) {
  TuneScope tune_scope;
  ptx_status st = nf_check("ptx_nearfield_step", H, W, h, w, B, transfer);
  if (st != PTX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  if (!grad_sample || !grad_probe) {
    set_error("ptx_nearfield_step: null gradient output");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  timing_begin(s);
  if (B == 0) {
    if (!(flags & PTX_FLAG_ACCUMULATE)) {
      fill_zero2((float2*)grad_sample, H * W, s);
      fill_zero2((float2*)grad_probe, h * w, s);
    }
    return check_launch("zero fill", s);
  }
  if (!sample || !probe || !pos || !measured || !loss_per_position) {
    set_error("ptx_nearfield_step: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  ScatterPlan sp{};
  dispatch_ne<ScatterPlanQuery>(PASS_ROW_LOOP, w, h, B, &sp);
  const ScanWorkspace ws = carve_scan(workspace, H, W, h, w, B, flags, true, sp, h);
  if (!workspace || ws_bytes < ws.bytes) {
    set_error("ptx_nearfield_step: workspace of %zu bytes needed, %zu given", ws.bytes, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  const float inv_n = 1.0f / (float)(h * w);
  ScanGatherRows<false> r1{(const float2*)sample, H, W, (const float2*)probe, (int)h, pos, ws.tb};
  if ((st = dispatch_ne<GatherRowsRun>(PASS_ROW, w, r1, B * h, s)) != PTX_OK) return st;
  if ((st = dispatch_ne<TransferColsRun>(PASS_COL, h, ws.tb, (const float2*)transfer, B, w, false, s)) != PTX_OK) return st;
  if ((st = dispatch_ne<StepRowsRun>(PASS_ROW, w, ws.tb, measured, amplitude, ws.loss_part, B * h, 2.0f * loss_scale, inv_n,
                                    s)) != PTX_OK)
    return st;
  if ((st = dispatch_ne<TransferColsRun>(PASS_COL, h, ws.tb, (const float2*)transfer, B, w, true, s)) != PTX_OK) return st;
  const LossRows loss{ws.loss_part, h, inv_n, loss_per_position, B};
  return scan_scatter<false>(ws, sp, (const float2*)sample, H, W, (const float2*)probe, h, w, pos, B,
                             (float2*)grad_sample, (float2*)grad_probe, flags, inv_n, s, loss);
}

}  // extern "C"
