// Conventional ptychography (ptychoX_base): forward, adjoint and the fused
// amplitude-loss training step.
//
// Reference semantics: python/ptychoX/base/ops.py:60-66 (forward) and 93-105
// (backward); the CUDA extension python/ptychoX/csrc/base_kernel.cu:302-357.
//
// For even frame sides, fftshift(F x) = F(x * (-1)^(r+c)) and
// N ifft2(ifftshift(G)) = (-1)^(r+c) * N ifft2(G), so both shifts become a sign
// applied in registers in the gather / scatter passes (no index remap pass).
//
// Passes (h x w frames, B views):
//   forward:  R1 rows  : gather sample window * probe * sign, row FFT   -> far_field (scratch)
//             C1 cols  : column FFT, amplitude                          -> far_field, amplitude
//   backward: C2 cols  : Gamma = g Phi/max(|Phi|,eps), column iFFT     -> ws
//             R2 rows  : row iFFT, sign, chi conj(p) -> object (atomics or per-view buffer),
//                        chi conj(patch) accumulated over a chunk of views in registers -> probe partials
//   step:     R1, then C12 cols: column FFT, |Phi|, loss partials, Gamma, column iFFT (in place), then R2.
#include "buffer_io.hpp"
#include "scan_common.hpp"

namespace ptx {
namespace {

// ---------------------------------------------------------------- C1: column FFT + amplitude
struct BaseFarFieldCols {
  float2* ff;  // in: row-transformed, out: far field (fftshifted)
  float* amp;
  int64_t width;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const float2* p = ff + frame * N * width;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = p[(t + q * T) * (int)width + (int)col];
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const int64_t off = frame * N * width;  // frame base (uniform)
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int64_t k = off + ((t + q * T) * (int)width + (int)col);
      ff[k] = v[q];
      amp[k] = c_abs(v[q]);
    }
  }
};

// ---------------------------------------------------------------- C2: phase * g, column iFFT
struct BaseGammaCols {
  const float2* ff;
  const float* g;
  float2* out;
  int64_t width;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const int64_t off = frame * N * width;  // frame base (uniform)
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int64_t k = off + ((t + q * T) * (int)width + (int)col);
      const float2 phi = ff[k];
      const float s = g[k] * fast_rcp(fmaxf(c_abs(phi), kFltEps));
      v[q] = c_scale(phi, s);
    }
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    float2* p = out + frame * N * width;
#pragma unroll
    for (int q = 0; q < E; ++q) p[(t + q * T) * (int)width + (int)col] = v[q];
  }
};

// ---------------------------------------------------------------- C12: fused forward/loss/adjoint column pass
// Persistent and software-pipelined: one resident work-group per CU walks the
// column tiles blockIdx.x, + gridDim.x, ...; while it transforms tile k one of
// two register sets already holds the loads of tile k + gridDim.x (row-FFT
// output and measured amplitudes), so the HBM traffic of the next tile
// overlaps the LDS exchanges of this one; the sets alternate, nothing is
// copied between them.  Buffer accesses keep the addressing to one VGPR per
// lane.  Per tile: column FFT, A = |Phi|, loss partials (one per wave),
// Gamma = gscale (A - Y) Phi / max(A, eps), column iFFT.
//
// The exchanges alternate between two LDS regions, one barrier each: with an
// even number of exchanges per transform, consecutive exchanges never share a
// region, so no barrier separates the transforms or the tiles; with an odd
// number one barrier follows each transform.
template <int N, int E, int C>
struct StepColsShape {
  static constexpr int T = ColCfg<N, E>::T;
  static constexpr int THREADS = T * C;
  static constexpr int WAVES = (THREADS + 63) / 64;
  static constexpr int EXCHANGES = FftShape<N, E>::STAGES - 1;
  static constexpr size_t LDS_BYTES =
      (2 * (size_t)C * col_stride<N, C>() + LoopTwiddles<N, E>::TABLE) * sizeof(float2);
};

template <int N, int E, int C, bool AMP>
__global__ void __launch_bounds__((ColCfg<N, E>::T * C))
    k_base_step_cols(float2* __restrict__ tb, const float* __restrict__ measured, const int64_t* __restrict__ order,
                     int64_t count, float* __restrict__ amp_out, float* __restrict__ loss_part, int64_t width,
                     int64_t tiles_per_frame, int64_t ntiles, float gscale) {
  using Sh = StepColsShape<N, E, C>;
  constexpr int T = Sh::T;
  static_assert(Sh::LDS_BYTES <= 160 * 1024, "LDS budget");
  __shared__ float2 lds0[C * col_stride<N, C>()];
  __shared__ float2 lds1[C * col_stride<N, C>()];
  __shared__ float2 s_tw[LoopTwiddles<N, E>::TABLE];
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int wave = threadIdx.x >> 6;
  const bool lane0 = (threadIdx.x & 63) == 0;
  const int64_t fsz = (int64_t)N * width;               // elements per frame
  const uint32_t fbytes = (uint32_t)(fsz * 8);
  const uint32_t lane = (uint32_t)(t * (int)width + c);  // element offset inside a tile (row t, column c)
  const uint32_t rstep = (uint32_t)(T * width);          // elements from row t + q T to t + (q + 1) T
  const unsigned nt = (unsigned)ntiles;
  LoopTwiddles<N, E> tw;  // thread t holds the same rows of every tile
  tw.init(t, s_tw);
  // Loads of tile kk: the frame rows q in [q0, q1) into dv (fetch_t), the
  // measured amplitudes into dm (fetch_m); the measured frame of a view is the
  // batch's own (order == null) or measured[order[frame]].  A tile past the
  // end gets zero-length buffers: its loads are issued but return zero
  // without touching memory, so the loop's memory instructions are the same
  // on every path and the compiler's wait counts (vmcnt) let the prefetch
  // stay in flight instead of draining it at every tile.
  // (tile and frame indices are compared as 32-bit scalars: a 64-bit ordered
  // compare needs vector registers, whose reuse would wait on in-flight stores)
  auto fetch_t = [&](int64_t kk, float2 (&dv)[E], int q0, int q1) {
    const bool live = (unsigned)kk < nt;
    const TileIdx ti = split_tile(xcd_paired_block(live ? (unsigned)kk : 0u, nt), tiles_per_frame);
    const uint32_t col0 = (uint32_t)(ti.tile * C);
    const auto rt = buffer_rsrc(tb + ti.frame * fsz, live ? fbytes : 0u);
#pragma unroll
    for (int q = 0; q < E; ++q)
      if (q >= q0 && q < q1) dv[q] = buf_load_f2(rt, lane * 8u, (q * rstep + col0) * 8u);
  };
  auto fetch_m = [&](int64_t kk, float (&dm)[E]) {
    const bool live = (unsigned)kk < nt;
    const TileIdx ti = split_tile(xcd_paired_block(live ? (unsigned)kk : 0u, nt), tiles_per_frame);
    const uint32_t col0 = (uint32_t)(ti.tile * C);
    int64_t mframe = ti.frame;
    if (order) {
      const uint64_t o = (uint64_t)order[ti.frame];
      // out of range: frame 0 (the positions of such a view are entry 0's too)
      mframe = ((o >> 32) == 0 && (uint32_t)o < (uint32_t)count) ? (int64_t)o : 0;
    }
    const auto rm = buffer_rsrc(measured + mframe * fsz, live ? fbytes / 2 : 0u);
#pragma unroll
    for (int q = 0; q < E; ++q) dm[q] = buf_load_f1(rm, lane * 4u, (q * rstep + col0) * 4u);
  };
  // Tile kk from (v, m); the loads of the next tile knext are spread over
  // this tile's forward transform: its first E/2 frame rows before this call,
  // the other frame rows after the transform's first stage, its measured
  // amplitudes after the transform.  Issued all at once, a tile's 2E loads
  // filled the address queue of the memory pipeline, and waves stalled at
  // their issue instead of starting the transform (column pass 69-71 -> 62
  // us at 2048^2/512^2/64 when spread this way).
  auto process = [&](int64_t kk, float2 (&v)[E], float (&m)[E], int64_t knext, float2 (&vnext)[E],
                     float (&mnext)[E]) {
    const TileIdx ti = split_tile(xcd_paired_block((unsigned)kk, nt), tiles_per_frame);
    const uint32_t col0 = (uint32_t)(ti.tile * C);
    // the exchange addresses depend only on (t, c); recompute them per tile
    // instead of letting them be hoisted out of the loop and held in registers
    int tt = t, cc = c;
    asm volatile("" : "+v"(tt), "+v"(cc));
    float2* l0 = &lds0[cc * col_stride<N, C>()];
    float2* l1 = &lds1[cc * col_stride<N, C>()];
    tw.template stages<-1, true, 0, 1>(v, tt, l0, l1);
    fetch_t(knext, vnext, E / 2, E);
    tw.template stages<-1, true, 1>(v, tt, l0, l1);
    if constexpr (Sh::EXCHANGES % 2 == 1) lds_barrier();
    fetch_m(knext, mnext);
    float lsum = 0.0f;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const AbsInv ai = c_abs_inv(v[q], 1.0f / kFltEps);
      const float d = ai.a - m[q];
      lsum += d * d;
      v[q] = c_scale(v[q], gscale * d * ai.inv);
      m[q] = ai.a;
    }
    // amplitudes (AMP: requested) and the wave's loss partial (lane 0; the
    // other lanes' offsets fall outside the 4-byte buffer): stores without
    // branches, for the same reason as above
    if constexpr (AMP) {
      const auto ra = buffer_rsrc(amp_out + ti.frame * fsz, fbytes / 2);
#pragma unroll
      for (int q = 0; q < E; ++q) buf_store_f1(m[q], ra, lane * 4u, (q * rstep + col0) * 4u);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      if (off < Sh::THREADS) lsum += __shfl_xor(lsum, off, 64);
    buf_store_f1(lsum, buffer_rsrc(loss_part + (ti.frame * tiles_per_frame + ti.tile) * Sh::WAVES + wave, 4u),
                 lane0 ? 0u : 64u, 0u);
    tw.template stages<+1>(v, tt, l0, l1);
    if constexpr (Sh::EXCHANGES % 2 == 1) lds_barrier();
    const auto rt = buffer_rsrc(tb + ti.frame * fsz, fbytes);
#pragma unroll
    for (int q = 0; q < E; ++q) buf_store_f2(v[q], rt, lane * 8u, (q * rstep + col0) * 8u);
  };
  // The first tile is peeled so that every path into the loop has issued the
  // same memory instructions (two tiles' loads, one tile's stores): the wait
  // for a tile's loads then never has to cover the previous tile's stores.
  float2 va[E], vb[E];
  float ma[E], mb[E];
  int64_t k = blockIdx.x;  // < ntiles: the grid never exceeds the tile count
  fetch_t(k, va, 0, E);
  fetch_m(k, ma);
  fetch_t(k + gridDim.x, vb, 0, E / 2);  // the next tile's first half, in flight during this tile
  process(k, va, ma, k + gridDim.x, vb, mb);
  k += gridDim.x;
  while ((unsigned)k < nt) {
    fetch_t(k + gridDim.x, va, 0, E / 2);
    process(k, vb, mb, k + gridDim.x, va, ma);
    k += gridDim.x;
    if ((unsigned)k >= nt) break;
    fetch_t(k + gridDim.x, vb, 0, E / 2);
    process(k, va, ma, k + gridDim.x, vb, mb);
    k += gridDim.x;
  }
}

// ---------------------------------------------------------------- host launchers

template <int N, int E>
struct FarFieldColsRun {
  static ptx_status run(const BaseFarFieldCols& op, int64_t frames, hipStream_t s) {
    launch_cols<N, E, -1>(op, frames, op.width, s);
    return check_launch("base far-field cols", s);
  }
};
template <int N, int E>
struct GammaColsRun {
  static ptx_status run(const BaseGammaCols& op, int64_t frames, hipStream_t s) {
    launch_cols<N, E, +1>(op, frames, op.width, s);
    return check_launch("base gamma cols", s);
  }
};


// Columns per work-group of k_base_step_cols: ColCfg's CMAX, or 8 for frames
// narrower than 16 columns.
template <int N, int E>
inline int step_cols_c(int64_t width) {
  using Cf = ColCfg<N, E>;
  return (Cf::CMAX > 8 && width < Cf::CMAX) ? 8 : Cf::CMAX;
}

// loss partials per frame written by k_base_step_cols: one per wave of each column tile
template <int N, int E>
struct StepLossPartsQuery {
  static ptx_status run(int64_t width, int64_t* parts) {
    using Cf = ColCfg<N, E>;
    const int c = step_cols_c<N, E>(width);
    *parts = (width / c) * ((Cf::T * c + 63) / 64);
    return PTX_OK;
  }
};

template <int N, int E, int C>
inline void launch_step_cols(float2* tb, const float* measured, const int64_t* order, int64_t count, float* amp,
                             float* loss_part, int64_t frames, int64_t width, float gscale, hipStream_t s) {
  using Sh = StepColsShape<N, E, C>;
  const int64_t tiles = width / C;
  int per_cu = 1024 / (Sh::T * C);
  const int by_lds = (int)((160 * 1024) / Sh::LDS_BYTES);
  if (by_lds < per_cu) per_cu = by_lds;
  const int64_t grid = persistent_grid(frames * tiles, per_cu);
  if (amp)
    k_base_step_cols<N, E, C, true><<<dim3((unsigned)grid), dim3(Sh::T * C), 0, s>>>(
        tb, measured, order, count, amp, loss_part, width, tiles, frames * tiles, gscale);
  else
    k_base_step_cols<N, E, C, false><<<dim3((unsigned)grid), dim3(Sh::T * C), 0, s>>>(
        tb, measured, order, count, amp, loss_part, width, tiles, frames * tiles, gscale);
}

template <int N, int E>
struct StepColsRun {
  static ptx_status run(float2* tb, const float* measured, const int64_t* order, int64_t count, float* amp,
                        float* loss_part, int64_t frames, int64_t width, float gscale, hipStream_t s) {
    using Cf = ColCfg<N, E>;
    if (frames * (width / 4) >= (int64_t(1) << 31)) {  // the kernel counts tiles in 32 bits
      set_error("base step cols: %lld frames of width %lld exceed the tile count limit", (long long)frames,
                (long long)width);
      return PTX_ERR_UNSUPPORTED;
    }
    if constexpr (Cf::CMAX > 8) {
      if (width < Cf::CMAX) {
        launch_step_cols<N, E, 8>(tb, measured, order, count, amp, loss_part, frames, width, gscale, s);
        return check_launch("base step cols", s);
      }
    }
    launch_step_cols<N, E, Cf::CMAX>(tb, measured, order, count, amp, loss_part, frames, width, gscale, s);
    return check_launch("base step cols", s);
  }
};

// ---------------------------------------------------------------- whole-frame fused step (frames <= 128^2)
// For N x N frames with N <= 128 the whole frame (N^2 complex, 128 KiB at
// N = 128) fits in LDS, so one work-group runs a view's complete step --
// gather, 2-D FFT, amplitude, loss, Gamma, 2-D inverse FFT, window
// contribution and probe partial -- and the frame never goes to HBM between
// the transforms (the three-pass path writes and reads it twice).  A
// work-group has N^2/16 threads holding 16 elements each (1024 threads at
// N = 128): thread (s, t) = (tid / T, tid % T), T = N/16, holds elements
// t + q T of sequence s.  Two orientations:
//   column orientation: sequence s is column s of the frame (element = row);
//   row orientation:    sequence s is row s (element = column).
// Global loads and stores in either orientation touch T x 8 contiguous bytes
// per sequence per instruction (64 B at N = 128).  The frame buffer in LDS
// (row stride lds_len<N>, the padded per-sequence layout of the FFT core)
// serves as every transform's exchange region and as the transpose buffer:
//   window * probe * sign loaded in column orientation -> column FFT ->
//   transpose -> row FFT (far field, fftshift by the sign) -> |Phi|, loss,
//   Gamma (measured amplitudes in row orientation) -> row iFFT -> transpose ->
//   column iFFT -> chi = sign * (.) in column orientation -> contribution
//   chi conj(probe) and probe partial chi conj(window).
// The probe partial sum over the work-group's chunk of views is kept in the
// work-group's own slice of gp_part (read-modify-write after the first view).  The loss of
// each view is a fixed-order work-group sum.  Work-groups past the views'
// build the tile lists of the sorted-segment gather, one tile per wave.
template <int N>
struct FrameCfg {
  static constexpr int E = 16;
  static constexpr int T = N / E;
  static constexpr int THREADS = N * T;
  static constexpr int LS = lds_len<N>();
  static_assert(N >= 32 && N <= 128, "whole-frame step: N in [32, 128]");
  static_assert((size_t)N * LS * sizeof(float2) <= 150 * 1024, "frame must fit in LDS");
};

template <int N>
__global__ void __launch_bounds__(FrameCfg<N>::THREADS)
    k_base_frame_step(const float2* __restrict__ sample, int64_t H, int64_t W, const float2* __restrict__ probe,
                      const int64_t* __restrict__ pos, const float* __restrict__ measured,
                      const int64_t* __restrict__ order, int64_t count, int64_t B, int chunk, int nchunks,
                      float gscale, float loss_scale, float* __restrict__ loss_out, float* __restrict__ amp_out,
                      float2* __restrict__ gobj, float2* __restrict__ contrib, float2* __restrict__ gp_part,
                      TileLists lists) {
  using Cf = FrameCfg<N>;
  constexpr int E = Cf::E, T = Cf::T, LS = Cf::LS;
  __shared__ float2 frame[N * LS];
  __shared__ float red[Cf::THREADS / 64 + 1];
  if ((int)blockIdx.x >= nchunks) {  // side work: tile lists of the gather, one tile per wave
    const int64_t tile = ((int64_t)blockIdx.x - nchunks) * (Cf::THREADS / 64) + (threadIdx.x >> 6);
    if (tile < lists.ntiles) lists.run_wave(tile);
    return;
  }
  const int s = (int)(threadIdx.x / T);  // sequence (column, then row)
  const int t = (int)(threadIdx.x % T);
  const int64_t b0 = (int64_t)blockIdx.x * chunk;
  const int64_t b1 = b0 + chunk < B ? b0 + chunk : B;
  constexpr uint32_t NN = (uint32_t)N * N;
  // Buffer addressing: one 32-bit lane offset per access pattern, the element
  // index's q term uniform (the sample below 2 GiB, checked by the host).
  const uint32_t lane_c = (uint32_t)(t * N + s);  // column orientation, frame layout (row j = t + q T, column s)
  const uint32_t lane_r = (uint32_t)(s * N + t);  // row orientation (row s, column t + q T)
  const uint32_t obj_bytes = (uint32_t)(H * W * 8);
  // the probe is read from L2 where it is needed (column orientation,
  // fftshift sign folded in: p[j][s] (-1)^(j+s), j = t + q T); keeping it in
  // registers for the chunk would exceed the 128-register budget of a
  // 1024-thread work-group
  const auto rp = buffer_rsrc(probe, NN * 8u);
  auto probe_signed = [&](int q) {
    const float2 p = buf_load_f2(rp, lane_c * 8u, (uint32_t)(q * T * N) * 8u);
    return ((t + q * T + s) & 1) ? c_neg(p) : p;
  };
  const auto rgp = buffer_rsrc(gp_part + (int64_t)blockIdx.x * NN, NN * 8u);
  bool first = true;  // no view of the chunk has written the probe slice yet
  for (int64_t b = b0; b < b1; ++b) {
    const int64_t x0 = pos[2 * b], y0 = pos[2 * b + 1];
    const bool ok = window_ok(x0, y0, N, N, H, W);
    int64_t mframe = b;
    if (order) {
      const uint64_t o = (uint64_t)order[b];
      mframe = ((o >> 32) == 0 && (uint32_t)o < (uint32_t)count) ? (int64_t)o : 0;
    }
    // window * probe * sign, column orientation (an out-of-range window reads zero)
    const auto rs = buffer_rsrc(sample, ok ? obj_bytes : 0u);
    const uint32_t wlane = ok ? (uint32_t)(((y0 + t) * W + x0 + s) * 8) : 0u;
    float2 v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = c_mul(buf_load_f2(rs, wlane, (uint32_t)(q * T * W * 8)), probe_signed(q));
    lds_barrier();  // the previous view's last reads of the frame buffer are done
    fft_seq<N, E, -1>(v, t, &frame[s * LS]);  // column FFT
    lds_barrier();
#pragma unroll
    for (int q = 0; q < E; ++q) frame[(t + q * T) * LS + lds_pad(s)] = v[q];  // transpose
    lds_barrier();
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = frame[s * LS + lds_pad(t + q * T)];  // row s
    lds_barrier();
    fft_seq<N, E, -1>(v, t, &frame[s * LS]);  // row FFT: Phi[s][t + q T]
    const auto rm = buffer_rsrc(measured + (int64_t)mframe * NN, NN * 4u);
    // amplitudes (when requested) through a zero-length buffer otherwise: no branch around the stores
    const auto ra = buffer_rsrc(amp_out ? amp_out + (int64_t)b * NN : amp_out, amp_out ? NN * 4u : 0u);
    float lsum = 0.0f;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const float y = buf_load_f1(rm, lane_r * 4u, (uint32_t)(q * T) * 4u);
      const AbsInv ai = c_abs_inv(v[q], 1.0f / kFltEps);
      const float d = ai.a - y;
      lsum += d * d;
      v[q] = c_scale(v[q], gscale * d * ai.inv);
      buf_store_f1(ai.a, ra, lane_r * 4u, (uint32_t)(q * T) * 4u);
    }
    // loss of the view: fixed-order work-group sum
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
      if (off < Cf::THREADS) lsum += __shfl_xor(lsum, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = lsum;
    lds_barrier();
    fft_seq<N, E, +1>(v, t, &frame[s * LS]);  // row iFFT
    if (threadIdx.x == 0) {
      float tot = 0.0f;
      for (int w = 0; w < Cf::THREADS / 64; ++w) tot += red[w];
      loss_out[b] = tot * loss_scale;
    }
    lds_barrier();
#pragma unroll
    for (int q = 0; q < E; ++q) frame[(t + q * T) * LS + lds_pad(s)] = v[q];  // transpose
    lds_barrier();
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = frame[s * LS + lds_pad(t + q * T)];  // column s
    lds_barrier();
    fft_seq<N, E, +1>(v, t, &frame[s * LS]);  // column iFFT: (-1)^(j+s) chi[j][s]
    if (!ok) continue;
    // contributions and probe partial, column orientation; the sign of chi is
    // the sign folded into pr, so chi conj(p) = v conj(pr)
    const auto rc = buffer_rsrc(contrib ? contrib + (int64_t)b * NN : contrib, contrib ? NN * 8u : 0u);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int j = t + q * T;
      const uint32_t qoff = (uint32_t)(q * T * N) * 8u;
      const float2 val = c_mulc(v[q], probe_signed(q));
      if (contrib) {
        buf_store_f2(val, rc, lane_c * 8u, qoff);
      } else {
        float2* g = gobj + (y0 + j) * W + x0 + s;
        unsafeAtomicAdd(&g->x, val.x);
        unsafeAtomicAdd(&g->y, val.y);
      }
      const float2 chi = ((j + s) & 1) ? c_neg(v[q]) : v[q];
      const float2 pp = c_mulc(chi, buf_load_f2(rs, wlane, (uint32_t)(q * T * W * 8)));
      // each thread owns its elements of the chunk's probe slice
      buf_store_f2(first ? pp : c_add(buf_load_f2(rgp, lane_c * 8u, qoff), pp), rgp, lane_c * 8u, qoff);
    }
    first = false;
  }
  if (first) {  // every view of the chunk was out of range: the slice is zero
#pragma unroll
    for (int q = 0; q < E; ++q) buf_store_f2(c_zero(), rgp, lane_c * 8u, (uint32_t)(q * T * N) * 8u);
  }
}

// Whole-frame step: the fused kernel (with the tile lists riding along), then
// the object-gradient gather with the probe-partial sum riding along, or the
// atomic scatter's zero fill before and the partial sum after.
template <int N>
ptx_status frame_step_launch(const ScanWorkspace& ws, const ScatterPlan& sp, const float2* sample, int64_t H,
                             int64_t W, const float2* probe, const int64_t* pos, const float* measured,
                             const int64_t* order, int64_t count, int64_t B, float loss_scale, float* loss,
                             float* amp, float2* gs, float2* gp, unsigned flags, hipStream_t s,
                             const AdamArgs& adam_obj, const AdamArgs& adam_probe) {
  using Cf = FrameCfg<N>;
  const bool det = use_segments(flags);
  const bool acc = flags & PTX_FLAG_ACCUMULATE;
  ptx_status st;
  if (!det && !acc) {
    fill_zero2(gs, H * W, s);
    if ((st = check_launch("zero fill", s)) != PTX_OK) return st;
  }
  TileLists lists{};
  if (det) lists = gather_lists_work(ws.gather, pos, B, N, N, H, W, 0, 0, true);
  constexpr int kWaves = Cf::THREADS / 64;
  const int64_t side = (lists.ntiles + kWaves - 1) / kWaves;
  k_base_frame_step<N><<<dim3((unsigned)(sp.nchunks + side)), dim3(Cf::THREADS), 0, s>>>(
      sample, H, W, probe, pos, measured, order, count, B, sp.chunk, sp.nchunks, 2.0f * loss_scale,
      1.0f / (float)(N * N), loss, amp, gs, det ? ws.tb : nullptr, ws.gp_part, lists);
  if ((st = check_launch("base frame step", s)) != PTX_OK) return st;
  SumParts2 probe_sum{ws.gp_part, sp.nchunks, (int64_t)N * N, gp, acc ? 1 : 0, 0};
  probe_sum.adam = adam_probe;
  if (!det) {
    if (probe_sum.blocks() > 0) k_sum_parts2<<<dim3((unsigned)probe_sum.blocks()), dim3(256), 0, s>>>(probe_sum);
    if ((st = check_launch("probe partial sum", s)) != PTX_OK) return st;
    if (adam_obj.param) return launch_adam(adam_obj, gs, H * W, s);  // after the atomic scatter
    return PTX_OK;
  }
  SideWork with_gather{};
  with_gather.sums = probe_sum;
  return run_tile_gather(ws.gather, ws.tb, N, N, pos, B, 0, 0, gs, H, W, acc ? 1 : 0, s, with_gather, adam_obj);
}

// Which path a square frame takes.  The whole-frame kernel serialises a
// view's work on one compute unit; measured against the three passes on
// MI355X (objects 4x the probe side, batches 32..1024) it is 4-8 % faster at
// 32^2, within -12..+7 % at 64^2 and 1.5-2.3x slower at 128^2 (there the
// frame's 16 values per thread exceed the 128-register budget of a
// 1024-thread work-group and spill), so it is the default at 32^2 only.
// PTX_FLAG_WHOLE_FRAME forces it for 32..128, PTX_FLAG_THREE_PASS forbids it.
inline bool frame_step_eligible(int64_t H, int64_t W, int64_t h, int64_t w, unsigned flags) {
  if (flags & PTX_FLAG_THREE_PASS) return false;
  if (h != w || H * W * 8 >= (int64_t(1) << 31)) return false;  // 32-bit buffer offsets into the sample
  if (flags & PTX_FLAG_WHOLE_FRAME) return h == 32 || h == 64 || h == 128;
  return h == 32;
}

}  // namespace
}  // namespace ptx

using namespace ptx;

extern "C" {

ptx_status ptx_base_forward(const void* sample, int64_t H, int64_t W, const void* probe, int64_t h, int64_t w,
                            const int64_t* pos, int64_t B, void* far_field, float* amplitude, void* stream) {
  TuneScope tune_scope;
  ptx_status st = check_shapes("ptx_base_forward", H, W, h, w, B);
  if (st != PTX_OK) return st;
  if (B == 0) return PTX_OK;
  if (!sample || !probe || !pos || !far_field || !amplitude) {
    set_error("ptx_base_forward: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  timing_begin(s);
  ScanGatherRows<true> r1{(const float2*)sample, H, W, (const float2*)probe, (int)h, pos, (float2*)far_field};
  st = dispatch_ne<GatherRowsRun>(PASS_ROW, w, r1, B * h, s);
  if (st != PTX_OK) return st;
  BaseFarFieldCols c1{(float2*)far_field, amplitude, w};
  return dispatch_ne<FarFieldColsRun>(PASS_COL, h, c1, B, s);
}

size_t ptx_base_backward_workspace(int64_t H, int64_t W, int64_t h, int64_t w, int64_t B, unsigned flags) {
  TuneScope tune_scope;
  if (check_shapes("ptx_base_backward_workspace", H, W, h, w, B) != PTX_OK) return 0;
  ScatterPlan sp{};
  dispatch_ne<ScatterPlanQuery>(PASS_ROW_LOOP, w, h, B, &sp);
  return carve_scan(nullptr, H, W, h, w, B, flags, false, sp, 0).bytes;
}

ptx_status ptx_base_backward(const float* grad_output, const void* far_field, const void* sample, int64_t H, int64_t W,
                             const void* probe, int64_t h, int64_t w, const int64_t* pos, int64_t B, void* grad_sample,
                             void* grad_probe, void* workspace, size_t ws_bytes, unsigned flags, void* stream) {
  TuneScope tune_scope;
  ptx_status st = check_shapes("ptx_base_backward", H, W, h, w, B);
  if (st != PTX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  if (!grad_sample || !grad_probe) {
    set_error("ptx_base_backward: null gradient output");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (B == 0) {
    if (!(flags & PTX_FLAG_ACCUMULATE)) {
      fill_zero2((float2*)grad_sample, H * W, s);
      fill_zero2((float2*)grad_probe, h * w, s);
    }
    return check_launch("zero fill", s);
  }
  if (!grad_output || !far_field || !sample || !probe || !pos) {
    set_error("ptx_base_backward: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  ScatterPlan sp{};
  dispatch_ne<ScatterPlanQuery>(PASS_ROW_LOOP, w, h, B, &sp);
  const ScanWorkspace ws = carve_scan(workspace, H, W, h, w, B, flags, false, sp, 0);
  if (!workspace || ws_bytes < ws.bytes) {
    set_error("ptx_base_backward: workspace of %zu bytes needed, %zu given", ws.bytes, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  timing_begin(s);
  BaseGammaCols c2{(const float2*)far_field, grad_output, ws.tb, w};
  st = dispatch_ne<GammaColsRun>(PASS_COL, h, c2, B, s);
  if (st != PTX_OK) return st;
  return scan_scatter<true>(ws, sp, (const float2*)sample, H, W, (const float2*)probe, h, w, pos, B,
                            (float2*)grad_sample, (float2*)grad_probe, flags, 1.0f, s);
}

size_t ptx_base_step_workspace(int64_t H, int64_t W, int64_t h, int64_t w, int64_t B, unsigned flags) {
  TuneScope tune_scope;
  if (check_shapes("ptx_base_step_workspace", H, W, h, w, B) != PTX_OK) return 0;
  ScatterPlan sp{};
  dispatch_ne<ScatterPlanQuery>(PASS_ROW_LOOP, w, h, B, &sp);
  int64_t loss_parts = 0;
  dispatch_ne<StepLossPartsQuery>(PASS_COL, h, w, &loss_parts);
  // room for the indexed variant's gathered positions too
  return carve_scan(nullptr, H, W, h, w, B, flags, true, sp, loss_parts, true).bytes;
}

}  // extern "C"

namespace ptx {

// Shared body of ptx_base_step, ptx_base_step_indexed and
// ptx_base_step_noalloc.  order == null: the batch's own positions / measured
// frames; otherwise view b is order[b] of the full tables
// positions_all[count] / measured_all[count].
ptx_status base_step_impl(const char* fn, const void* sample, int64_t H, int64_t W, const void* probe, int64_t h,
                          int64_t w, const int64_t* pos, int64_t count, const float* measured, const int64_t* order,
                          int64_t B, float loss_scale, float* loss_per_position, float* amplitude, void* grad_sample,
                          void* grad_probe, void* workspace, size_t ws_bytes, unsigned flags, hipStream_t s,
                          const AdamArgs* adam_obj, const AdamArgs* adam_probe) {
  ptx_status st = check_shapes(fn, H, W, h, w, B);
  if (st != PTX_OK) return st;
  if (!grad_sample || !grad_probe) {
    set_error("%s: null gradient output", fn);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  const AdamArgs ao = adam_obj ? *adam_obj : AdamArgs{};
  const AdamArgs ap = adam_probe ? *adam_probe : AdamArgs{};
  if ((ao.param || ap.param) && (flags & PTX_FLAG_ACCUMULATE)) {
    set_error("%s: the fused Adam update takes the whole gradient (no PTX_FLAG_ACCUMULATE)", fn);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (B == 0) {
    if (!(flags & PTX_FLAG_ACCUMULATE)) {
      fill_zero2((float2*)grad_sample, H * W, s);
      fill_zero2((float2*)grad_probe, h * w, s);
    }
    if ((st = check_launch("zero fill", s)) != PTX_OK) return st;
    if (ao.param && (st = launch_adam(ao, (const float2*)grad_sample, H * W, s)) != PTX_OK) return st;
    if (ap.param) return launch_adam(ap, (const float2*)grad_probe, h * w, s);
    return PTX_OK;
  }
  if (!sample || !probe || !pos || !measured || !loss_per_position) {
    set_error("%s: null pointer argument", fn);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  ScatterPlan sp{};
  dispatch_ne<ScatterPlanQuery>(PASS_ROW_LOOP, w, h, B, &sp);
  int64_t loss_parts = 0;
  dispatch_ne<StepLossPartsQuery>(PASS_COL, h, w, &loss_parts);
  const ScanWorkspace ws = carve_scan(workspace, H, W, h, w, B, flags, true, sp, loss_parts, order != nullptr);
  if (!workspace || ws_bytes < ws.bytes) {
    set_error("%s: workspace of %zu bytes needed, %zu given", fn, ws.bytes, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  timing_begin(s);
  const bool frame = frame_step_eligible(H, W, h, w, flags);
  // an indexed batch's positions are gathered by the gather-rows kernel itself
  // when it is the view-looping one, else by a kernel of their own first
  const bool order_in_gather = order && !frame && gather_loop_ok(H, W, h, w);
  const int64_t* bpos = order ? ws.pos_batch : pos;
  if (order && !order_in_gather &&
      (st = launch_gather_positions(order, pos, count, B, ws.pos_batch, nullptr, s)) != PTX_OK)
    return st;
  if (frame) {
    auto* smp = (const float2*)sample;
    auto* prb = (const float2*)probe;
    auto* gsp = (float2*)grad_sample;
    auto* gpp = (float2*)grad_probe;
    switch (h) {
      case 32:
        return frame_step_launch<32>(ws, sp, smp, H, W, prb, bpos, measured, order, count, B, loss_scale,
                                     loss_per_position, amplitude, gsp, gpp, flags, s, ao, ap);
      case 64:
        return frame_step_launch<64>(ws, sp, smp, H, W, prb, bpos, measured, order, count, B, loss_scale,
                                     loss_per_position, amplitude, gsp, gpp, flags, s, ao, ap);
      default:
        return frame_step_launch<128>(ws, sp, smp, H, W, prb, bpos, measured, order, count, B, loss_scale,
                                      loss_per_position, amplitude, gsp, gpp, flags, s, ao, ap);
    }
  }
  ScanGatherRows<true> r1{(const float2*)sample, H, W, (const float2*)probe, (int)h, bpos, ws.tb};
  if (order_in_gather) {
    r1.pos = pos;
    r1.order = order;
    r1.count = count;
    r1.pos_out = ws.pos_batch;
  }
  st = dispatch_ne<GatherRowsRun>(PASS_ROW, w, r1, B * h, s);
  if (st != PTX_OK) return st;
  st = dispatch_ne<StepColsRun>(PASS_COL, h, ws.tb, measured, order, count, amplitude, ws.loss_part, B, w,
                                2.0f * loss_scale, s);
  if (st != PTX_OK) return st;
  const LossRows loss{ws.loss_part, loss_parts, 1.0f / (float)(h * w), loss_per_position, B};
  return scan_scatter<true>(ws, sp, (const float2*)sample, H, W, (const float2*)probe, h, w, bpos, B,
                            (float2*)grad_sample, (float2*)grad_probe, flags, 1.0f, s, loss, ao, ap);
}

}  // namespace ptx

extern "C" {

ptx_status ptx_base_step(const void* sample, int64_t H, int64_t W, const void* probe, int64_t h, int64_t w,
                         const int64_t* pos, int64_t B, const float* measured, float loss_scale,
                         float* loss_per_position, float* amplitude, void* grad_sample, void* grad_probe,
                         void* workspace, size_t ws_bytes, unsigned flags, void* stream) {
  TuneScope tune_scope;
  return base_step_impl("ptx_base_step", sample, H, W, probe, h, w, pos, B, measured, nullptr, B, loss_scale,
                        loss_per_position, amplitude, grad_sample, grad_probe, workspace, ws_bytes, flags,
                        (hipStream_t)stream);
}

ptx_status ptx_base_step_indexed(const void* sample, int64_t H, int64_t W, const void* probe, int64_t h, int64_t w,
                                 const int64_t* positions_all, const float* measured_all, int64_t count,
                                 const int64_t* order, int64_t B, float loss_scale, float* loss_per_position,
                                 float* amplitude, void* grad_sample, void* grad_probe, void* workspace,
                                 size_t ws_bytes, unsigned flags, void* stream) {
  TuneScope tune_scope;
  if (B > 0 && !order) {
    set_error("ptx_base_step_indexed: null view order");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  // the kernels index the tables with 32-bit unsigned compares; an empty
  // table has no entry 0 to fall back on
  if (B > 0 && (count <= 0 || count > INT32_MAX)) {
    set_error("ptx_base_step_indexed: table of %lld views (need 1 .. 2^31 - 1)", (long long)count);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  return base_step_impl("ptx_base_step_indexed", sample, H, W, probe, h, w, positions_all, count, measured_all,
                        order, B, loss_scale, loss_per_position, amplitude, grad_sample, grad_probe, workspace,
                        ws_bytes, flags, (hipStream_t)stream);
}

ptx_status ptx_base_train_step(void* sample, int64_t H, int64_t W, void* probe, int64_t h, int64_t w,
                               const int64_t* positions_all, const float* measured_all, int64_t count,
                               const int64_t* order, int64_t B, float loss_scale, float* loss_per_position,
                               float* amplitude, void* grad_sample, void* grad_probe, const ptx_adam* adam_sample,
                               const ptx_adam* adam_probe, void* workspace, size_t ws_bytes, unsigned flags,
                               void* stream) {
  TuneScope tune_scope;
  const char* fn = "ptx_base_train_step";
  AdamArgs upd[2]{};
  const ptx_adam* cfg[2] = {adam_sample, adam_probe};
  void* params[2] = {sample, probe};
  for (int k = 0; k < 2; ++k) {
    const ptx_adam* a = cfg[k];
    if (!a) continue;
    if (!a->exp_avg || !a->exp_avg_sq || !a->hyper || !params[k]) {
      set_error("%s: null Adam state", fn);
      return PTX_ERR_INVALID_ARGUMENT;
    }
    if (!(a->beta1 >= 0.0f && a->beta1 < 1.0f && a->beta2 >= 0.0f && a->beta2 < 1.0f)) {
      set_error("%s: betas must lie in [0, 1)", fn);
      return PTX_ERR_INVALID_ARGUMENT;
    }
    upd[k] = AdamArgs{(float2*)params[k], a->exp_avg, a->exp_avg_sq, a->hyper, a->beta1, a->beta2};
  }
  if (order && B > 0 && (count <= 0 || count > INT32_MAX)) {
    set_error("%s: table of %lld views (need 1 .. 2^31 - 1)", fn, (long long)count);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  return base_step_impl(fn, sample, H, W, probe, h, w, positions_all, order ? count : B, measured_all, order, B,
                        loss_scale, loss_per_position, amplitude, grad_sample, grad_probe, workspace, ws_bytes, flags,
                        (hipStream_t)stream, &upd[0], &upd[1]);
}

}  // extern "C"
