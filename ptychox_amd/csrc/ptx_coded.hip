// Coded ptychography (ptychoX_coded): forward, adjoint and the fused
// amplitude-loss training step, full-frame H x W per view.
//
// Reference semantics: python/ptychoX/coded/ops.py:136-171 (forward under
// autograd) and the CUDA extension python/ptychoX/csrc/coded_kernel.cu:283-424:
//   O^ = F O;  R_b = exp(i 2 pi (dx_b fx + dy_b fy)), f = signed integer frequency
//   X_b = F^-1(O^ H1 R_b);  Z_b = F^-1(F(X_b C) H2);  A_b = sqrt(max(pool_ds |Z_b|^2, 0))
//   gZ = Z G_up/(ds^2 max(A_up, 1e-7));  gY = F^-1(conj(H2) F gZ)
//   grad_C = sum_b gY conj(X_b);  gX = gY conj(C);  grad_O^ = sum_b conj(H1 R_b) F(gX_b)/N
//   grad_O = N F^-1(grad_O^)
//
// The ramp is separable, R_b = ex_b[fx] ey_b[fy], with both factors computed in
// double precision from the exact turn count (the reference evaluates the
// argument 2 pi (dx fx + dy fy) in float32, coded_kernel.cu:117-122).
//
// Per view: P1 columns (O^ H1 R_b, column iFFT) | P2 rows (row iFFT -> X_b,
// * C, row FFT) | P3 columns (FFT, * H2, iFFT) | P4 rows (row iFFT -> Z_b,
// amplitude).  Backward: Q1 rows (gZ, row FFT) | Q2 columns (FFT, * conj H2,
// iFFT) | Q3 rows over a chunk of views (row iFFT -> gY, mask partials, * conj C,
// row FFT) | Q4 columns over a chunk of views (column FFT, spectrum partials),
// then fixed-order partial sums and the object-size inverse FFT.  Both shared
// gradients accumulate in registers over the chunk: no atomics anywhere, so
// the operator is deterministic.  The row passes that feed a column pass (P2,
// Q1, Q3) write their frames in the column-strip layout (passes.hpp) into a
// second frame buffer, so every column pass that reads frames (P3, Q2, Q4)
// reads its tiles from contiguous strips; the column passes write the plain
// layout, which the row passes read along their rows.
#include "buffer_io.hpp"
#include "common_kernels.hpp"
#include "passes.hpp"
#include "ptx_internal.hpp"

namespace ptx {
namespace {

constexpr float kCodedAmpEps = 1.0e-7f;  // python/ptychoX/csrc/coded_kernel.cu:227
// The MATLAB near-field backward divides by (A + 0.001) and drops the 1/ds^2
// of the exact derivative (matlab/PtychoX_updated/scripts/kernels/kernels_nearf.cuh:507-530).
constexpr float kMatlabAmpEps = 1.0e-3f;
constexpr double kTwoPi = 6.283185307179586476925286766559;

// ex[b, kx] = exp(i 2 pi dx_b fx),  ey[b, ky] = exp(i 2 pi dy_b fy)   (tables [B, W] and [B, H])
static __global__ void __launch_bounds__(256) k_ramp_tables(const float* __restrict__ shifts, int64_t B, int64_t H,
                                                            int64_t W, float2* __restrict__ ex,
                                                            float2* __restrict__ ey) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t per = H + W;
  if (i >= B * per) return;
  const int64_t b = i / per, k = i % per;
  const bool is_x = k < W;
  const int64_t n = is_x ? W : H;
  const int64_t j = is_x ? k : k - W;
  const int64_t f = 2 * j >= n ? j - n : j;
  const double turns = (double)shifts[2 * b + (is_x ? 0 : 1)] * (double)f;
  const double frac = turns - rint(turns);
  double sn, cs;
  sincos(kTwoPi * frac, &sn, &cs);
  const float2 v = make_float2((float)cs, (float)sn);
  if (is_x)
    ex[b * W + j] = v;
  else
    ey[b * H + j] = v;
}

// P1: O^ H1 R_b -> column iFFT
struct CodedSpecCols {
  const float2* spec;
  const float2* h1;
  const float2* ex;
  const float2* ey;
  float2* out;
  int64_t H, width;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const float2 rx = ex[frame * width + col];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int64_t r = t + q * T;
      const int64_t k = r * width + col;
      const float2 o = h1 ? c_mul(spec[k], h1[k]) : spec[k];  // h1 == null: H1 = 1
      v[q] = c_mul(o, c_mul(ey[frame * H + r], rx));
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

// P2: row iFFT -> X = v/N (stored), * C, row FFT -> out in the column-strip layout
template <int N, int E>
__global__ void __launch_bounds__(256)
    k_coded_x_rows(const float2* __restrict__ tb, float2* __restrict__ out, float2* __restrict__ xfwd,
                   const float2* __restrict__ mask, int64_t H, int64_t nrows, float inv_n) {
  using Cf = RowCfg<N, E>;
  constexpr int T = Cf::T, S = Cf::S;
  __shared__ float2 lds[S * lds_len<N>()];
  const int seq = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int64_t r0 = (int64_t)blockIdx.x * S;
  const int64_t row = r0 + seq;
  const bool valid = row < nrows;
  const int64_t y = row & (H - 1);  // H is a power of two
  // the work-group's rows as one buffer of each array: rows past nrows lie
  // beyond it, so their loads return zero and their stores are dropped -- the
  // loads need no branches (which would make the compiler wait for each load
  // before issuing the next)
  const uint32_t live = (uint32_t)((nrows - r0 < S ? nrows - r0 : S) * N);
  const uint32_t el = (uint32_t)(seq * N + t);
  float2 v[E];
  {
    const auto rt = buffer_rsrc(tb + r0 * N, live * 8u);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = buf_load_f2(rt, el * 8u, (uint32_t)(q * T * 8));
  }
  fft_row<N, E, +1>(v, t, &lds[seq * lds_len<N>()]);
  const auto rx = buffer_rsrc(xfwd + r0 * N, live * 8u);
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const float2 x = c_scale(v[q], inv_n);
    buf_store_f2(x, rx, el * 8u, (uint32_t)(q * T * 8));
    v[q] = c_mul(x, mask[y * N + t + q * T]);
  }
  seq_sync<row_in_wave<N>(E)>();  // the second transform rewrites the row's LDS
  fft_row<N, E, -1>(v, t, &lds[seq * lds_len<N>()]);
  if (valid) store_row_strip<N, E>(out + (row - y) * N, (uint32_t)H, (uint32_t)y, t, v);
}

// P4 (forward): row iFFT -> Z = v/N stored; ds == 1: amplitude |Z|
struct CodedZRows {
  const float2* in;
  float2* z;
  float* amp;  // null when ds > 1 (pooled separately)
  float inv_n;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = valid ? in[row * N + t + q * T] : c_zero();
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    if (!valid) return;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int64_t k = row * N + t + q * T;
      const float2 e = c_scale(v[q], inv_n);
      z[k] = e;
      if (amp) amp[k] = c_abs(e);
    }
  }
};

// RMS pooling of |Z|^2 over ds x ds cells.  mode 0: amplitude only;
// mode 1 (backward): sc = G/(ds^2 max(A, eps)); mode 2 (step): amplitude, loss
// partial per pooled row and sc = gscale (A - Y)/(ds^2 max(A, eps));
// mode 3 (MATLAB near-field backward): sc = G/(A + 0.001).
static __global__ void __launch_bounds__(256)
    k_coded_pool(const float2* __restrict__ z, int64_t B, int64_t H, int64_t W, int ds, const float* __restrict__ g,
                 float* __restrict__ amp, float* __restrict__ sc, float* __restrict__ loss_part, float gscale,
                 int mode) {
  const int64_t lh = H / ds, lw = W / ds;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= B * lh * lw) return;
  const int64_t b = i / (lh * lw), rem = i % (lh * lw);
  const int64_t ly = rem / lw, lx = rem % lw;
  float s = 0.0f;
  for (int yy = 0; yy < ds; ++yy)
    for (int xx = 0; xx < ds; ++xx) {
      const float2 v = z[(b * H + ly * ds + yy) * W + lx * ds + xx];
      s += v.x * v.x + v.y * v.y;
    }
  const float inv_cell = 1.0f / (float)(ds * ds);
  const float a = sqrtf(fmaxf(s * inv_cell, 0.0f));
  if (amp) amp[i] = a;
  if (mode == 1) {
    sc[i] = g[i] * inv_cell * fast_rcp(fmaxf(a, kCodedAmpEps));
  } else if (mode == 2) {
    const float d = a - g[i];
    sc[i] = gscale * d * inv_cell * fast_rcp(fmaxf(a, kCodedAmpEps));
    loss_part[i] = d * d;
  } else if (mode == 3) {
    sc[i] = g[i] / (a + kMatlabAmpEps);
  }
}

// Q1: gZ = Z * sc_up (ds > 1, scale map from k_coded_pool) or
// gZ = Z G/max(|Z|, eps) (ds == 1, straight from the upstream gradient); row FFT
// -> out in the column-strip layout
struct CodedGradZRows {
  const float2* z;
  const float* sc;  // ds > 1: [B, H/ds, W/ds] scale map;  ds == 1: G [B, H, W]
  float2* out;
  int64_t H;
  int ds;
  int matlab;  // ds == 1: gZ = Z G/(|Z| + 0.001) (MATLAB near-field backward)
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    // (H is a power of two)
    const int64_t b = valid ? row >> __builtin_ctzll((uint64_t)H) : 0, y = valid ? row & (H - 1) : 0;
    const int64_t lw = N / ds;
    const float* srow = sc + (b * (H / ds) + y / ds) * lw;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * T;
      if (!valid) {
        v[q] = c_zero();
      } else if (ds == 1) {
        const float2 e = z[row * N + i];
        v[q] = matlab ? c_scale(e, srow[i] / (c_abs(e) + kMatlabAmpEps))
                      : c_scale(e, srow[i] * fast_rcp(fmaxf(c_abs(e), kCodedAmpEps)));
      } else {
        v[q] = c_scale(z[row * N + i], srow[i / ds]);
      }
    }
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    if (!valid) return;
    const int64_t y = row & (H - 1);  // H is a power of two
    float2* d = out + (row - y) * N;  // the frame, column-strip layout
    store_row_strip<N, E>(d, (uint32_t)H, (uint32_t)y, t, v);
  }
};

// P4 + Q1 fused for the ds == 1 step: row iFFT -> Z, A = |Z|, loss partial per
// row, gZ = Z gscale (A - Y)/max(A, eps), row FFT -> out in the column-strip layout
template <int N, int E>
__global__ void __launch_bounds__(256)
    k_coded_step_rows(const float2* __restrict__ tb, float2* __restrict__ out, const float* __restrict__ measured,
                      float* __restrict__ amp_out, float* __restrict__ loss_part, int64_t H, int64_t nrows,
                      float gscale, float inv_n) {
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
  float2 v[E];
  float m[E];
  {
    const auto rt = buffer_rsrc(tb + r0 * N, live * 8u);
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
    const AbsInv ai = c_abs_inv(e, 1.0f / kCodedAmpEps);
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
  // row sum: butterfly within each wave, then the row's waves in order (a row
  // is T >= 64 threads, whole waves; shorter rows share a wave, T-lane groups)
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
  if (valid) {
    const int64_t y = row & (H - 1);  // H is a power of two
    store_row_strip<N, E>(out + (row - y) * N, (uint32_t)H, (uint32_t)y, t, v);
  }
}

// Q3: per (row group, chunk of views): row iFFT -> gY = v/N; mask partial += gY conj(X_b);
// gX = gY conj(C); row FFT -> out in the column-strip layout
template <int N, int E>
__global__ void __launch_bounds__(256)
    k_coded_y_rows(const float2* __restrict__ tb, float2* __restrict__ out, const float2* __restrict__ xfwd,
                   const float2* __restrict__ mask, int64_t H, int64_t B, int chunk, int ngroups, float inv_n,
                   float2* __restrict__ gc_part) {
  // Views of a chunk are pipelined: the next view's row (T) is loaded while
  // this view's transforms run, and its X row once this view's X is consumed;
  // buffer addressing (one VGPR per lane), LDS-only barriers and LDS twiddles
  // keep those loads in flight.
  using Cf = RowCfg<N, E>;
  constexpr int T = Cf::T, S = Cf::S;
  __shared__ float2 lds[S * lds_len<N>()];
  __shared__ float2 s_tw[N];
  load_lds_twiddles<N>(s_tw);
  const int seq = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int group = blockIdx.x % ngroups;
  const int ch = blockIdx.x / ngroups;
  const int64_t y = (int64_t)group * S + seq;
  const bool yv = y < H;
  const int64_t b0 = (int64_t)ch * chunk;
  const int64_t b1 = b0 + chunk < B ? b0 + chunk : B;
  const uint32_t fbytes = (uint32_t)(H * N * 8);
  const uint32_t lane = (uint32_t)(((yv ? y : 0) * N + t) * 8);  // rows past H read row 0, never stored
  float2 cm[E], acc[E], nv[E], nx[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    cm[q] = yv ? mask[y * N + t + q * T] : c_zero();
    acc[q] = c_zero();
  }
  auto fetch_t = [&](int64_t b, float2 (&dv)[E]) {
    const auto r = buffer_rsrc(tb + b * H * N, fbytes);
#pragma unroll
    for (int q = 0; q < E; ++q) dv[q] = buf_load_f2(r, lane, (uint32_t)(q * T * 8));
  };
  auto fetch_x = [&](int64_t b, float2 (&dx)[E]) {
    const auto r = buffer_rsrc(xfwd + b * H * N, fbytes);
#pragma unroll
    for (int q = 0; q < E; ++q) dx[q] = buf_load_f2(r, lane, (uint32_t)(q * T * 8));
  };
  if (b0 < b1) {
    fetch_t(b0, nv);
    fetch_x(b0, nx);
  }
  for (int64_t b = b0; b < b1; ++b) {
    float2 v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = nv[q];
    if (b + 1 < b1) fetch_t(b + 1, nv);
    int tt = t;
    asm volatile("" : "+v"(tt));  // keep the exchange addresses per view, not hoisted into registers
    float2* sl = &lds[seq * lds_len<N>()];
    fft_row_tw<N, E, +1>(v, tt, sl, s_tw);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const float2 gy = c_scale(v[q], inv_n);
      acc[q] = c_add(acc[q], c_mulc(gy, nx[q]));
      v[q] = c_mulc(gy, cm[q]);
    }
    if (b + 1 < b1) fetch_x(b + 1, nx);
    seq_sync<row_in_wave<N>(E)>();  // the second transform rewrites the row's LDS
    fft_row_tw<N, E, -1>(v, tt, sl, s_tw);
    if (yv) store_row_strip<N, E>(out + b * H * N, (uint32_t)H, (uint32_t)y, t, v);  // column-strip layout
    seq_sync<row_in_wave<N>(E)>();  // next view's first LDS write vs this view's last read
  }
  if (yv) {
#pragma unroll
    for (int q = 0; q < E; ++q) gc_part[((int64_t)ch * H + y) * N + t + q * T] = acc[q];
  }
}

// Q4: per (column tile, chunk of views; frames in the column-strip layout):
// column FFT; spectrum partial
// += conj(H1 R_b) v / N, with the view-independent conj(H1)/N applied once
// after the view loop.  About one work-group per CU (ColChunkQuery), each on
// one column tile and a chunk of consecutive views, software-pipelined like
// k_transfer_cols: two register sets alternate, the next view's tile and ramp
// values in flight during this view's transform; a view past the chunk gets
// zero-length buffers (loads issued, no memory touched), so every path issues
// the same memory instructions and the wait counts keep the prefetch in
// flight; the first view is peeled; the exchanges alternate two LDS regions
// (with an odd number of exchanges per transform, every other view starts on
// the second region).
template <int N, int E, int C>
__global__ void __launch_bounds__((ColCfg<N, E>::T * C))
    k_coded_spec_cols(const float2* __restrict__ tb, const float2* __restrict__ h1, const float2* __restrict__ ex,
                      const float2* __restrict__ ey, int64_t B, int64_t width, int64_t tiles, int chunk, float inv_n,
                      float2* __restrict__ gs_part) {
  using Sh = PipeColsShape<N, E, C, false>;
  constexpr int T = Sh::T;
  constexpr bool ODD = Sh::EXCHANGES % 2 == 1;
  __shared__ float2 lds0[C * col_stride<N, C>()];
  __shared__ float2 lds1[C * col_stride<N, C>()];
  __shared__ float2 s_tw[LoopTwiddles<N, E, false>::TABLE];
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  LoopTwiddles<N, E, false> tw;  // the accumulator fills the register budget: LDS table
  tw.init(t, s_tw);
  const TileIdx ti = split_tile(xcd_paired_block(blockIdx.x, gridDim.x), tiles);
  const int64_t tile = ti.tile;
  const unsigned ch = (unsigned)ti.frame;  // chunk
  const int64_t col = tile * C + c;
  const unsigned b0 = ch * (unsigned)chunk;                            // < B
  const unsigned b1 = b0 + chunk < (unsigned)B ? b0 + chunk : (unsigned)B;
  const uint32_t fbytes = (uint32_t)(N * width * 8);
  const uint32_t col0 = (uint32_t)(tile * C * 8);
  // the frames are in the column-strip layout: the tile is a contiguous block
  const uint32_t lg = strip_lg(width);
  const uint32_t sbase = strip_off(0u, (uint32_t)(tile * C), (uint32_t)N, lg);
  const uint32_t lane = (strip_off((uint32_t)t, (uint32_t)col, (uint32_t)N, lg) - sbase) * 8u;
  float2 acc[E];
#pragma unroll
  for (int q = 0; q < E; ++q) acc[q] = c_zero();
  auto fetch = [&](unsigned b, float2 (&dv)[E], float2 (&dy)[E], float2& drx) {
    const bool live = b < b1;
    const unsigned bb = live ? b : b0;
    drx = buf_load_f2(buffer_rsrc(ex + (int64_t)bb * width, live ? (uint32_t)(width * 8) : 0u), (uint32_t)c * 8u,
                      col0);
    const auto ry = buffer_rsrc(ey + (int64_t)bb * N, live ? (uint32_t)(N * 8) : 0u);
#pragma unroll
    for (int q = 0; q < E; ++q) dy[q] = buf_load_f2(ry, (uint32_t)t * 8u, (uint32_t)(q * T * 8));
    const auto r = buffer_rsrc(tb + (int64_t)bb * N * width, live ? fbytes : 0u);
#pragma unroll
    for (int q = 0; q < E; ++q) dv[q] = buf_load_f2(r, lane, (sbase + ((uint32_t)(q * T) << lg)) * 8u);
  };
  auto process = [&](float2 (&v)[E], const float2 (&ry)[E], float2 rx, bool second) {
    // the exchange addresses depend only on (t, c): recomputed per view from the
    // thread index instead of being hoisted out of the loop into registers
    unsigned tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int tt = (int)(tid / C), cc = (int)(tid % C);
    float2* l0 = &lds0[cc * col_stride<N, C>()];
    float2* l1 = &lds1[cc * col_stride<N, C>()];
    if (second)
      tw.template stages<-1>(v, tt, l1, l0);
    else
      tw.template stages<-1>(v, tt, l0, l1);
#pragma unroll
    for (int q = 0; q < E; ++q) acc[q] = c_add(acc[q], c_mulc(v[q], c_mul(ry[q], rx)));  // v conj(R_b)
  };
  float2 va[E], vb[E], ya[E], yb[E], xa, xb;
  unsigned b = b0;
  fetch(b, va, ya, xa);
  fetch(b + 1, vb, yb, xb);  // in flight during this view's transform
  process(va, ya, xa, false);
  ++b;
  while (b < b1) {
    fetch(b + 1, va, ya, xa);
    process(vb, yb, xb, ODD);
    ++b;
    if (b >= b1) break;
    fetch(b + 1, vb, yb, xb);
    process(va, ya, xa, false);
    ++b;
  }
  float2* d = gs_part + (int64_t)ch * N * width;  // slice base; the lane adds its column
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const float2 hc = h1 ? h1[(t + q * T) * (int)width + (int)col] : make_float2(1.0f, 0.0f);  // null: H1 = 1
    d[(t + q * T) * (int)width + (int)col] = c_scale(c_mulc(acc[q], hc), inv_n);
  }
}

// P1 as a column-tile walker: a work-group owns one column tile and a chunk of
// consecutive views; the tile's O^ H1 is loaded once, and per view only the
// ramp factors are read (next view's in flight during this view's transform):
// v = O^ H1 ey_b ex_b, column iFFT, store.  Zero-length buffers past the chunk,
// the first view peeled and two-region LDS exchanges as in k_coded_spec_cols.
template <int N, int E, int C>
__global__ void __launch_bounds__((ColCfg<N, E>::T * C))
    k_coded_spec_fwd_cols(const float2* __restrict__ spec, const float2* __restrict__ h1,
                          const float2* __restrict__ ex, const float2* __restrict__ ey, int64_t B, int64_t width,
                          int64_t tiles, int chunk, float2* __restrict__ out) {
  using Sh = PipeColsShape<N, E, C>;
  constexpr int T = Sh::T;
  constexpr bool ODD = Sh::EXCHANGES % 2 == 1;
  __shared__ float2 lds0[C * col_stride<N, C>()];
  __shared__ float2 lds1[C * col_stride<N, C>()];
  __shared__ float2 s_tw[LoopTwiddles<N, E>::TABLE];
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  LoopTwiddles<N, E> tw;  // thread t holds the same rows of every view
  tw.init(t, s_tw);
  const TileIdx ti = split_tile(xcd_paired_block(blockIdx.x, gridDim.x), tiles);
  const int64_t col = ti.tile * C + c;
  const unsigned b0 = (unsigned)ti.frame * (unsigned)chunk;  // < B
  const unsigned b1 = b0 + chunk < (unsigned)B ? b0 + chunk : (unsigned)B;
  const uint32_t fbytes = (uint32_t)(N * width * 8);
  const uint32_t lane = (uint32_t)((t * width + c) * 8);
  const uint32_t col0 = (uint32_t)(ti.tile * C * 8);
  const uint32_t rstep = (uint32_t)(T * width * 8);
  float2 base[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int64_t k = (int64_t)(t + q * T) * width + col;
    base[q] = h1 ? c_mul(spec[k], h1[k]) : spec[k];  // h1 == null: H1 = 1
  }
  auto fetch = [&](unsigned b, float2 (&dy)[E], float2& drx) {
    const bool live = b < b1;
    const unsigned bb = live ? b : b0;
    drx = buf_load_f2(buffer_rsrc(ex + (int64_t)bb * width, live ? (uint32_t)(width * 8) : 0u), (uint32_t)c * 8u,
                      col0);
    const auto ry = buffer_rsrc(ey + (int64_t)bb * N, live ? (uint32_t)(N * 8) : 0u);
#pragma unroll
    for (int q = 0; q < E; ++q) dy[q] = buf_load_f2(ry, (uint32_t)t * 8u, (uint32_t)(q * T * 8));
  };
  auto process = [&](unsigned b, const float2 (&ry)[E], float2 rx, bool second) {
    unsigned tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int tt = (int)(tid / C), cc = (int)(tid % C);
    float2* l0 = &lds0[cc * col_stride<N, C>()];
    float2* l1 = &lds1[cc * col_stride<N, C>()];
    float2 v[E];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = c_mul(base[q], c_mul(ry[q], rx));
    if (second)
      tw.template stages<+1>(v, tt, l1, l0);
    else
      tw.template stages<+1>(v, tt, l0, l1);
    const auto r = buffer_rsrc(out + (int64_t)b * N * width, fbytes);
#pragma unroll
    for (int q = 0; q < E; ++q) buf_store_f2(v[q], r, lane, q * rstep + col0);
  };
  float2 ya[E], yb[E], xa, xb;
  unsigned b = b0;
  fetch(b, ya, xa);
  fetch(b + 1, yb, xb);
  process(b, ya, xa, false);
  ++b;
  while (b < b1) {
    fetch(b + 1, ya, xa);
    process(b, yb, xb, ODD);
    ++b;
    if (b >= b1) break;
    fetch(b + 1, yb, xb);
    process(b, ya, xa, false);
    ++b;
  }
}

// ---------------------------------------------------------------- host launchers

// P1: column tiles x chunks of views, about 256 work-groups
template <int N, int E>
struct SpecColsRun {
  template <int C>
  static void launch(const CodedSpecCols& op, int64_t frames, hipStream_t s) {
    if (frames <= 0) return;
    const int64_t tiles = op.width / C;
    int64_t nchunks = (tuned_groups(PASS_COL, N, 256) + tiles - 1) / tiles;
    if (nchunks > frames) nchunks = frames;
    const int64_t chunk = (frames + nchunks - 1) / nchunks;
    nchunks = (frames + chunk - 1) / chunk;
    k_coded_spec_fwd_cols<N, E, C><<<dim3((unsigned)(tiles * nchunks)), dim3(ColCfg<N, E>::T * C), 0, s>>>(
        op.spec, op.h1, op.ex, op.ey, frames, op.width, tiles, (int)chunk, op.out);
  }
  static ptx_status run(const CodedSpecCols& op, int64_t frames, hipStream_t s) {
    constexpr int CM = ColCfg<N, E>::CMAX;
    bool done = false;
    if constexpr (CM > 8) {
      if (op.width < CM) {  // frames narrower than CMAX columns: tiles of 8
        launch<8>(op, frames, s);
        done = true;
      }
    }
    if (!done) launch<CM>(op, frames, s);
    return check_launch("coded spectrum cols", s);
  }
};
template <int N, int E>
struct XRowsRun {
  static ptx_status run(const float2* tb, float2* out, float2* xfwd, const float2* mask, int64_t H, int64_t nrows,
                        float inv_n, hipStream_t s) {
    const int64_t blocks = (nrows + RowCfg<N, E>::S - 1) / RowCfg<N, E>::S;
    k_coded_x_rows<N, E><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(tb, out, xfwd, mask, H, nrows, inv_n);
    return check_launch("coded x rows", s);
  }
};
// src (column-strip layout) -> tb (plain layout)
template <int N, int E>
struct TransferRun {
  static ptx_status run(float2* tb, const float2* src, const float2* tr, int64_t frames, int64_t width, bool conj,
                        hipStream_t s) {
    launch_transfer_cols<N, E>(tb, tr, frames, width, conj, s, src);
    return check_launch(conj ? "coded transfer* cols" : "coded transfer cols", s);
  }
};
template <int N, int E>
struct ZRowsRun {
  static ptx_status run(const CodedZRows& op, int64_t nrows, hipStream_t s) {
    launch_rows<N, E, +1>(op, nrows, s);
    return check_launch("coded z rows", s);
  }
};
template <int N, int E>
struct GradZRowsRun {
  static ptx_status run(const CodedGradZRows& op, int64_t nrows, hipStream_t s) {
    launch_rows<N, E, -1>(op, nrows, s);
    return check_launch("coded grad-z rows", s);
  }
};
template <int N, int E>
struct StepRowsRun {
  static ptx_status run(const float2* tb, float2* out, const float* measured, float* amp, float* loss_part, int64_t H,
                        int64_t nrows, float gscale, float inv_n, hipStream_t s) {
    const int64_t blocks = (nrows + RowCfg<N, E>::S - 1) / RowCfg<N, E>::S;
    k_coded_step_rows<N, E><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(tb, out, measured, amp, loss_part, H, nrows,
                                                                      gscale, inv_n);
    return check_launch("coded step rows", s);
  }
};

struct Chunks {
  int row_groups, row_chunk, row_nchunks;  // Q3
  int col_tiles, col_c, col_chunk, col_nchunks;  // Q4
};

template <int N, int E>
struct RowChunkQuery {
  static ptx_status run(int64_t H, int64_t B, Chunks* c) {
    c->row_groups = (int)((H + RowCfg<N, E>::S - 1) / RowCfg<N, E>::S);
    int64_t want = (tuned_groups(PASS_VIEW_ROWS, N, 2048) + c->row_groups - 1) / c->row_groups;
    // (an empty batch plans one chunk of one view: nothing is launched for it)
    want = want < 1 ? 1 : (B > 0 && want > B ? B : want);
    c->row_chunk = B > 0 ? (int)((B + want - 1) / want) : 1;
    c->row_nchunks = B > 0 ? (int)((B + c->row_chunk - 1) / c->row_chunk) : 1;
    return PTX_OK;
  }
};
// Q4 chunks: about 256 work-groups by default (one per CU of the target, one
// resident work-group each; tunable), so each walks a long chunk of views and
// few partial spectra are written and summed.
template <int N, int E>
struct ColChunkQuery {
  static ptx_status run(int64_t W, int64_t B, Chunks* c) {
    c->col_c = W >= ColCfg<N, E>::CMAX ? ColCfg<N, E>::CMAX : 8;
    c->col_tiles = (int)(W / c->col_c);
    int64_t want = (tuned_groups(PASS_COL_LOOP, N, 256) + c->col_tiles - 1) / c->col_tiles;
    want = want < 1 ? 1 : (B > 0 && want > B ? B : want);
    c->col_chunk = B > 0 ? (int)((B + want - 1) / want) : 1;
    c->col_nchunks = B > 0 ? (int)((B + c->col_chunk - 1) / c->col_chunk) : 1;
    return PTX_OK;
  }
};
template <int N, int E>
struct YRowsRun {
  static ptx_status run(const float2* tb, float2* out, const float2* xfwd, const float2* mask, int64_t H, int64_t B,
                        const Chunks& c, float inv_n, float2* gc_part, hipStream_t s) {
    k_coded_y_rows<N, E><<<dim3((unsigned)(c.row_groups * c.row_nchunks)), dim3(256), 0, s>>>(
        tb, out, xfwd, mask, H, B, c.row_chunk, c.row_groups, inv_n, gc_part);
    return check_launch("coded y rows", s);
  }
};
template <int N, int E>
struct SpecGradColsRun {
  static ptx_status run(const float2* tb, const float2* h1, const float2* ex, const float2* ey, int64_t B, int64_t W,
                        const Chunks& c, float inv_n, float2* gs_part, hipStream_t s) {
    const dim3 grid((unsigned)(c.col_tiles * c.col_nchunks));
    constexpr int CM = ColCfg<N, E>::CMAX;
    if constexpr (CM > 8) {
      if (c.col_c != CM) {  // frames narrower than CMAX columns: tiles of 8
        k_coded_spec_cols<N, E, 8><<<grid, dim3(ColCfg<N, E>::T * 8), 0, s>>>(tb, h1, ex, ey, B, W, c.col_tiles,
                                                                         c.col_chunk, inv_n, gs_part);
        return check_launch("coded spectrum-grad cols", s);
      }
    }
    k_coded_spec_cols<N, E, CM><<<grid, dim3(ColCfg<N, E>::T * CM), 0, s>>>(tb, h1, ex, ey, B, W, c.col_tiles,
                                                                       c.col_chunk, inv_n, gs_part);
    return check_launch("coded spectrum-grad cols", s);
  }
};
template <int N, int E>
struct RowsOOPRun {
  static ptx_status run(const float2* in, float2* out, int64_t nrows, int dir, hipStream_t s) {
    RowsOutOfPlace op{in, out, 1.0f};
    if (dir < 0)
      launch_rows<N, E, -1>(op, nrows, s);
    else
      launch_rows<N, E, +1>(op, nrows, s);
    return check_launch(dir < 0 ? "fft rows" : "ifft rows", s);
  }
};
// columns of one [N, width] frame from `in` into `out` (= or +=)
struct ColsInto {
  const float2* in;
  float2* out;
  int64_t width;
  int accumulate;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = in[(t + q * T) * (int)width + (int)col];
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      float2* d = out + (t + q * T) * (int)width + (int)col;
      *d = accumulate ? c_add(*d, v[q]) : v[q];
    }
  }
};
template <int N, int E>
struct ColsIntoRun {
  static ptx_status run(const ColsInto& op, int dir, hipStream_t s) {
    if (dir < 0)
      launch_cols<N, E, -1>(op, 1, op.width, s);
    else
      launch_cols<N, E, +1>(op, 1, op.width, s);
    return check_launch(dir < 0 ? "fft cols" : "ifft cols", s);
  }
};

ptx_status fft2_frame(const float2* in, float2* out, int64_t H, int64_t W, int dir, int accumulate, float2* scratch,
                      hipStream_t s) {
  // rows out of place into scratch, columns from scratch into out
  ptx_status st = dispatch_ne<RowsOOPRun>(PASS_ROW, W, in, scratch, H, dir, s);
  if (st != PTX_OK) return st;
  ColsInto c{scratch, out, W, accumulate};
  return dispatch_ne<ColsIntoRun>(PASS_COL, H, c, dir, s);
}

struct CodedWorkspace {
  float2* spec;     // [H, W] (step)
  float2* ex;       // [B, W]
  float2* ey;       // [B, H]
  float2* xfwd;     // [B, H, W] (step)
  float2* tb;       // [B, H, W]
  float2* strip;    // [B, H, W] in the column-strip layout (inputs of the column passes)
  float2* gc_part;  // [row chunks, H, W]
  float2* gs_part;  // [col chunks, H, W]
  float2* gspec;    // [H, W]
  float* sc;        // [B, H/ds, W/ds] (ds > 1)
  float* loss_part; // step: [B * H] (ds == 1) or [B * H/ds * W/ds]
  size_t bytes;
};

CodedWorkspace carve(void* base, int64_t H, int64_t W, int64_t B, int64_t ds, bool step, const Chunks& c) {
  Carver cv{(char*)base, 0};
  CodedWorkspace ws{};
  ws.spec = step ? cv.take<float2>((size_t)(H * W)) : nullptr;
  ws.ex = cv.take<float2>((size_t)(B * W));
  ws.ey = cv.take<float2>((size_t)(B * H));
  ws.xfwd = step ? cv.take<float2>((size_t)(B * H * W)) : nullptr;
  ws.tb = cv.take<float2>((size_t)(B * H * W));
  ws.strip = cv.take<float2>((size_t)(B * H * W));
  ws.gc_part = cv.take<float2>((size_t)(c.row_nchunks * H * W));
  ws.gs_part = cv.take<float2>((size_t)(c.col_nchunks * H * W));
  ws.gspec = cv.take<float2>((size_t)(H * W));
  const int64_t low = B * (H / ds) * (W / ds);
  ws.sc = ds > 1 ? cv.take<float>((size_t)low) : nullptr;
  ws.loss_part = step ? cv.take<float>((size_t)(ds > 1 ? low : B * H)) : nullptr;
  ws.bytes = cv.used;
  return ws;
}

ptx_status check(const char* fn, int64_t H, int64_t W, int64_t B, int64_t ds) {
  if (B < 0) {
    set_error("%s: negative batch", fn);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (!size_ok(H) || !size_ok(W)) {
    set_error("%s: frame %lld x %lld unsupported (sides must be powers of two in [%d, %d])", fn, (long long)H,
              (long long)W, PTX_MIN_N, PTX_MAX_N);
    return PTX_ERR_UNSUPPORTED;
  }
  if (ds < 1 || H % ds != 0 || W % ds != 0) {
    set_error("%s: object_field shape must be divisible by downsample (%lld)", fn, (long long)ds);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  return PTX_OK;
}

Chunks plan(int64_t H, int64_t W, int64_t B) {
  Chunks c{};
  dispatch_ne<RowChunkQuery>(PASS_VIEW_ROWS, W, H, B, &c);
  dispatch_ne<ColChunkQuery>(PASS_COL_LOOP, H, W, B, &c);
  return c;
}

ptx_status ramps(const float* shifts, int64_t B, int64_t H, int64_t W, const CodedWorkspace& ws, hipStream_t s) {
  const int64_t n = B * (H + W);
  k_ramp_tables<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(shifts, B, H, W, ws.ex, ws.ey);
  return check_launch("coded ramp tables", s);
}

// P1..P3 of the forward: spectrum -> X (stored) -> tb holds F_cols^-1 (F(X C) H2) row-transformed
ptx_status forward_to_t(const float2* spec, const float2* mask, const float2* t1, const float2* t2, int64_t H,
                        int64_t W, int64_t B, const CodedWorkspace& ws, float2* xfwd, hipStream_t s) {
  const float inv_n = 1.0f / (float)(H * W);
  CodedSpecCols p1{spec, t1, ws.ex, ws.ey, ws.tb, H, W};
  ptx_status st = dispatch_ne<SpecColsRun>(PASS_COL, H, p1, B, s);
  if (st != PTX_OK) return st;
  if ((st = dispatch_ne<XRowsRun>(PASS_ROW, W, ws.tb, ws.strip, xfwd, mask, H, B * H, inv_n, s)) != PTX_OK) return st;
  return dispatch_ne<TransferRun>(PASS_COL, H, ws.tb, ws.strip, t2, B, W, false, s);
}

// Q2..end: strip holds the row FFT of gZ (column-strip layout); produces
// grad_object and grad_coded.
// matlab: both gradients conjugated, as the MATLAB near-field backward returns
// them -- grad_coded = conj(sum_b gY conj(X_b)) and
// grad_object = conj(N ifft2(S)) = fft2(conj(S)) for the spectrum gradient S
// (matlab/PtychoX_updated/scripts/kernels/kernels_nearf.cuh:186-240,394-436 and
// scripts/backward/bwd_adPtychoX_nearf.cu, whose last transform is forward).
ptx_status backward_from_t(const float2* xfwd, const float2* mask, const float2* t1, const float2* t2, int64_t H,
                           int64_t W, int64_t B, const Chunks& c, const CodedWorkspace& ws, float2* gobj,
                           float2* gcoded, int accumulate, hipStream_t s, bool matlab = false) {
  const float inv_n = 1.0f / (float)(H * W);
  ptx_status st = dispatch_ne<TransferRun>(PASS_COL, H, ws.tb, ws.strip, t2, B, W, true, s);
  if (st != PTX_OK) return st;
  if ((st = dispatch_ne<YRowsRun>(PASS_VIEW_ROWS, W, ws.tb, ws.strip, xfwd, mask, H, B, c, inv_n, ws.gc_part, s)) !=
      PTX_OK)
    return st;
  if ((st = dispatch_ne<SpecGradColsRun>(PASS_COL_LOOP, H, ws.strip, t1, ws.ex, ws.ey, B, W, c, inv_n, ws.gs_part,
                                         s)) != PTX_OK)
    return st;
  sum_parts2(ws.gc_part, c.row_nchunks, H * W, gcoded, accumulate, s, matlab ? 1 : 0);
  if ((st = check_launch("mask partial sum", s)) != PTX_OK) return st;
  sum_parts2(ws.gs_part, c.col_nchunks, H * W, ws.gspec, 0, s, matlab ? 1 : 0);
  if ((st = check_launch("spectrum partial sum", s)) != PTX_OK) return st;
  // grad_O = N ifft2(grad_O^) = unnormalised inverse DFT (MATLAB: the forward
  // DFT of the conjugate); rows in place, columns into the output
  return fft2_frame(ws.gspec, gobj, H, W, matlab ? -1 : +1, accumulate, ws.gspec, s);
}

}  // namespace
}  // namespace ptx

using namespace ptx;

extern "C" {

ptx_status ptx_coded_forward(const void* object_field, int64_t H, int64_t W, const void* coded_surface,
                             const void* transfer_1, const void* transfer_2, const float* shifts_xy, int64_t B,
                             int64_t downsample, void* object_spectrum, void* x_forward, void* sensor_field,
                             float* amplitude, void* workspace, size_t ws_bytes, void* stream) {
  TuneScope tune_scope;
  ptx_status st = check("ptx_coded_forward", H, W, B, downsample);
  if (st != PTX_OK) return st;
  if (!object_field || !object_spectrum || (B > 0 && (!coded_surface || !transfer_2 || !shifts_xy ||
                                                      !x_forward || !sensor_field || !amplitude))) {
    set_error("ptx_coded_forward: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  const Chunks c = plan(H, W, B);
  const CodedWorkspace ws = carve(workspace, H, W, B, downsample, false, c);
  if (B > 0 && (!workspace || ws_bytes < ws.bytes)) {
    set_error("ptx_coded_forward: workspace of %zu bytes needed, %zu given", ws.bytes, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  timing_begin(s);
  float2* spec = (float2*)object_spectrum;
  if ((st = fft2_frame((const float2*)object_field, spec, H, W, -1, 0, spec, s)) != PTX_OK) return st;
  if (B == 0) return PTX_OK;
  if ((st = ramps(shifts_xy, B, H, W, ws, s)) != PTX_OK) return st;
  if ((st = forward_to_t(spec, (const float2*)coded_surface, (const float2*)transfer_1, (const float2*)transfer_2, H,
                         W, B, ws, (float2*)x_forward, s)) != PTX_OK)
    return st;
  CodedZRows p4{ws.tb, (float2*)sensor_field, downsample == 1 ? amplitude : nullptr, 1.0f / (float)(H * W)};
  if ((st = dispatch_ne<ZRowsRun>(PASS_ROW, W, p4, B * H, s)) != PTX_OK) return st;
  if (downsample > 1) {
    const int64_t n = B * (H / downsample) * (W / downsample);
    k_coded_pool<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(
        (const float2*)sensor_field, B, H, W, (int)downsample, nullptr, amplitude, nullptr, nullptr, 0.0f, 0);
    if ((st = check_launch("coded pool", s)) != PTX_OK) return st;
  }
  return PTX_OK;
}

size_t ptx_coded_workspace(int64_t H, int64_t W, int64_t B, int64_t downsample, int step) {
  TuneScope tune_scope;
  if (check("ptx_coded_workspace", H, W, B, downsample) != PTX_OK) return 0;
  return carve(nullptr, H, W, B, downsample, step != 0, plan(H, W, B)).bytes;
}

ptx_status ptx_coded_backward(const float* grad_output, const void* object_spectrum, const void* x_forward,
                              const void* sensor_field, const void* coded_surface, const void* transfer_1,
                              const void* transfer_2, const float* shifts_xy, int64_t B, int64_t H, int64_t W,
                              int64_t downsample, void* grad_object, void* grad_coded, void* workspace,
                              size_t ws_bytes, unsigned flags, void* stream) {
  TuneScope tune_scope;
  ptx_status st = check("ptx_coded_backward", H, W, B, downsample);
  if (st != PTX_OK) return st;
  if (!grad_object || !grad_coded) {
    set_error("ptx_coded_backward: null gradient output");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  timing_begin(s);
  const int acc = (flags & PTX_FLAG_ACCUMULATE) ? 1 : 0;
  if (B == 0) {
    if (!acc) {
      fill_zero2((float2*)grad_object, H * W, s);
      fill_zero2((float2*)grad_coded, H * W, s);
    }
    return check_launch("zero fill", s);
  }
  if (!grad_output || !object_spectrum || !x_forward || !sensor_field || !coded_surface || !transfer_2 ||
      !shifts_xy) {
    set_error("ptx_coded_backward: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  const Chunks c = plan(H, W, B);
  const CodedWorkspace ws = carve(workspace, H, W, B, downsample, false, c);
  if (!workspace || ws_bytes < ws.bytes) {
    set_error("ptx_coded_backward: workspace of %zu bytes needed, %zu given", ws.bytes, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  const bool matlab = (flags & PTX_FLAG_MATLAB_NEARF) != 0;
  if ((st = ramps(shifts_xy, B, H, W, ws, s)) != PTX_OK) return st;
  const float* scmap = grad_output;
  if (downsample > 1) {
    // scale map G/(ds^2 max(A, eps)) from the recomputed amplitude (coded_kernel.cu:370-377),
    // or G/(A + 0.001) for the MATLAB near-field backward
    const int64_t n = B * (H / downsample) * (W / downsample);
    k_coded_pool<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(
        (const float2*)sensor_field, B, H, W, (int)downsample, grad_output, nullptr, ws.sc, nullptr, 0.0f,
        matlab ? 3 : 1);
    if ((st = check_launch("coded pool", s)) != PTX_OK) return st;
    scmap = ws.sc;
  }
  CodedGradZRows q1{(const float2*)sensor_field, scmap, ws.strip, H, (int)downsample, matlab ? 1 : 0};
  if ((st = dispatch_ne<GradZRowsRun>(PASS_ROW, W, q1, B * H, s)) != PTX_OK) return st;
  return backward_from_t((const float2*)x_forward, (const float2*)coded_surface, (const float2*)transfer_1,
                         (const float2*)transfer_2, H, W, B, c, ws, (float2*)grad_object, (float2*)grad_coded, acc, s,
                         matlab);
}

ptx_status ptx_coded_step(const void* object_field, int64_t H, int64_t W, const void* coded_surface,
                          const void* transfer_1, const void* transfer_2, const float* shifts_xy, int64_t B,
                          int64_t downsample, const float* measured, float loss_scale, float* loss_per_position,
                          float* amplitude, void* grad_object, void* grad_coded, void* workspace, size_t ws_bytes,
                          unsigned flags, void* stream) {
  TuneScope tune_scope;
  ptx_status st = check("ptx_coded_step", H, W, B, downsample);
  if (st != PTX_OK) return st;
  if (!grad_object || !grad_coded) {
    set_error("ptx_coded_step: null gradient output");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  timing_begin(s);
  const int acc = (flags & PTX_FLAG_ACCUMULATE) ? 1 : 0;
  if (B == 0) {
    if (!acc) {
      fill_zero2((float2*)grad_object, H * W, s);
      fill_zero2((float2*)grad_coded, H * W, s);
    }
    return check_launch("zero fill", s);
  }
  if (!object_field || !coded_surface || !transfer_2 || !shifts_xy || !measured ||
      !loss_per_position) {
    set_error("ptx_coded_step: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  const Chunks c = plan(H, W, B);
  const CodedWorkspace ws = carve(workspace, H, W, B, downsample, true, c);
  if (!workspace || ws_bytes < ws.bytes) {
    set_error("ptx_coded_step: workspace of %zu bytes needed, %zu given", ws.bytes, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  const int64_t ds = downsample;
  const int64_t lh = H / ds, lw = W / ds;
  if ((st = fft2_frame((const float2*)object_field, ws.spec, H, W, -1, 0, ws.spec, s)) != PTX_OK) return st;
  if ((st = ramps(shifts_xy, B, H, W, ws, s)) != PTX_OK) return st;
  if ((st = forward_to_t(ws.spec, (const float2*)coded_surface, (const float2*)transfer_1, (const float2*)transfer_2,
                         H, W, B, ws, ws.xfwd, s)) != PTX_OK)
    return st;
  const float inv_n = 1.0f / (float)(H * W);
  if (ds == 1) {
    if ((st = dispatch_ne<StepRowsRun>(PASS_LOSS_ROWS, W, ws.tb, ws.strip, measured, amplitude, ws.loss_part, H, B * H,
                                       2.0f * loss_scale, inv_n, s)) != PTX_OK)
      return st;
    loss_rows(ws.loss_part, H, B, inv_n, loss_per_position, s);
    if ((st = check_launch("loss rows", s)) != PTX_OK) return st;
  } else {
    // Z in place in T, then pooled amplitude / loss / scale map, then gZ row FFT
    CodedZRows p4{ws.tb, ws.tb, nullptr, inv_n};
    if ((st = dispatch_ne<ZRowsRun>(PASS_ROW, W, p4, B * H, s)) != PTX_OK) return st;
    const int64_t n = B * lh * lw;
    k_coded_pool<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(ws.tb, B, H, W, (int)ds, measured, amplitude,
                                                                        ws.sc, ws.loss_part, 2.0f * loss_scale, 2);
    if ((st = check_launch("coded pool", s)) != PTX_OK) return st;
    loss_rows(ws.loss_part, lh * lw, B, 1.0f / (float)(lh * lw), loss_per_position, s);
    if ((st = check_launch("loss rows", s)) != PTX_OK) return st;
    CodedGradZRows q1{ws.tb, ws.sc, ws.strip, H, (int)ds, 0};
    if ((st = dispatch_ne<GradZRowsRun>(PASS_ROW, W, q1, B * H, s)) != PTX_OK) return st;
  }
  return backward_from_t(ws.xfwd, (const float2*)coded_surface, (const float2*)transfer_1, (const float2*)transfer_2,
                         H, W, B, c, ws, (float2*)grad_object, (float2*)grad_coded, acc, s);
}

}  // extern "C"
