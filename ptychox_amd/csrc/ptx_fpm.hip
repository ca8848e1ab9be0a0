// Fourier ptychographic microscopy (ptychoX_fourier): forward, adjoint and the
// fused amplitude-loss training step.
//
// Reference semantics: python/ptychoX/fpm/ops.py:56-59 (forward), 85-97
// (backward), 113-124 (wrap-around sub-spectrum indices); CUDA extension
// python/ptychoX/csrc/fpm_kernel.cu:233-317.
//
//   O^ = F O;  sub_b[r,c] = O^[(y_b + r + H/2) mod H, (x_b + c + W/2) mod W]
//   E_b = ifft2(fftshift(sub_b q)),  A_b = |E_b|
//   g_b = ifftshift(fft2(G E/max(|E|,eps))) / (h w)
//   grad_q = sum_b g_b conj(sub_b);  grad_O^[idx_b] += g_b conj(q);  grad_O = H W ifft2(grad_O^)
//
// For even h, w:  ifft2(fftshift(X)) = (-1)^(r+c) ifft2(X)  and
// ifftshift(fft2(Y)) = fft2(Y (-1)^(r+c)), so the shifts are signs in registers
// and the sub-spectrum is gathered / scattered in its natural order.
//
// Passes: object FFT (rows, cols), R1 gather O^ q + row iFFT, C1 column iFFT +
// scale/sign + amplitude | C2 Gamma sign + column FFT, R2 row FFT + pupil
// partials + spectrum scatter (atomics or sorted segments), spectrum iFFT (rows,
// cols).  The step fuses C1 and C2 through the loss.
#include "buffer_io.hpp"
#include "common_kernels.hpp"
#include "passes.hpp"
#include "ptx_internal.hpp"

namespace ptx {
namespace {

// ---------------------------------------------------------------- R1: gather sub-spectrum * pupil, row iFFT
struct FpmGatherRows {
  const float2* spec;  // [H, W] object spectrum
  int64_t H, W;
  const float2* pupil;
  int h;
  const int64_t* led;
  float2* out;  // [B, h, w]
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const int64_t b = valid ? row / h : 0;
    const int r = valid ? (int)(row - b * h) : 0;
    const int64_t sy = (led[2 * b + 1] + r + H / 2) & (H - 1);
    const int64_t sx0 = led[2 * b] + W / 2;
    const float2* srow = spec + sy * W;
    const float2* prow = pupil + (int64_t)r * N;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * T;
      v[q] = valid ? c_mul(srow[(sx0 + i) & (W - 1)], prow[i]) : c_zero();
    }
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    if (!valid) return;
    float2* d = out + row * N;
#pragma unroll
    for (int q = 0; q < E; ++q) d[t + q * T] = v[q];
  }
};

// The same gather with the pupil rows resident: work-group (row group, chunk
// of LEDs) keeps S pupil rows in registers and walks its LEDs, the next LED's
// spectrum row in flight during this LED's transform (two register sets);
// every element's spectrum column wraps modulo W (a power of two), so each
// element has its own buffer offset; rows past h read zero and are stored to
// an out-of-range offset (dropped) -- no branches around memory accesses.
// The exchanges alternate two LDS regions.  Needs the spectrum below 2 GiB.
template <int N, int E>
__global__ void __launch_bounds__(256)
    k_fpm_gather_loop(const float2* __restrict__ spec, int64_t H, int64_t W, const float2* __restrict__ pupil, int h,
                      const int64_t* __restrict__ led, int64_t B, int chunk, int ngroups, float2* __restrict__ out) {
  using Cf = RowCfg<N, E>;
  constexpr int T = Cf::T, S = Cf::S;
  constexpr bool ODD = (FftShape<N, E>::STAGES - 1) % 2 == 1;
  __shared__ float2 lds0[S * lds_len<N>()];
  __shared__ float2 lds1[S * lds_len<N>()];
  __shared__ float2 s_tw[N];
  load_lds_twiddles<N>(s_tw);
  const int seq = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int group = blockIdx.x % ngroups;
  const unsigned ch = blockIdx.x / ngroups;
  const int r = group * S + seq;
  const bool rv = r < h;
  const unsigned b0 = ch * (unsigned)chunk;  // < B
  const unsigned b1 = b0 + chunk < (unsigned)B ? b0 + chunk : (unsigned)B;
  const uint32_t spec_bytes = (uint32_t)(H * W * 8);
  const uint32_t frame_bytes = (uint32_t)((int64_t)h * N * 8);
  constexpr uint32_t kOut = 0x80000000u;  // an offset past every buffer: reads zero, drops stores
  const auto rs = buffer_rsrc(spec, spec_bytes);
  float2 pr[E];
#pragma unroll
  for (int q = 0; q < E; ++q) pr[q] = rv ? pupil[(int64_t)r * N + t + q * T] : c_zero();
  const uint32_t out_lane = rv ? (uint32_t)(((int64_t)r * N + t) * 8) : kOut;
  auto fetch = [&](unsigned b, float2 (&dv)[E]) {
    const unsigned bb = b < b1 ? b : b0;  // past the chunk: a valid LED's row, never used
    const int64_t sy = (led[2 * bb + 1] + r + H / 2) & (H - 1);
    const int64_t sx0 = led[2 * bb] + W / 2;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const uint32_t off = rv ? (uint32_t)((sy * W + ((sx0 + t + q * T) & (W - 1))) * 8) : kOut;
      dv[q] = buf_load_f2(rs, off, 0u);
    }
  };
  auto process = [&](unsigned b, float2 (&v)[E], bool second) {
    int tt = t, sq = seq;
    asm volatile("" : "+v"(tt), "+v"(sq));  // exchange addresses per LED, not hoisted into registers
    float2* l0 = &lds0[sq * lds_len<N>()];
    float2* l1 = &lds1[sq * lds_len<N>()];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = c_mul(v[q], pr[q]);
    constexpr bool WAVE = row_in_wave<N>(E);  // each row within one wave
    if (second)
      fft_stages<N, E, +1, true, 0, true, true, WAVE>(v, tt, l1, l0, s_tw);
    else
      fft_stages<N, E, +1, true, 0, true, true, WAVE>(v, tt, l0, l1, s_tw);
    const auto ro = buffer_rsrc(out + (int64_t)b * h * N, frame_bytes);
#pragma unroll
    for (int q = 0; q < E; ++q) buf_store_f2(v[q], ro, out_lane, (uint32_t)(q * T * 8));
  };
  float2 va[E], vb[E];
  unsigned b = b0;
  fetch(b, va);
  fetch(b + 1, vb);
  process(b, va, false);
  ++b;
  while (b < b1) {
    fetch(b + 1, va);
    process(b, vb, ODD);
    ++b;
    if (b >= b1) break;
    fetch(b + 1, vb);
    process(b, va, false);
    ++b;
  }
}

// ---------------------------------------------------------------- C1: column iFFT -> image field, amplitude
struct FpmImageCols {
  float2* field;  // in: row-transformed, out: image field E
  float* amp;
  int64_t width;
  float inv_n;  // 1/(h w)
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const float2* p = field + frame * N * width;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = p[(t + q * T) * (int)width + (int)col];
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const int64_t off = frame * N * width;  // frame base (uniform)
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int r = t + q * T;
      const float s = ((r + col) & 1) ? -inv_n : inv_n;
      const float2 e = c_scale(v[q], s);
      const int64_t k = off + (r * (int)width + (int)col);
      field[k] = e;
      amp[k] = c_abs(e);
    }
  }
};

// ---------------------------------------------------------------- C2: Gamma * sign, column FFT
struct FpmGammaCols {
  const float2* field;
  const float* g;
  float2* out;
  int64_t width;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const int64_t off = frame * N * width;  // frame base (uniform)
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int r = t + q * T;
      const int64_t k = off + (r * (int)width + (int)col);
      const float2 e = field[k];
      float s = g[k] * fast_rcp(fmaxf(c_abs(e), kFltEps));
      if ((r + col) & 1) s = -s;
      v[q] = c_scale(e, s);
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

// ---------------------------------------------------------------- C12: fused column pass of the step
// column iFFT -> A = |v|/(hw), loss, Gamma*sign = G v/(hw)/max(A, eps) (the two signs cancel) -> column FFT
template <int N, int E, int C>
__global__ void __launch_bounds__((ColCfg<N, E>::T * C))
    k_fpm_step_cols(float2* __restrict__ tb, const float* __restrict__ measured, float* __restrict__ amp_out,
                    float* __restrict__ loss_part, int64_t width, int64_t tiles_per_frame, float gscale, float inv_n) {
  using Cf = ColCfg<N, E>;
  constexpr int T = Cf::T;
  constexpr int THREADS = T * C, WAVES = (THREADS + 63) / 64;
  __shared__ float2 lds[C * col_stride<N, C>()];
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int64_t blk = xcd_paired_block(blockIdx.x, gridDim.x);
  const TileIdx ti = split_tile(blk, tiles_per_frame);
  const int64_t frame = ti.frame;
  const int64_t tile = ti.tile;
  const int64_t col = tile * C + c;
  const int64_t off = frame * N * width;  // frame base (uniform)
  // the frame's buffers: one 32-bit offset per lane, the row steps as scalar
  // offsets; the measured amplitudes are requested with the frame column, so
  // their latency is hidden behind the forward transform
  const uint32_t fbytes = (uint32_t)(N * width * 8);
  const uint32_t lane = (uint32_t)(t * width + col);
  const uint32_t rstep = (uint32_t)(T * width);
  const auto rt = buffer_rsrc(tb + off, fbytes);
  float2 v[E];
  float m[E];
  {
    const auto rm = buffer_rsrc(measured + off, fbytes / 2);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = buf_load_f2(rt, lane * 8u, (uint32_t)q * rstep * 8u);
#pragma unroll
    for (int q = 0; q < E; ++q) m[q] = buf_load_f1(rm, lane * 4u, (uint32_t)q * rstep * 4u);
  }
  // twiddles by separate table loads: measured faster here than the pair table (52 vs 60 us at 256 points)
  fft_seq<N, E, +1, false, false>(v, t, &lds[c * col_stride<N, C>()]);
  float lsum = 0.0f;
#pragma unroll
  for (int q = 0; q < E; ++q) {
    // a = inv_n |v|; inv_n / max(a, eps) = 1 / max(|v|, eps / inv_n)
    const AbsInv ai = c_abs_inv(v[q], inv_n / kFltEps);
    const float a = ai.a * inv_n;
    const float d = a - m[q];
    lsum += d * d;
    m[q] = a;
    v[q] = c_scale(v[q], gscale * d * ai.inv);
  }
  if (amp_out) {
    const auto ra = buffer_rsrc(amp_out + off, fbytes / 2);
#pragma unroll
    for (int q = 0; q < E; ++q) buf_store_f1(m[q], ra, lane * 4u, (uint32_t)q * rstep * 4u);
  }
  // one loss partial per wave (no work-group barrier)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    if (o < THREADS) lsum += __shfl_xor(lsum, o, 64);
  if ((threadIdx.x & 63) == 0) loss_part[(frame * tiles_per_frame + tile) * WAVES + (threadIdx.x >> 6)] = lsum;
  lds_barrier();  // the second transform rewrites the LDS
  fft_seq<N, E, -1, false, false>(v, t, &lds[c * col_stride<N, C>()]);
#pragma unroll
  for (int q = 0; q < E; ++q) buf_store_f2(v[q], rt, lane * 8u, (uint32_t)q * rstep * 8u);
}

// ---------------------------------------------------------------- R2: row FFT, pupil partials, spectrum scatter
// H and W are powers of two: the sub-spectrum wraps with masks.
template <int N, int E>
__global__ void __launch_bounds__(256)
    k_fpm_scatter_rows(const float2* __restrict__ tb, const float2* __restrict__ spec, int64_t H, int64_t W,
                       const float2* __restrict__ pupil, int h, const int64_t* __restrict__ led, int64_t B, int chunk,
                       int ngroups, float inv_n, float2* __restrict__ gspec, float2* __restrict__ contrib,
                       float2* __restrict__ gq_part, SideWork side, int64_t main_blocks) {
  using Cf = RowCfg<N, E>;
  constexpr int T = Cf::T, S = Cf::S;
  if ((int64_t)blockIdx.x >= main_blocks) {  // work-groups riding along
    side.run((int64_t)blockIdx.x - main_blocks);
    return;
  }
  __shared__ float2 lds[S * lds_len<N>()];
  const int seq = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int group = blockIdx.x % ngroups;
  const int ch = blockIdx.x / ngroups;
  const int r = group * S + seq;
  const bool rv = r < h;
  const int64_t b0 = (int64_t)ch * chunk;
  const int64_t b1 = b0 + chunk < B ? b0 + chunk : B;
  float2 qc[E], acc[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    qc[q] = rv ? pupil[(int64_t)r * N + t + q * T] : c_zero();
    acc[q] = c_zero();
  }
  for (int64_t b = b0; b < b1; ++b) {
    float2 v[E];
    const float2* src = tb + (b * h + r) * N;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = rv ? src[t + q * T] : c_zero();
    fft_row<N, E, -1>(v, t, &lds[seq * lds_len<N>()]);
    if (rv) {
      const int64_t sy = (led[2 * b + 1] + r + H / 2) & (H - 1);
      const int64_t sx0 = led[2 * b] + W / 2;
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int i = t + q * T;
        const float2 g = c_scale(v[q], inv_n);
        const int64_t k = sy * W + ((sx0 + i) & (W - 1));
        acc[q] = c_add(acc[q], c_mulc(g, spec[k]));
        const float2 val = c_mulc(g, qc[q]);
        if (contrib) {
          contrib[(b * h + r) * N + i] = val;
        } else {
          unsafeAtomicAdd(&gspec[k].x, val.x);
          unsafeAtomicAdd(&gspec[k].y, val.y);
        }
      }
    }
    seq_sync<row_in_wave<N>(E)>();  // next LED's first LDS write vs this LED's last LDS read
  }
  if (rv) {
    float2* d = gq_part + ((int64_t)ch * h + r) * N;
#pragma unroll
    for (int q = 0; q < E; ++q) d[t + q * T] = acc[q];
  }
}

// columns of [N, width] into `out` (= or +=), used for the last pass of grad_O = HW ifft2(grad_O^)
struct ColsToOut {
  const float2* in;
  float2* out;
  int64_t width;
  int accumulate;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const float2* p = in + frame * N * width;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = p[(t + q * T) * (int)width + (int)col];
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    float2* p = out + frame * N * width;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      float2* d = p + (t + q * T) * (int)width + (int)col;
      *d = accumulate ? c_add(*d, v[q]) : v[q];
    }
  }
};

// ---------------------------------------------------------------- host launchers

template <int N, int E>
struct RowsRun {
  template <int DIR, class Op>
  static ptx_status go(const Op& op, int64_t nrows, hipStream_t s, const char* what) {
    launch_rows<N, E, DIR>(op, nrows, s);
    return check_launch(what, s);
  }
};
template <int N, int E>
struct ObjFftRows {
  static ptx_status run(const float2* in, float2* out, int64_t nrows, int dir, hipStream_t s) {
    RowsOutOfPlace op{in, out, 1.0f};
    return dir < 0 ? RowsRun<N, E>::template go<-1>(op, nrows, s, "fft rows")
                   : RowsRun<N, E>::template go<+1>(op, nrows, s, "ifft rows");
  }
};
template <int N, int E>
struct ObjFftColsOut {
  static ptx_status run(const ColsToOut& op, int dir, hipStream_t s) {
    if (dir < 0)
      launch_cols<N, E, -1>(op, 1, op.width, s);
    else
      launch_cols<N, E, +1>(op, 1, op.width, s);
    return check_launch(dir < 0 ? "fft cols" : "ifft cols", s);
  }
};
template <int N, int E>
struct FpmGatherRun {
  static ptx_status run(const FpmGatherRows& op, int64_t nrows, hipStream_t s) {
    using Cf = RowCfg<N, E>;
    const int64_t B = op.h > 0 ? nrows / op.h : 0;
    if (B > 0 && op.H * op.W * 8 < (int64_t(1) << 31) && (int64_t)op.h * N * 8 < (int64_t(1) << 31)) {
      // pupil rows resident: ~1024 work-groups (tunable), each walking a chunk of LEDs
      const int ngroups = (int)((op.h + Cf::S - 1) / Cf::S);
      int64_t nchunks = (tuned_groups(PASS_ROW, N, 1024) + ngroups - 1) / ngroups;
      if (nchunks > B) nchunks = B;
      const int64_t chunk = (B + nchunks - 1) / nchunks;
      nchunks = (B + chunk - 1) / chunk;
      k_fpm_gather_loop<N, E><<<dim3((unsigned)(ngroups * nchunks)), dim3(256), 0, s>>>(
          op.spec, op.H, op.W, op.pupil, op.h, op.led, B, (int)chunk, ngroups, op.out);
    } else {
      launch_rows<N, E, +1>(op, nrows, s);
    }
    return check_launch("fpm gather rows", s);
  }
};
template <int N, int E>
struct FpmImageColsRun {
  static ptx_status run(const FpmImageCols& op, int64_t frames, hipStream_t s) {
    launch_cols<N, E, +1>(op, frames, op.width, s);
    return check_launch("fpm image cols", s);
  }
};
template <int N, int E>
struct FpmGammaColsRun {
  static ptx_status run(const FpmGammaCols& op, int64_t frames, hipStream_t s) {
    launch_cols<N, E, -1>(op, frames, op.width, s);
    return check_launch("fpm gamma cols", s);
  }
};
template <int N, int E>
struct FpmStepColsRun {
  static ptx_status run(float2* tb, const float* measured, float* amp, float* loss_part, int64_t frames, int64_t width,
                        float gscale, float inv_n, hipStream_t s) {
    using Cf = ColCfg<N, E>;
    if (width >= Cf::CMAX) {
      constexpr int C = Cf::CMAX;
      const int64_t tiles = width / C;
      k_fpm_step_cols<N, E, C><<<dim3((unsigned)(frames * tiles)), dim3(Cf::T * C), 0, s>>>(
          tb, measured, amp, loss_part, width, tiles, gscale, inv_n);
    } else {
      constexpr int C = 8;
      const int64_t tiles = width / C;
      k_fpm_step_cols<N, E, C><<<dim3((unsigned)(frames * tiles)), dim3(Cf::T * C), 0, s>>>(
          tb, measured, amp, loss_part, width, tiles, gscale, inv_n);
    }
    return check_launch("fpm step cols", s);
  }
};
// loss partials per frame of k_fpm_step_cols: one per wave of each column tile
template <int N, int E>
struct LossPartsQuery {
  static ptx_status run(int64_t width, int64_t* parts) {
    const int c = width >= ColCfg<N, E>::CMAX ? ColCfg<N, E>::CMAX : 8;
    *parts = (width / c) * ((ColCfg<N, E>::T * c + 63) / 64);
    return PTX_OK;
  }
};

struct Plan {
  int S, ngroups, chunk, nchunks;
};
template <int N, int E>
struct PlanQuery {
  static ptx_status run(int64_t h, int64_t B, Plan* p) {
    p->S = RowCfg<N, E>::S;
    p->ngroups = (int)((h + p->S - 1) / p->S);
    int64_t want = (tuned_groups(PASS_ROW_LOOP, N, kRowLoopGroups) + p->ngroups - 1) / p->ngroups;
    if (want > B) want = B > 0 ? B : 1;
    if (want < 1) want = 1;
    p->chunk = (int)((B + want - 1) / want);
    if (p->chunk < 1) p->chunk = 1;
    p->nchunks = (int)((B + p->chunk - 1) / p->chunk);
    if (p->nchunks < 1) p->nchunks = 1;
    return PTX_OK;
  }
};
template <int N, int E>
struct ScatterRun {
  static ptx_status run(const float2* tb, const float2* spec, int64_t H, int64_t W, const float2* pupil, int h,
                        const int64_t* led, int64_t B, const Plan& p, float inv_n, float2* gspec, float2* contrib,
                        float2* gq_part, const SideWork& side, hipStream_t s) {
    const int64_t main_blocks = (int64_t)p.ngroups * p.nchunks;
    k_fpm_scatter_rows<N, E><<<dim3((unsigned)(main_blocks + side.blocks())), dim3(256), 0, s>>>(
        tb, spec, H, W, pupil, h, led, B, p.chunk, p.ngroups, inv_n, gspec, contrib, gq_part, side, main_blocks);
    return check_launch("fpm scatter rows", s);
  }
};

struct FpmWorkspace {
  float2* spec;      // [H, W] object spectrum (step only)
  float2* gspec;     // [H, W] spectrum gradient
  float2* tb;        // [B, h, w]
  float2* gq_part;   // [nchunks, h, w]
  float* loss_part;  // [B, loss partials per frame] (step only)
  GatherScratch gather;  // deterministic only
  size_t bytes;
};

FpmWorkspace carve(void* base, int64_t H, int64_t W, int64_t h, int64_t w, int64_t B, unsigned flags, bool step,
                   const Plan& p, int64_t loss_parts) {
  Carver cv{(char*)base, 0};
  FpmWorkspace ws{};
  ws.spec = step ? cv.take<float2>((size_t)(H * W)) : nullptr;
  ws.gspec = cv.take<float2>((size_t)(H * W));
  ws.tb = cv.take<float2>((size_t)(B * h * w));
  ws.gq_part = cv.take<float2>((size_t)(p.nchunks * h * w));
  ws.loss_part = step ? cv.take<float>((size_t)(B * loss_parts)) : nullptr;
  if (use_segments(flags)) ws.gather = carve_gather(cv, H, W, h, w, B);
  ws.bytes = cv.used;
  return ws;
}

ptx_status check(const char* fn, int64_t H, int64_t W, int64_t h, int64_t w, int64_t B) {
  if (B < 0) {
    set_error("%s: negative batch", fn);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (!size_ok(H) || !size_ok(W) || !size_ok(h) || !size_ok(w)) {
    set_error("%s: object %lld x %lld / pupil %lld x %lld unsupported (sides must be powers of two in [%d, %d])", fn,
              (long long)H, (long long)W, (long long)h, (long long)w, PTX_MIN_N, PTX_MAX_N);
    return PTX_ERR_UNSUPPORTED;
  }
  if (h > H || w > W) {
    set_error("%s: pupil %lld x %lld larger than the object spectrum %lld x %lld", fn, (long long)h, (long long)w,
              (long long)H, (long long)W);
    return PTX_ERR_OUT_OF_RANGE;
  }
  return PTX_OK;
}

// object spectrum: rows (out of place) then columns into `spec`
ptx_status spectrum_of(const float2* obj, float2* spec, int64_t H, int64_t W, hipStream_t s) {
  ptx_status st = dispatch_ne<ObjFftRows>(PASS_ROW, W, obj, spec, H, -1, s);
  if (st != PTX_OK) return st;
  ColsToOut c{spec, spec, W, 0};
  return dispatch_ne<ObjFftColsOut>(PASS_COL, H, c, -1, s);
}

// R2 + pupil reduction + spectrum scatter + grad_O = HW ifft2(grad_O^).  The
// per-LED loss sums (a step's) and the tile lists ride along with the scatter
// launch, the pupil-partial sum with the tile gather (deterministic mode).
ptx_status fpm_tail(const FpmWorkspace& ws, const Plan& p, const float2* spec, int64_t H, int64_t W,
                    const float2* pupil, int64_t h, int64_t w, const int64_t* led, int64_t B, float2* gobj,
                    float2* gpupil, unsigned flags, hipStream_t s, const LossRows& loss = LossRows{}) {
  const bool det = use_segments(flags);
  const bool acc = flags & PTX_FLAG_ACCUMULATE;
  ptx_status st = PTX_OK;
  if (!det) {
    fill_zero2(ws.gspec, H * W, s);
    if ((st = check_launch("zero fill", s)) != PTX_OK) return st;
  }
  const float inv_n = 1.0f / (float)(h * w);
  SideWork with_scatter{};
  if (det) with_scatter.lists = gather_lists_work(ws.gather, led, B, h, w, H, W, H / 2, W / 2, false);
  with_scatter.loss = loss;
  st = dispatch_ne<ScatterRun>(PASS_ROW_LOOP, w, ws.tb, spec, H, W, pupil, (int)h, led, B, p, inv_n, ws.gspec,
                              det ? ws.tb : nullptr, ws.gq_part, with_scatter, s);
  if (st != PTX_OK) return st;
  const SumParts2 pupil_sum{ws.gq_part, p.nchunks, h * w, gpupil, acc ? 1 : 0, 0};
  if (det) {
    SideWork with_gather{};
    with_gather.sums = pupil_sum;
    st = run_tile_gather(ws.gather, ws.tb, h, w, led, B, H / 2, W / 2, ws.gspec, H, W, 0, s, with_gather);
    if (st != PTX_OK) return st;
  } else {
    sum_parts2(ws.gq_part, p.nchunks, h * w, gpupil, acc ? 1 : 0, s);
    if ((st = check_launch("pupil partial sum", s)) != PTX_OK) return st;
  }
  st = dispatch_ne<ObjFftRows>(PASS_ROW, W, ws.gspec, ws.gspec, H, +1, s);
  if (st != PTX_OK) return st;
  ColsToOut c{ws.gspec, gobj, W, acc ? 1 : 0};
  return dispatch_ne<ObjFftColsOut>(PASS_COL, H, c, +1, s);
}

}  // namespace
}  // namespace ptx

using namespace ptx;

extern "C" {

ptx_status ptx_fpm_forward(const void* object_field, int64_t H, int64_t W, const void* pupil, int64_t h, int64_t w,
                           const int64_t* led, int64_t B, void* object_fft, void* image_field, float* amplitude,
                           void* stream) {
  TuneScope tune_scope;
  ptx_status st = check("ptx_fpm_forward", H, W, h, w, B);
  if (st != PTX_OK) return st;
  if (!object_field || !pupil || !object_fft || (B > 0 && (!led || !image_field || !amplitude))) {
    set_error("ptx_fpm_forward: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  timing_begin(s);
  st = spectrum_of((const float2*)object_field, (float2*)object_fft, H, W, s);
  if (st != PTX_OK || B == 0) return st;
  FpmGatherRows r1{(const float2*)object_fft, H, W, (const float2*)pupil, (int)h, led, (float2*)image_field};
  st = dispatch_ne<FpmGatherRun>(PASS_ROW, w, r1, B * h, s);
  if (st != PTX_OK) return st;
  FpmImageCols c1{(float2*)image_field, amplitude, w, 1.0f / (float)(h * w)};
  return dispatch_ne<FpmImageColsRun>(PASS_COL, h, c1, B, s);
}

size_t ptx_fpm_backward_workspace(int64_t H, int64_t W, int64_t h, int64_t w, int64_t B, unsigned flags) {
  TuneScope tune_scope;
  if (check("ptx_fpm_backward_workspace", H, W, h, w, B) != PTX_OK) return 0;
  Plan p{};
  dispatch_ne<PlanQuery>(PASS_ROW_LOOP, w, h, B, &p);
  return carve(nullptr, H, W, h, w, B, flags, false, p, 0).bytes;
}

ptx_status ptx_fpm_backward(const float* grad_output, const void* image_field, const void* object_fft, int64_t H,
                            int64_t W, const void* pupil, int64_t h, int64_t w, const int64_t* led, int64_t B,
                            void* grad_object, void* grad_pupil, void* workspace, size_t ws_bytes, unsigned flags,
                            void* stream) {
  TuneScope tune_scope;
  ptx_status st = check("ptx_fpm_backward", H, W, h, w, B);
  if (st != PTX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  if (!grad_object || !grad_pupil) {
    set_error("ptx_fpm_backward: null gradient output");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  timing_begin(s);
  if (B == 0) {
    if (!(flags & PTX_FLAG_ACCUMULATE)) {
      fill_zero2((float2*)grad_object, H * W, s);
      fill_zero2((float2*)grad_pupil, h * w, s);
    }
    return check_launch("zero fill", s);
  }
  if (!grad_output || !image_field || !object_fft || !pupil || !led) {
    set_error("ptx_fpm_backward: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  Plan p{};
  dispatch_ne<PlanQuery>(PASS_ROW_LOOP, w, h, B, &p);
  const FpmWorkspace ws = carve(workspace, H, W, h, w, B, flags, false, p, 0);
  if (!workspace || ws_bytes < ws.bytes) {
    set_error("ptx_fpm_backward: workspace of %zu bytes needed, %zu given", ws.bytes, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  FpmGammaCols c2{(const float2*)image_field, grad_output, ws.tb, w};
  st = dispatch_ne<FpmGammaColsRun>(PASS_COL, h, c2, B, s);
  if (st != PTX_OK) return st;
  return fpm_tail(ws, p, (const float2*)object_fft, H, W, (const float2*)pupil, h, w, led, B, (float2*)grad_object,
                  (float2*)grad_pupil, flags, s);
}

size_t ptx_fpm_step_workspace(int64_t H, int64_t W, int64_t h, int64_t w, int64_t B, unsigned flags) {
  TuneScope tune_scope;
  if (check("ptx_fpm_step_workspace", H, W, h, w, B) != PTX_OK) return 0;
  Plan p{};
  dispatch_ne<PlanQuery>(PASS_ROW_LOOP, w, h, B, &p);
  int64_t loss_parts = 0;
  dispatch_ne<LossPartsQuery>(PASS_COL, h, w, &loss_parts);
  return carve(nullptr, H, W, h, w, B, flags, true, p, loss_parts).bytes;
}

ptx_status ptx_fpm_step(const void* object_field, int64_t H, int64_t W, const void* pupil, int64_t h, int64_t w,
                        const int64_t* led, int64_t B, const float* measured, float loss_scale,
                        float* loss_per_position, float* amplitude, void* grad_object, void* grad_pupil,
                        void* workspace, size_t ws_bytes, unsigned flags, void* stream) {
  TuneScope tune_scope;
  ptx_status st = check("ptx_fpm_step", H, W, h, w, B);
  if (st != PTX_OK) return st;
  hipStream_t s = (hipStream_t)stream;
  if (!grad_object || !grad_pupil) {
    set_error("ptx_fpm_step: null gradient output");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  timing_begin(s);
  if (B == 0) {
    if (!(flags & PTX_FLAG_ACCUMULATE)) {
      fill_zero2((float2*)grad_object, H * W, s);
      fill_zero2((float2*)grad_pupil, h * w, s);
    }
    return check_launch("zero fill", s);
  }
  if (!object_field || !pupil || !led || !measured || !loss_per_position) {
    set_error("ptx_fpm_step: null pointer argument");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  Plan p{};
  dispatch_ne<PlanQuery>(PASS_ROW_LOOP, w, h, B, &p);
  int64_t loss_parts = 0;
  dispatch_ne<LossPartsQuery>(PASS_COL, h, w, &loss_parts);
  const FpmWorkspace ws = carve(workspace, H, W, h, w, B, flags, true, p, loss_parts);
  if (!workspace || ws_bytes < ws.bytes) {
    set_error("ptx_fpm_step: workspace of %zu bytes needed, %zu given", ws.bytes, ws_bytes);
    return PTX_ERR_WORKSPACE;
  }
  st = spectrum_of((const float2*)object_field, ws.spec, H, W, s);
  if (st != PTX_OK) return st;
  FpmGatherRows r1{ws.spec, H, W, (const float2*)pupil, (int)h, led, ws.tb};
  st = dispatch_ne<FpmGatherRun>(PASS_ROW, w, r1, B * h, s);
  if (st != PTX_OK) return st;
  const float inv_n = 1.0f / (float)(h * w);
  st = dispatch_ne<FpmStepColsRun>(PASS_COL, h, ws.tb, measured, amplitude, ws.loss_part, B, w, 2.0f * loss_scale, inv_n, s);
  if (st != PTX_OK) return st;
  return fpm_tail(ws, p, ws.spec, H, W, (const float2*)pupil, h, w, led, B, (float2*)grad_object,
                  (float2*)grad_pupil, flags, s, LossRows{ws.loss_part, loss_parts, inv_n, loss_per_position, B});
}

}  // extern "C"
