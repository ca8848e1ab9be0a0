// Pieces shared by the scan-window operators (conventional and near-field):
// the gather of sample window * probe into row transforms, the row pass that
// scatters the object gradient and accumulates the probe gradient over a chunk
// of views in registers, the launch plan and the workspace layout.
#pragma once

#include <type_traits>

#include "buffer_io.hpp"
#include "common_kernels.hpp"
#include "passes.hpp"
#include "ptx_internal.hpp"

namespace ptx {
namespace {

// ---------------------------------------------------------------- gather sample window * probe (* sign), row FFT
template <bool SIGN>
struct ScanGatherRows {
  const float2* sample;
  int64_t H, W;
  const float2* probe;
  int h;
  const int64_t* pos;
  float2* out;  // [B, h, w]
  // Indexed batches (the view-looping kernel only): view b is entry order[b]
  // of pos[count] (an entry outside [0, count) reads entry 0), and work-group
  // 0 writes the batch's positions to pos_out for the passes that follow.
  const int64_t* order = nullptr;
  int64_t count = 0;
  int64_t* pos_out = nullptr;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const int64_t b = valid ? row / h : 0;
    const int r = valid ? (int)(row - b * h) : 0;
    const int64_t x0 = pos[2 * b], y0 = pos[2 * b + 1];
    const bool ok = valid && window_ok(x0, y0, h, N, H, W);
    const float2* srow = sample + (y0 + r) * W + x0;
    const float2* prow = probe + (int64_t)r * N;
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int i = t + q * T;
      float2 e = ok ? c_mul(srow[i], prow[i]) : c_zero();
      v[q] = (SIGN && ((r + i) & 1)) ? c_neg(e) : e;
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

// ---------------------------------------------------------------- the same gather, probe rows resident
// Work-group (row group, chunk of views): S probe rows (with the fftshift sign
// folded in) stay in registers while the work-group walks its views, so the
// probe is read once per chunk instead of once per view.  The views are
// software-pipelined: two register sets alternate, the next view's window
// row in flight during this view's transform.  Loads and stores are raw
// buffer accesses without branches: an invalid window, a row past h or a
// view past the chunk reads zero through a zero-length buffer or an
// out-of-range offset, and such rows are stored to an out-of-range offset
// (dropped), so the compiler's wait counts keep the prefetch in flight.  The
// exchanges alternate two LDS regions.  Same arithmetic per element as
// ScanGatherRows.  Needs the object below 2 GiB (32-bit buffer offsets).
// order != null: view b is position order[b] of pos[count] (entry 0 for an
// order entry outside [0, count)), and work-group 0 also writes the batch's
// positions to pos_out, so no separate gather of the positions runs before.
inline bool gather_loop_ok(int64_t H, int64_t W, int64_t h, int64_t n) {
  return H * W * 8 < (int64_t(1) << 31) && h * n * 8 < (int64_t(1) << 31);
}
template <int N, int E, bool SIGN>
__global__ void __launch_bounds__(256)
    k_scan_gather_loop(const float2* __restrict__ sample, int64_t H, int64_t W, const float2* __restrict__ probe,
                       int h, const int64_t* __restrict__ pos, int64_t B, int chunk, int ngroups,
                       float2* __restrict__ out, const int64_t* __restrict__ order, int64_t count,
                       int64_t* __restrict__ pos_out) {
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
  const uint32_t obj_bytes = (uint32_t)(H * W * 8);
  const uint32_t frame_bytes = (uint32_t)((int64_t)h * N * 8);
  constexpr uint32_t kOut = 0x80000000u;  // an offset past every buffer: reads zero, drops stores
  float2 pr[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int i = t + q * T;
    const float2 p = rv ? probe[(int64_t)r * N + i] : c_zero();
    pr[q] = (SIGN && ((r + i) & 1)) ? c_neg(p) : p;
  }
  const uint32_t out_lane = rv ? (uint32_t)(((int64_t)r * N + t) * 8) : kOut;
  auto entry = [&](int64_t b) -> int64_t {
    if (!order) return b;
    const int64_t k = order[b];
    return (uint64_t)k < (uint64_t)count ? k : 0;
  };
  if (pos_out && blockIdx.x == 0) {
    for (int64_t b = threadIdx.x; b < B; b += 256) {
      const int64_t k = entry(b);
      pos_out[2 * b] = pos[2 * k];
      pos_out[2 * b + 1] = pos[2 * k + 1];
    }
  }
  auto fetch = [&](unsigned b, float2 (&dv)[E]) {
    const bool live = b < b1;
    const int64_t k = entry(live ? b : b0);
    const int64_t x0 = pos[2 * k], y0 = pos[2 * k + 1];
    const bool ok = live && window_ok(x0, y0, h, N, H, W);
    const auto rs = buffer_rsrc(sample, ok ? obj_bytes : 0u);
    const uint32_t voff = (ok && rv) ? (uint32_t)(((y0 + r) * W + x0 + t) * 8) : kOut;
#pragma unroll
    for (int q = 0; q < E; ++q) dv[q] = buf_load_f2(rs, voff, (uint32_t)(q * T * 8));
  };
  auto process = [&](unsigned b, float2 (&v)[E], bool second) {
    int tt = t, sq = seq;
    asm volatile("" : "+v"(tt), "+v"(sq));  // exchange addresses per view, not hoisted into registers
    float2* l0 = &lds0[sq * lds_len<N>()];
    float2* l1 = &lds1[sq * lds_len<N>()];
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = c_mul(v[q], pr[q]);
    constexpr bool WAVE = row_in_wave<N>(E);  // each row within one wave
    if (second)
      fft_stages<N, E, -1, true, 0, true, true, WAVE>(v, tt, l1, l0, s_tw);
    else
      fft_stages<N, E, -1, true, 0, true, true, WAVE>(v, tt, l0, l1, s_tw);
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

// ---------------------------------------------------------------- row iFFT, (sign | scale), object scatter, probe partials
template <int N, int E, bool SIGN>
__global__ void __launch_bounds__(256)
    k_scan_scatter_rows(const float2* __restrict__ tb, const float2* __restrict__ sample, int64_t H, int64_t W,
                        const float2* __restrict__ probe, int h, const int64_t* __restrict__ pos, int64_t B,
                        int chunk, int ngroups, float2* __restrict__ gobj, float2* __restrict__ contrib,
                        float2* __restrict__ gp_part, float scale, SideWork side, int64_t main_blocks) {
  using Cf = RowCfg<N, E>;
  constexpr int T = Cf::T, S = Cf::S;
  if ((int64_t)blockIdx.x >= main_blocks) {  // work-groups riding along
    side.run((int64_t)blockIdx.x - main_blocks);
    return;
  }
  const int blk = (int)blockIdx.x;
  __shared__ float2 lds[S * lds_len<N>()];
  const int seq = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int group = blk % ngroups;
  const int ch = blk / ngroups;
  const int r = group * S + seq;
  const bool rv = r < h;
  const int64_t b0 = (int64_t)ch * chunk;
  const int64_t b1 = b0 + chunk < B ? b0 + chunk : B;
  float2 pr[E], acc[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    pr[q] = rv ? probe[(int64_t)r * N + t + q * T] : c_zero();
    acc[q] = c_zero();
  }
  for (int64_t b = b0; b < b1; ++b) {
    float2 v[E];
    const float2* src = tb + (b * h + r) * N;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = rv ? src[t + q * T] : c_zero();
    fft_row<N, E, +1>(v, t, &lds[seq * lds_len<N>()]);
    const int64_t x0 = pos[2 * b], y0 = pos[2 * b + 1];
    if (rv && window_ok(x0, y0, h, N, H, W)) {
      const int64_t obase = (y0 + r) * W + x0;
#pragma unroll
      for (int q = 0; q < E; ++q) {
        const int i = t + q * T;
        const float2 chi = SIGN ? (((r + i) & 1) ? c_neg(v[q]) : v[q]) : c_scale(v[q], scale);
        acc[q] = c_add(acc[q], c_mulc(chi, sample[obase + i]));
        const float2 val = c_mulc(chi, pr[q]);
        if (contrib) {
          contrib[(b * h + r) * N + i] = val;
        } else {
          unsafeAtomicAdd(&gobj[obase + i].x, val.x);
          unsafeAtomicAdd(&gobj[obase + i].y, val.y);
        }
      }
    }
    seq_sync<row_in_wave<N>(E)>();  // next view's first LDS write vs this view's last LDS read
  }
  if (rv) {
    float2* d = gp_part + ((int64_t)ch * h + r) * N;
#pragma unroll
    for (int q = 0; q < E; ++q) d[t + q * T] = acc[q];
  }
}

// ---------------------------------------------------------------- host side

template <int N, int E>
struct GatherRowsRun {
  template <bool SIGN>
  static ptx_status run(const ScanGatherRows<SIGN>& op, int64_t nrows, hipStream_t s) {
    using Cf = RowCfg<N, E>;
    const int64_t B = op.h > 0 ? nrows / op.h : 0;
    if (B > 0 && gather_loop_ok(op.H, op.W, op.h, N)) {
      // probe rows resident: ~1024 work-groups (tunable), each walking a chunk of views
      const int ngroups = (int)((op.h + Cf::S - 1) / Cf::S);
      int64_t nchunks = (tuned_groups(PASS_ROW, N, 1024) + ngroups - 1) / ngroups;
      if (nchunks > B) nchunks = B;
      const int64_t chunk = (B + nchunks - 1) / nchunks;
      nchunks = (B + chunk - 1) / chunk;
      k_scan_gather_loop<N, E, SIGN><<<dim3((unsigned)(ngroups * nchunks)), dim3(256), 0, s>>>(
          op.sample, op.H, op.W, op.probe, op.h, op.pos, B, (int)chunk, ngroups, op.out, op.order, op.count,
          op.pos_out);
    } else {
      if (op.order) {
        set_error("scan gather rows: indexed views need the view-looping kernel");
        return PTX_ERR_INVALID_ARGUMENT;
      }
      launch_rows<N, E, -1>(op, nrows, s);
    }
    return check_launch(SIGN ? "base gather rows" : "scan gather rows", s);
  }
};

struct ScatterPlan {
  int S;        // rows per work-group
  int ngroups;  // row groups
  int chunk;    // views per work-group
  int nchunks;  // probe partial slices
};

template <int N, int E>
struct ScatterPlanQuery {
  static ptx_status run(int64_t h, int64_t B, ScatterPlan* p) {
    p->S = RowCfg<N, E>::S;
    p->ngroups = (int)((h + p->S - 1) / p->S);
    // aim for kRowLoopGroups work-groups (tunable); each probe-partial slice covers `chunk` views
    int64_t want = (tuned_groups(PASS_ROW_LOOP, N, kRowLoopGroups) + p->ngroups - 1) / p->ngroups;
    if (want < 1) want = 1;
    if (want > B) want = B > 0 ? B : 1;
    p->chunk = (int)((B + want - 1) / want);
    if (p->chunk < 1) p->chunk = 1;
    p->nchunks = (int)((B + p->chunk - 1) / p->chunk);
    if (p->nchunks < 1) p->nchunks = 1;
    return PTX_OK;
  }
};

template <int N, int E>
struct ScatterRowsRun {
  template <bool SIGN>
  static ptx_status run(const float2* tb, const float2* sample, int64_t H, int64_t W, const float2* probe, int h,
                        const int64_t* pos, int64_t B, const ScatterPlan& p, float2* gobj, float2* contrib,
                        float2* gp_part, float scale, const SideWork& side, hipStream_t s,
                        std::integral_constant<bool, SIGN>) {
    const int64_t main_blocks = (int64_t)p.ngroups * p.nchunks;
    k_scan_scatter_rows<N, E, SIGN><<<dim3((unsigned)(main_blocks + side.blocks())), dim3(256), 0, s>>>(
        tb, sample, H, W, probe, h, pos, B, p.chunk, p.ngroups, gobj, contrib, gp_part, scale, side, main_blocks);
    return check_launch(SIGN ? "base scatter rows" : "scan scatter rows", s);
  }
};

struct ScanWorkspace {
  float2* tb;        // [B, h, w]
  float2* gp_part;   // [nchunks, h, w]
  float* loss_part;  // [B, loss partials per frame]   (step only)
  GatherScratch gather;  // deterministic only
  int64_t* pos_batch;  // [B, 2] gathered positions (indexed step only)
  size_t bytes;
};

ScanWorkspace carve_scan(void* base, int64_t H, int64_t W, int64_t h, int64_t w, int64_t B, unsigned flags,
                         bool step, const ScatterPlan& sp, int64_t loss_parts, bool indexed = false) {
  Carver cv{(char*)base, 0};
  ScanWorkspace ws{};
  ws.tb = cv.take<float2>((size_t)(B * h * w));
  ws.gp_part = cv.take<float2>((size_t)(sp.nchunks * h * w));
  ws.loss_part = step ? cv.take<float>((size_t)(B * loss_parts)) : nullptr;
  if (use_segments(flags)) ws.gather = carve_gather(cv, H, W, h, w, B);
  if (indexed) ws.pos_batch = cv.take<int64_t>((size_t)(2 * B));
  ws.bytes = cv.used;
  return ws;
}

ptx_status check_shapes(const char* fn, int64_t H, int64_t W, int64_t h, int64_t w, int64_t B) {
  if (B < 0 || H <= 0 || W <= 0) {
    set_error("%s: invalid sizes (sample %lld x %lld, batch %lld)", fn, (long long)H, (long long)W, (long long)B);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (!size_ok(h) || !size_ok(w)) {
    set_error("%s: probe %lld x %lld unsupported (sides must be powers of two in [%d, %d])", fn, (long long)h,
              (long long)w, PTX_MIN_N, PTX_MAX_N);
    return PTX_ERR_UNSUPPORTED;
  }
  if (h > H || w > W) {
    set_error("%s: probe %lld x %lld larger than sample %lld x %lld", fn, (long long)h, (long long)w, (long long)H,
              (long long)W);
    return PTX_ERR_OUT_OF_RANGE;
  }
  return PTX_OK;
}

// shared tail of backward and step: R2 scatter, probe reduction, deterministic
// gather.  The tile lists (they need only the positions) and the step's
// per-view loss sums (`loss`, batch 0 for none) ride along with the scatter
// launch, the probe-partial sum with the gather launch.
// With adam_obj / adam_probe set (param non-null) the gradients feed those
// Adam updates: fused into the gather / probe-sum epilogues, or run as the
// stand-alone update after the atomic scatter (gs, gp are then scratch).
template <bool SIGN>
ptx_status scan_scatter(const ScanWorkspace& ws, const ScatterPlan& sp, const float2* sample, int64_t H, int64_t W,
                        const float2* probe, int64_t h, int64_t w, const int64_t* pos, int64_t B, float2* gs,
                        float2* gp, unsigned flags, float scale, hipStream_t s, const LossRows& loss = LossRows{},
                        const AdamArgs& adam_obj = AdamArgs{}, const AdamArgs& adam_probe = AdamArgs{}) {
  const bool det = use_segments(flags);
  const bool acc = flags & PTX_FLAG_ACCUMULATE;
  ptx_status st = PTX_OK;
  if (!det && !acc) {
    fill_zero2(gs, H * W, s);
    if ((st = check_launch("zero fill", s)) != PTX_OK) return st;
  }
  SideWork with_scatter{};
  if (det) with_scatter.lists = gather_lists_work(ws.gather, pos, B, h, w, H, W, 0, 0, true);
  with_scatter.loss = loss;
  st = dispatch_ne<ScatterRowsRun>(PASS_ROW_LOOP, w, ws.tb, sample, H, W, probe, (int)h, pos, B, sp, gs,
                                  det ? ws.tb : nullptr, ws.gp_part, scale, with_scatter, s,
                                  std::integral_constant<bool, SIGN>{});
  if (st != PTX_OK) return st;
  SumParts2 probe_sum{ws.gp_part, sp.nchunks, h * w, gp, acc ? 1 : 0, 0};
  probe_sum.adam = adam_probe;
  if (!det) {
    if (probe_sum.blocks() > 0) k_sum_parts2<<<dim3((unsigned)probe_sum.blocks()), dim3(256), 0, s>>>(probe_sum);
    if ((st = check_launch("probe partial sum", s)) != PTX_OK) return st;
    if (adam_obj.param) return launch_adam(adam_obj, gs, H * W, s);  // after the atomic scatter
    return PTX_OK;
  }
  SideWork with_gather{};
  with_gather.sums = probe_sum;
  return run_tile_gather(ws.gather, ws.tb, h, w, pos, B, 0, 0, gs, H, W, acc ? 1 : 0, s, with_gather, adam_obj);
}

}  // namespace
}  // namespace ptx
