// Small helper kernels shared by the operators: fills, fixed-order partial
// reductions, block reductions, and the sorted-segment (deterministic)
// scatter-add of per-view windows into the object.
#pragma once

#include "fft_core.hpp"
#include "ptx_internal.hpp"

namespace ptx {

constexpr float kFltEps = 1.1920928955078125e-07f;  // FLT_EPSILON, python/ptychoX/base/ops.py:93

__device__ __forceinline__ bool window_ok(int64_t x0, int64_t y0, int h, int w, int64_t H, int64_t W) {
  return x0 >= 0 && y0 >= 0 && x0 + w <= W && y0 + h <= H;
}

// graph-safe zero fill (no hipMemsetAsync node)
static __global__ void __launch_bounds__(256) k_fill_zero2(float2* __restrict__ p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = c_zero();
}
inline void fill_zero2(float2* p, int64_t n, hipStream_t s) {
  if (n > 0) k_fill_zero2<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(p, n);
}

// One Adam update of a complex element, per real component, with
// torch.optim.Adam's foreach arithmetic as the drivers configure it
// (python/train_ptychography.py:500-507):
//   m = lerp(m, g, 1 - b1);  v = b2 v + (1 - b2) g^2
//   p += -(lr / (1 - b1^t)) * m / (sqrt(v) / sqrt(1 - b2^t) + eps)
// hyper: device float[3] = {lr / (1 - b1^t), sqrt(1 - b2^t), eps}.  Used by
// the stand-alone Adam kernel and by the gradient epilogues that fuse the
// update (param == null: no update, the epilogue stores the gradient).
// Every rounding is spelled out (contraction off, the fused multiply-adds
// written as such), so the update is bitwise the same in every kernel that
// inlines it: one fma per foreach op (lerp, addcmul after the scale by b2,
// addcdiv), the square root, the divisions and the eps add rounded separately.
__device__ __forceinline__ float adam_component(float p, float g, float& m, float& v, float b1, float b2,
                                                float step, float bc2, float eps) {
#pragma clang fp contract(off)
  const float mi = __builtin_fmaf(1.0f - b1, g - m, m);        // lerp(m, g, 1 - b1), weight < 0.5
  const float vi = __builtin_fmaf((1.0f - b2) * g, g, v * b2);  // v * b2, then addcmul(g, g, 1 - b2)
  m = mi;
  v = vi;
  const float denom = sqrtf(vi) / bc2 + eps;
  return __builtin_fmaf(-step, mi / denom, p);  // addcdiv(m, denom, -step)
}

struct AdamArgs {
  float2* param;  // complex parameter, updated in place
  float* m;       // first moments, 2 floats per element
  float* v;       // second moments
  const float* hyper;
  float beta1, beta2;
  __device__ __forceinline__ void apply(int64_t i, float2 g) const {
    const float step = hyper[0], bc2 = hyper[1], eps = hyper[2];
    float2 p = param[i];
    float m0 = m[2 * i], m1 = m[2 * i + 1], v0 = v[2 * i], v1 = v[2 * i + 1];
    p.x = adam_component(p.x, g.x, m0, v0, beta1, beta2, step, bc2, eps);
    p.y = adam_component(p.y, g.y, m1, v1, beta1, beta2, step, bc2, eps);
    param[i] = p;
    m[2 * i] = m0;
    m[2 * i + 1] = m1;
    v[2 * i] = v0;
    v[2 * i + 1] = v1;
  }
};

// out[i] (+)= sum_{c < nparts} part[c*n + i], summed in ascending c (deterministic)
// conj: out (+)= conj(sum of the parts).  One 256-thread work-group per 256 outputs.
struct SumParts2 {
  const float2* part;
  int64_t nparts, n;
  float2* out;
  int accumulate, conj;
  AdamArgs adam{};  // adam.param set: the sum is the gradient of an Adam update (out unused)
  __host__ __device__ int64_t blocks() const { return n > 0 ? (n + 255) / 256 : 0; }
  __device__ __forceinline__ void run(int64_t blk) const {
    const int64_t i = blk * 256 + threadIdx.x;
    if (i >= n) return;
    // eight slices' loads issued before their additions, which stay in ascending c
    auto sum_into = [&](float2 s) {
      int64_t c = 0;
      for (; c + 8 <= nparts; c += 8) {
        float2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = part[(c + u) * n + i];
#pragma unroll
        for (int u = 0; u < 8; ++u) s = c_add(s, v[u]);
      }
      for (; c < nparts; ++c) s = c_add(s, part[c * n + i]);
      return s;
    };
    if (adam.param) {
      adam.apply(i, sum_into(c_zero()));
    } else if (!conj) {
      out[i] = sum_into(accumulate ? out[i] : c_zero());
    } else {
      float2 t = sum_into(c_zero());
      t.y = -t.y;
      out[i] = accumulate ? c_add(out[i], t) : t;
    }
  }
};
static __global__ void __launch_bounds__(256) k_sum_parts2(SumParts2 w) { w.run(blockIdx.x); }
inline void sum_parts2(const float2* part, int64_t nparts, int64_t n, float2* out, int accumulate, hipStream_t s,
                       int conj = 0) {
  const SumParts2 w{part, nparts, n, out, accumulate, conj};
  if (w.blocks() > 0) k_sum_parts2<<<dim3((unsigned)w.blocks()), dim3(256), 0, s>>>(w);
}


// Sum of x over the work-group (blockDim.x a power of two <= 1024); result valid in
// every thread.  Shuffle offsets stay below blockDim.x so that work-groups smaller
// than a wave never read inactive lanes.
__device__ __forceinline__ float block_sum(float x, float* red /* >= 16 floats of LDS */) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
    if (off < (int)blockDim.x) x += __shfl_xor(x, off, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = (blockDim.x + 63) >> 6;
  if (lane == 0) red[wave] = x;
  lds_barrier();
  float s = 0.0f;
  for (int w = 0; w < nwaves; ++w) s += red[w];  // fixed order
  lds_barrier();
  return s;
}

// out[b] = scale * sum_{k < nparts} part[b*nparts + k]: one work-group per b,
// each thread summing parts k = thread, thread + 256, ... in ascending order,
// then the fixed-order work-group sum (deterministic)
struct LossRows {
  const float* part;
  int64_t nparts;
  float scale;
  float* out;
  int64_t batch;  // 0: nothing to do
  __host__ __device__ int64_t blocks() const { return batch; }
  __device__ __forceinline__ void run(int64_t b, float* red /* 16 floats of LDS */) const {
    float s = 0.0f;
    for (int64_t k = threadIdx.x; k < nparts; k += 256) s += part[b * nparts + k];
    const float tot = block_sum(s, red);
    if (threadIdx.x == 0) out[b] = tot * scale;
  }
};
static __global__ void __launch_bounds__(256) k_loss_rows(LossRows w) {
  __shared__ float red[16];
  w.run(blockIdx.x, red);
}
inline void loss_rows(const float* part, int64_t nparts, int64_t batch, float scale, float* out, hipStream_t s) {
  const LossRows w{part, nparts, scale, out, batch};
  if (w.blocks() > 0) k_loss_rows<<<dim3((unsigned)w.blocks()), dim3(256), 0, s>>>(w);
}

// ---------------------------------------------------------------- sorted-segment scatter

// Gather tiles are 128 object columns wide.  Their height depends on the
// gather's mode: 16 rows when it is segmented (fewer, larger partial tiles to
// combine); 4 rows in the single-pass mode, where a work-group's time is its
// tile's list length, so small tiles spread a densely overlapped band of the
// object over more work-groups (a mini-batch of consecutive raster positions
// covers such a band: 40.0 -> 34.0 us for object 2048^2, probe 512^2, 64 views).
constexpr int kTileW = 128;
constexpr int kSegTileH = 16;
constexpr int kFlatTileH = 4;

__device__ __forceinline__ int64_t pmod(int64_t a, int64_t n) {
  const int64_t r = a % n;
  return r < 0 ? r + n : r;
}

// Windows are h x w rectangles whose top-left corner is
//   ((y_b + off_y) mod H, (x_b + off_x) mod W)
// on an H x W grid, wrapping around its edges: off = 0 gives the conventional
// scan window (never wraps for valid positions), off = (H/2, W/2) the Fourier
// ptychography sub-spectrum of python/ptychoX/fpm/ops.py:113-124.

// For every grid tile, the views whose window overlaps it, in ascending view
// order.  lists: [ntiles, batch] int32; counts: [ntiles].  One 256-thread
// work-group per tile.  With inside_only (scan windows, which must not wrap)
// a view whose window leaves the grid is listed nowhere: the scatter writes
// no contribution for it, so its rows of the contribution stack hold stale
// data that the gather must never read.
struct TileLists {
  const int64_t* pos;
  int64_t batch;
  int h, w;
  int64_t H, W, off_y, off_x, tiles_x;
  int tile_h;  // tile rows
  int* lists;
  int* counts;
  int64_t ntiles;  // 0: nothing to do
  int inside_only;
  __host__ __device__ int64_t blocks() const { return ntiles; }
  // any work-group size that is a multiple of 64, up to 1024
  __device__ __forceinline__ void run(int64_t tile, int* wave_tot /* blockDim.x / 64 ints of LDS */) const {
    const int64_t ty0 = (tile / tiles_x) * tile_h, tx0 = (tile % tiles_x) * kTileW;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nwaves = (int)(blockDim.x >> 6);
    int count = 0;
    for (int64_t base = 0; base < batch; base += blockDim.x) {
      const int64_t b = base + threadIdx.x;
      bool hit = false;
      if (b < batch) {
        const int64_t px = pos[2 * b], py = pos[2 * b + 1];
        const int64_t oy = pmod(py + off_y, H), ox = pmod(px + off_x, W);
        // two arcs on a circle overlap iff one starts inside the other
        const bool ry = pmod(ty0 - oy, H) < h || pmod(oy - ty0, H) < tile_h;
        const bool rx = pmod(tx0 - ox, W) < w || pmod(ox - tx0, W) < kTileW;
        hit = ry && rx && (!inside_only || window_ok(px, py, h, w, H, W));
      }
      const unsigned long long m = __ballot(hit);
      const int before = __popcll(m & ((1ull << lane) - 1ull));
      if (lane == 0) wave_tot[wave] = __popcll(m);
      __syncthreads();
      int off = count;
      for (int k = 0; k < wave; ++k) off += wave_tot[k];
      if (hit) lists[tile * batch + off + before] = (int)b;
      for (int k = 0; k < nwaves; ++k) count += wave_tot[k];
      __syncthreads();
    }
    if (threadIdx.x == 0) counts[tile] = count;
  }
  // The same list built by one wave alone (no LDS, no work-group barrier), so
  // each wave of a work-group can build the list of a different tile.
  __device__ __forceinline__ void run_wave(int64_t tile) const {
    const int64_t ty0 = (tile / tiles_x) * tile_h, tx0 = (tile % tiles_x) * kTileW;
    const int lane = threadIdx.x & 63;
    int count = 0;
    for (int64_t base = 0; base < batch; base += 64) {
      const int64_t b = base + lane;
      bool hit = false;
      if (b < batch) {
        const int64_t px = pos[2 * b], py = pos[2 * b + 1];
        const int64_t oy = pmod(py + off_y, H), ox = pmod(px + off_x, W);
        const bool ry = pmod(ty0 - oy, H) < h || pmod(oy - ty0, H) < tile_h;
        const bool rx = pmod(tx0 - ox, W) < w || pmod(ox - tx0, W) < kTileW;
        hit = ry && rx && (!inside_only || window_ok(px, py, h, w, H, W));
      }
      const unsigned long long m = __ballot(hit);
      if (hit) lists[tile * batch + count + __popcll(m & ((1ull << lane) - 1ull))] = (int)b;
      count += __popcll(m);
    }
    if (lane == 0) counts[tile] = count;
  }
};
static __global__ void __launch_bounds__(256) k_tile_lists(TileLists w) {
  __shared__ int wave_tot[4];
  w.run(blockIdx.x, wave_tot);
}

// Independent small jobs that ride along as extra work-groups of a larger
// 256-thread launch (the work-groups after the launch's own): per-tile view
// lists, per-view loss sums, probe-partial sums.  Each work-group does exactly
// what the stand-alone kernel's work-group would, so results are unchanged;
// the launches disappear (neutral on the conventional step, +3 % on the
// near-field step).  Placed before the launch's own work-groups they measured
// slower: they hold the big kernel's LDS allocation and delay its work.
struct SideWork {
  TileLists lists{};
  LossRows loss{};
  SumParts2 sums{};
  __host__ __device__ int64_t blocks() const { return lists.blocks() + loss.blocks() + sums.blocks(); }
  __device__ __forceinline__ void run(int64_t blk) const {
    __shared__ int wave_tot[4];
    __shared__ float red[16];
    if (blk < lists.blocks()) {
      lists.run(blk, wave_tot);
      return;
    }
    blk -= lists.blocks();
    if (blk < loss.blocks()) {
      loss.run(blk, red);
      return;
    }
    blk -= loss.blocks();
    sums.run(blk);
  }
};

// grad[Y, X] (+)= sum over listed views b (ascending) of contrib[b, cy, cx] with
// (cy, cx) = ((Y - oy_b) mod H, (X - ox_b) mod W) inside the h x w window.
//
// Load balance: a tile's ascending view list is cut into segments of
// kSegViews views, one work-group per segment (the plan below enumerates
// them), so heavily overlapped tiles -- the Fourier spectrum near DC, the
// near-field object under the probe -- spread over many work-groups.  Segment
// 0 writes the tile's output, segments 1.. write partial tiles, and the
// combine pass adds those to the output in segment order: a fixed summation
// order, independent of scheduling.
constexpr int kSegViews = 8;
// Views gathered per iteration: their 8-row loads are all in flight before the
// ordered additions.
constexpr int kGatherUnroll = 4;
constexpr int kSegTilePx = kSegTileH * kTileW;

// Work items and partial-tile slots a gather can need: every tile has at
// least one segment, and the segments beyond the first are bounded by the
// list entries divided by kSegViews.  A view's window meets at most h/16 + 2
// rows of the segmented gather's 16-row tiles (wrap-around on a grid of whole
// tiles included) and the list test can add the grid's last, partial tile
// row; likewise for columns, hence (h/16 + 3) x (w/128 + 3) list entries per
// view.
struct GatherBounds {
  int64_t items, slots;
};
inline GatherBounds gather_bounds(int64_t ntiles, int64_t h, int64_t w, int64_t batch) {
  int64_t per_view = (h / kSegTileH + 3) * (w / kTileW + 3);
  if (per_view > ntiles) per_view = ntiles;
  const int64_t slots = (batch * per_view + kSegViews - 1) / kSegViews;
  return GatherBounds{ntiles + slots, slots};
}

// One work-group: per tile nseg = max(1, ceil(count / kSegViews)); exclusive
// scans give each tile its first work item (seg_off) and its first partial
// slot (slot_off); item_tile[i] names the tile of work item i; *n_items the total.
static __global__ void __launch_bounds__(1024)
    k_gather_plan(const int* __restrict__ counts, int64_t ntiles, int* __restrict__ seg_off,
                  int* __restrict__ slot_off, int* __restrict__ item_tile, int* __restrict__ n_items) {
  __shared__ int s_seg[1024], s_slot[1024];
  int carry_seg = 0, carry_slot = 0;
  for (int64_t base = 0; base < ntiles; base += 1024) {
    const int64_t t = base + threadIdx.x;
    const int c = t < ntiles ? counts[t] : 0;
    const int nseg = t < ntiles ? (c > kSegViews ? (c + kSegViews - 1) / kSegViews : 1) : 0;
    s_seg[threadIdx.x] = nseg;
    s_slot[threadIdx.x] = nseg > 0 ? nseg - 1 : 0;
    __syncthreads();
    // inclusive Hillis-Steele scans of both counts
    for (int off = 1; off < 1024; off <<= 1) {
      const int a = threadIdx.x >= (unsigned)off ? s_seg[threadIdx.x - off] : 0;
      const int b = threadIdx.x >= (unsigned)off ? s_slot[threadIdx.x - off] : 0;
      __syncthreads();
      s_seg[threadIdx.x] += a;
      s_slot[threadIdx.x] += b;
      __syncthreads();
    }
    if (t < ntiles) {
      const int first = carry_seg + s_seg[threadIdx.x] - nseg;
      seg_off[t] = first;
      slot_off[t] = carry_slot + s_slot[threadIdx.x] - (nseg - 1);
      for (int k = 0; k < nseg; ++k) item_tile[first + k] = (int)t;
    }
    carry_seg += s_seg[1023];
    carry_slot += s_slot[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *n_items = carry_seg;
}

// SEG: one work item per segment (k_gather_plan's enumeration); otherwise one
// work item per tile over its whole list, staged through LDS kListChunk views
// at a time.
// ADAM: the tile's finished gradient feeds `adam` (a separate instantiation,
// so the plain gather carries none of the update's code).
constexpr int kListChunk = 64;
template <bool SEG, bool ADAM>
static __global__ void __launch_bounds__(256) k_tile_gather(const float2* __restrict__ contrib, int h, int w,
                                                            const int64_t* __restrict__ pos, int64_t batch,
                                                            int64_t off_y, int64_t off_x,
                                                            const int* __restrict__ lists,
                                                            const int* __restrict__ counts, int64_t tiles_x,
                                                            const int* __restrict__ seg_off,
                                                            const int* __restrict__ slot_off,
                                                            const int* __restrict__ item_tile,
                                                            const int* __restrict__ n_items,
                                                            float2* __restrict__ partial,
                                                            float2* __restrict__ grad, int64_t H, int64_t W,
                                                            int accumulate, SideWork side, int64_t main_blocks,
                                                            AdamArgs adam) {
  constexpr int TH = SEG ? kSegTileH : kFlatTileH;
  constexpr int RPT = TH * kTileW / 256;  // rows per thread
  constexpr int RSTEP = 256 / kTileW;
  constexpr int CHUNK = SEG ? kSegViews : kListChunk;
  if ((int64_t)blockIdx.x >= main_blocks) {  // work-groups riding along
    side.run((int64_t)blockIdx.x - main_blocks);
    return;
  }
  const unsigned blk = blockIdx.x;
  __shared__ int s_b[CHUNK];
  __shared__ int s_oy[CHUNK];
  __shared__ int s_ox[CHUNK];
  int64_t tile;
  int seg = 0;
  if (SEG) {
    const int item = (int)blk;
    if (item >= *n_items) return;  // the grid is the host's upper bound
    tile = item_tile[item];
    seg = item - seg_off[tile];
  } else {
    tile = blk;
  }
  const int64_t ty0 = (tile / tiles_x) * TH, tx0 = (tile % tiles_x) * kTileW;
  const int64_t X = tx0 + (threadIdx.x % kTileW);
  const int64_t Y0 = ty0 + threadIdx.x / kTileW;
  const bool xin = X < W;
  float2 acc[RPT];
#pragma unroll
  for (int j = 0; j < RPT; ++j) acc[j] = c_zero();
  const int n = counts[tile];
  const int vbeg = seg * kSegViews;
  const int vend = SEG ? (n - vbeg < kSegViews ? n : vbeg + kSegViews) : n;
  const int* list = lists + tile * batch;
  for (int c0 = vbeg; c0 < vend; c0 += CHUNK) {
    const int m = vend - c0 < CHUNK ? vend - c0 : CHUNK;
    if (c0 > vbeg) __syncthreads();  // the previous chunk is consumed
    if (threadIdx.x < (unsigned)m) {
      const int64_t b = list[c0 + threadIdx.x];
      s_b[threadIdx.x] = (int)b;
      s_oy[threadIdx.x] = (int)pmod(pos[2 * b + 1] + off_y, H);
      s_ox[threadIdx.x] = (int)pmod(pos[2 * b] + off_x, W);
    }
    __syncthreads();
    for (int k = 0; k < m; k += kGatherUnroll) {
      float2 tv_all[kGatherUnroll][RPT];
      bool uv_all[kGatherUnroll][RPT];
#pragma unroll
      for (int v = 0; v < kGatherUnroll; ++v) {
        float2* tv = tv_all[v];
        bool* uv = uv_all[v];
        const int kk = k + v;
        const bool have = kk < m;
        const int64_t b = have ? s_b[kk] : 0;
        int64_t cx = X - (have ? s_ox[kk] : 0);
        if (cx < 0) cx += W;
        int64_t cy = Y0 - (have ? s_oy[kk] : 0);
        if (cy < 0) cy += H;
        const bool colok = have && xin && cx < w;
        const float2* src = contrib + (b * h) * w + cx;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
          const int64_t Y = Y0 + j * RSTEP;
          uv[j] = colok && Y < H && cy < h;
          tv[j] = uv[j] ? src[cy * w] : c_zero();
          cy += RSTEP;
          if (cy >= H) cy -= H;
        }
      }
#pragma unroll
      for (int v = 0; v < kGatherUnroll; ++v)
#pragma unroll
        for (int j = 0; j < RPT; ++j)
          if (uv_all[v][j]) acc[j] = c_add(acc[j], tv_all[v][j]);  // ascending view order
    }
  }
  if (SEG && seg > 0) {  // a partial tile, added by the combine pass
    float2* pt = partial + (int64_t)(slot_off[tile] + seg - 1) * kSegTilePx;
#pragma unroll
    for (int j = 0; j < RPT; ++j) pt[j * 256 + threadIdx.x] = acc[j];
    return;
  }
  if (!xin) return;
  // the tile's final gradient (single pass, or a segmented tile with one
  // segment) feeds the fused Adam update; otherwise it is stored (and the
  // combine pass finishes segmented tiles)
  const bool final_here = !SEG || n <= kSegViews;
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
    const int64_t Y = Y0 + j * RSTEP;
    if (Y < H) {
      if (ADAM && final_here) {
        adam.apply(Y * W + X, acc[j]);
      } else {
        float2* d = grad + Y * W + X;
        *d = accumulate ? c_add(*d, acc[j]) : acc[j];
      }
    }
  }
}

// grad += the partial tiles of segments 1.. of every tile, in segment order.
// Work-group (tile, j) adds pixel chunk j (256 pixels) of the tile, eight
// partial tiles' loads in flight per step.
constexpr int kCombineBatch = 8;
static __global__ void __launch_bounds__(256) k_gather_combine(const int* __restrict__ counts,
                                                               const int* __restrict__ slot_off, int64_t tiles_x,
                                                               const float2* __restrict__ partial,
                                                               float2* __restrict__ grad, int64_t H, int64_t W,
                                                               AdamArgs adam) {
  constexpr int RPT = kSegTilePx / 256;
  constexpr int RSTEP = 256 / kTileW;
  const int64_t tile = blockIdx.x / RPT;
  const int j = blockIdx.x % RPT;
  const int n = counts[tile];
  const int nseg = n > kSegViews ? (n + kSegViews - 1) / kSegViews : 1;
  if (nseg < 2) return;
  const int64_t ty0 = (tile / tiles_x) * kSegTileH, tx0 = (tile % tiles_x) * kTileW;
  const int64_t X = tx0 + (threadIdx.x % kTileW);
  const int64_t Y = ty0 + threadIdx.x / kTileW + j * RSTEP;
  if (X >= W || Y >= H) return;
  const float2* pt = partial + (int64_t)slot_off[tile] * kSegTilePx + j * 256 + threadIdx.x;
  float2 s = grad[Y * W + X];
  for (int sg = 0; sg < nseg - 1; sg += kCombineBatch) {
    float2 v[kCombineBatch];
#pragma unroll
    for (int k = 0; k < kCombineBatch; ++k) v[k] = sg + k < nseg - 1 ? pt[(int64_t)(sg + k) * kSegTilePx] : c_zero();
#pragma unroll
    for (int k = 0; k < kCombineBatch; ++k)
      if (sg + k < nseg - 1) s = c_add(s, v[k]);
  }
  if (adam.param)
    adam.apply(Y * W + X, s);
  else
    grad[Y * W + X] = s;
}

struct TileGrid {
  int64_t tiles_x, tiles_y;
  int tile_h;
  int64_t count() const { return tiles_x * tiles_y; }
};

// ---------------------------------------------------------------- workspace carving

struct Carver {
  char* base;
  size_t used;
  template <class T>
  T* take(size_t count) {
    used = (used + 255) & ~size_t(255);
    T* p = base ? reinterpret_cast<T*>(base + used) : nullptr;
    used += count * sizeof(T);
    return p;
  }
};

// Segments pay for the plan and combine passes once the windows overlap by
// at least a segment's worth of views on average.
inline bool gather_segmented(int64_t h, int64_t w, int64_t B, int64_t H, int64_t W) {
  return B * h * w >= (int64_t)kSegViews * H * W;
}
inline TileGrid tile_grid(int64_t h, int64_t w, int64_t B, int64_t H, int64_t W) {
  const int th = gather_segmented(h, w, B, H, W) ? kSegTileH : kFlatTileH;
  return TileGrid{(W + kTileW - 1) / kTileW, (H + th - 1) / th, th};
}

// Scratch of the deterministic gradient gather: per-tile view lists, the
// segment plan and the partial tiles of segments beyond the first.
struct GatherScratch {
  int* lists;      // [ntiles, B]
  int* counts;     // [ntiles]
  int* seg_off;    // [ntiles]
  int* slot_off;   // [ntiles]
  int* item_tile;  // [bound on work items]
  int* n_items;    // [1]
  float2* partial; // [bound on partial slots, kSegTilePx]
};

inline GatherScratch carve_gather(Carver& cv, int64_t H, int64_t W, int64_t h, int64_t w, int64_t B) {
  const TileGrid tg = tile_grid(h, w, B, H, W);
  const GatherBounds gb = gather_bounds(tg.count(), h, w, B);
  GatherScratch g{};
  g.lists = cv.take<int>((size_t)(tg.count() * B));
  g.counts = cv.take<int>((size_t)tg.count());
  if (!gather_segmented(h, w, B, H, W)) return g;
  g.seg_off = cv.take<int>((size_t)tg.count());
  g.slot_off = cv.take<int>((size_t)tg.count());
  g.item_tile = cv.take<int>((size_t)gb.items);
  g.n_items = cv.take<int>(1);
  g.partial = cv.take<float2>((size_t)(gb.slots * kSegTilePx));
  return g;
}

// grad (+)= the windows' contributions (contrib [B, h, w], window corners
// pos + (off_x, off_y), wrapping modulo (H, W)): tile lists, segment plan,
// segmented gather, combine.
// The list job of a gather (to run as side work of an earlier launch).
// inside_only: the windows are scan windows (no wrap-around); a view whose
// window leaves the grid contributes nothing.
inline TileLists gather_lists_work(const GatherScratch& g, const int64_t* pos, int64_t B, int64_t h, int64_t w,
                                   int64_t H, int64_t W, int64_t off_y, int64_t off_x, bool inside_only) {
  const TileGrid tg = tile_grid(h, w, B, H, W);
  return TileLists{pos,           B,        (int)h,   (int)w,    H,          W,
                   off_y,         off_x,    tg.tiles_x, tg.tile_h, g.lists,    g.counts,
                   tg.count(),    inside_only ? 1 : 0};
}

// grad (+)= the windows' contributions, from the tile lists an earlier
// launch built (gather_lists_work as its side work): the single-pass gather,
// or plan, segmented gather and combine.  `side` rides along with the gather
// launch.
// adam.param set: the object gradient feeds that Adam update instead of
// being stored (grad is then scratch of segmented tiles; no accumulate).
inline ptx_status run_tile_gather(const GatherScratch& g, const float2* contrib, int64_t h, int64_t w,
                                  const int64_t* pos, int64_t B, int64_t off_y, int64_t off_x, float2* grad,
                                  int64_t H, int64_t W, int accumulate, hipStream_t s,
                                  const SideWork& side = SideWork{}, const AdamArgs& adam = AdamArgs{}) {
  const TileGrid tg = tile_grid(h, w, B, H, W);
  ptx_status st;
  if (!gather_segmented(h, w, B, H, W)) {
    const dim3 grid((unsigned)(tg.count() + side.blocks()));
    if (adam.param)
      k_tile_gather<false, true><<<grid, dim3(256), 0, s>>>(
          contrib, (int)h, (int)w, pos, B, off_y, off_x, g.lists, g.counts, tg.tiles_x, nullptr, nullptr, nullptr,
          nullptr, nullptr, grad, H, W, accumulate, side, tg.count(), adam);
    else
      k_tile_gather<false, false><<<grid, dim3(256), 0, s>>>(
          contrib, (int)h, (int)w, pos, B, off_y, off_x, g.lists, g.counts, tg.tiles_x, nullptr, nullptr, nullptr,
          nullptr, nullptr, grad, H, W, accumulate, side, tg.count(), adam);
    return check_launch("tile gather", s);
  }
  const GatherBounds gb = gather_bounds(tg.count(), h, w, B);
  k_gather_plan<<<dim3(1), dim3(1024), 0, s>>>(g.counts, tg.count(), g.seg_off, g.slot_off, g.item_tile, g.n_items);
  if ((st = check_launch("gather plan", s)) != PTX_OK) return st;
  const dim3 grid((unsigned)(gb.items + side.blocks()));
  if (adam.param)
    k_tile_gather<true, true><<<grid, dim3(256), 0, s>>>(
        contrib, (int)h, (int)w, pos, B, off_y, off_x, g.lists, g.counts, tg.tiles_x, g.seg_off, g.slot_off,
        g.item_tile, g.n_items, g.partial, grad, H, W, accumulate, side, gb.items, adam);
  else
    k_tile_gather<true, false><<<grid, dim3(256), 0, s>>>(
        contrib, (int)h, (int)w, pos, B, off_y, off_x, g.lists, g.counts, tg.tiles_x, g.seg_off, g.slot_off,
        g.item_tile, g.n_items, g.partial, grad, H, W, accumulate, side, gb.items, adam);
  if ((st = check_launch("tile gather", s)) != PTX_OK) return st;
  if (gb.slots > 0) {
    k_gather_combine<<<dim3((unsigned)(tg.count() * (kSegTilePx / 256))), dim3(256), 0, s>>>(g.counts, g.slot_off,
                                                                                          tg.tiles_x, g.partial,
                                                                                          grad, H, W, adam);
    if ((st = check_launch("gather combine", s)) != PTX_OK) return st;
  }
  return PTX_OK;
}

}  // namespace ptx
