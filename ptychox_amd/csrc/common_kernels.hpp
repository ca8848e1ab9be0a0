// Small helper kernels shared by the operators: fills, fixed-order partial
// reductions, block reductions, and the sorted-segment (deterministic)
// scatter-add of per-view windows into the object.
#pragma once

#include "fft_core.hpp"

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

// out[i] (+)= sum_{c < nparts} part[c*n + i], summed in ascending c (deterministic)
// conj: out (+)= conj(sum of the parts)
static __global__ void __launch_bounds__(256) k_sum_parts2(const float2* __restrict__ part, int64_t nparts, int64_t n,
                                                    float2* __restrict__ out, int accumulate, int conj) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
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
  if (!conj) {
    out[i] = sum_into(accumulate ? out[i] : c_zero());
  } else {
    float2 t = sum_into(c_zero());
    t.y = -t.y;
    out[i] = accumulate ? c_add(out[i], t) : t;
  }
}
inline void sum_parts2(const float2* part, int64_t nparts, int64_t n, float2* out, int accumulate, hipStream_t s,
                       int conj = 0) {
  if (n > 0)
    k_sum_parts2<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(part, nparts, n, out, accumulate, conj);
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
static __global__ void __launch_bounds__(256) k_loss_rows(const float* __restrict__ part, int64_t nparts,
                                                   float scale, float* __restrict__ out) {
  __shared__ float red[16];
  const int64_t b = blockIdx.x;
  float s = 0.0f;
  for (int64_t k = threadIdx.x; k < nparts; k += 256) s += part[b * nparts + k];
  const float tot = block_sum(s, red);
  if (threadIdx.x == 0) out[b] = tot * scale;
}
inline void loss_rows(const float* part, int64_t nparts, int64_t batch, float scale, float* out, hipStream_t s) {
  if (batch > 0) k_loss_rows<<<dim3((unsigned)batch), dim3(256), 0, s>>>(part, nparts, scale, out);
}

// ---------------------------------------------------------------- sorted-segment scatter

constexpr int kTileH = 16;   // object rows per gather tile
constexpr int kTileW = 128;  // object columns per gather tile

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
// order.  lists: [ntiles, batch] int32; counts: [ntiles].
static __global__ void __launch_bounds__(256) k_tile_lists(const int64_t* __restrict__ pos, int64_t batch, int h, int w,
                                                           int64_t H, int64_t W, int64_t off_y, int64_t off_x,
                                                           int64_t tiles_x, int* __restrict__ lists,
                                                           int* __restrict__ counts) {
  __shared__ int wave_tot[4];
  const int64_t tile = blockIdx.x;
  const int64_t ty0 = (tile / tiles_x) * kTileH, tx0 = (tile % tiles_x) * kTileW;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int count = 0;
  for (int64_t base = 0; base < batch; base += 256) {
    const int64_t b = base + threadIdx.x;
    bool hit = false;
    if (b < batch) {
      const int64_t oy = pmod(pos[2 * b + 1] + off_y, H), ox = pmod(pos[2 * b] + off_x, W);
      // two arcs on a circle overlap iff one starts inside the other
      const bool ry = pmod(ty0 - oy, H) < h || pmod(oy - ty0, H) < kTileH;
      const bool rx = pmod(tx0 - ox, W) < w || pmod(ox - tx0, W) < kTileW;
      hit = ry && rx;
    }
    const unsigned long long m = __ballot(hit);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = __popcll(m);
    __syncthreads();
    int off = count;
    for (int k = 0; k < wave; ++k) off += wave_tot[k];
    if (hit) lists[tile * batch + off + before] = (int)b;
    count += wave_tot[0] + wave_tot[1] + wave_tot[2] + wave_tot[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[tile] = count;
}

// grad[Y, X] (+)= sum over listed views b (ascending) of contrib[b, cy, cx] with
// (cy, cx) = ((Y - oy_b) mod H, (X - ox_b) mod W) inside the h x w window.
// The work-group first stages (b, oy_b, ox_b) of up to kViewsPerPass listed
// views in LDS, then walks them two at a time so that sixteen independent
// 8-byte loads per thread are in flight (the per-view latency was the limit).
constexpr int kViewsPerPass = 256;
// Views gathered per iteration: their 8-row loads are all in flight before the
// ordered additions.  Heavily overlapped tiles (the Fourier spectrum near DC,
// near-field windows) sum tens of views in one work-group, which this bounds.
constexpr int kGatherUnroll = 4;

static __global__ void __launch_bounds__(256) k_tile_gather(const float2* __restrict__ contrib, int h, int w,
                                                            const int64_t* __restrict__ pos, int64_t batch,
                                                            int64_t off_y, int64_t off_x,
                                                            const int* __restrict__ lists,
                                                            const int* __restrict__ counts, int64_t tiles_x,
                                                            float2* __restrict__ grad, int64_t H, int64_t W,
                                                            int accumulate) {
  constexpr int RPT = kTileH * kTileW / 256;  // rows per thread
  constexpr int RSTEP = 256 / kTileW;
  __shared__ int s_b[kViewsPerPass];
  __shared__ int s_oy[kViewsPerPass];
  __shared__ int s_ox[kViewsPerPass];
  const int64_t tile = blockIdx.x;
  const int64_t ty0 = (tile / tiles_x) * kTileH, tx0 = (tile % tiles_x) * kTileW;
  const int64_t X = tx0 + (threadIdx.x % kTileW);
  const int64_t Y0 = ty0 + threadIdx.x / kTileW;
  const bool xin = X < W;
  float2 acc[RPT];
#pragma unroll
  for (int j = 0; j < RPT; ++j) acc[j] = c_zero();
  const int n = counts[tile];
  const int* list = lists + tile * batch;
  for (int base = 0; base < n; base += kViewsPerPass) {
    const int m = n - base < kViewsPerPass ? n - base : kViewsPerPass;
    __syncthreads();  // previous pass finished reading the staged views
    for (int k = threadIdx.x; k < m; k += 256) {
      const int64_t b = list[base + k];
      s_b[k] = (int)b;
      s_oy[k] = (int)pmod(pos[2 * b + 1] + off_y, H);
      s_ox[k] = (int)pmod(pos[2 * b] + off_x, W);
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
  if (!xin) return;
#pragma unroll
  for (int j = 0; j < RPT; ++j) {
    const int64_t Y = Y0 + j * RSTEP;
    if (Y < H) {
      float2* d = grad + Y * W + X;
      *d = accumulate ? c_add(*d, acc[j]) : acc[j];
    }
  }
}

struct TileGrid {
  int64_t tiles_x, tiles_y;
  int64_t count() const { return tiles_x * tiles_y; }
};
inline TileGrid tile_grid(int64_t H, int64_t W) {
  return TileGrid{(W + kTileW - 1) / kTileW, (H + kTileH - 1) / kTileH};
}

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

}  // namespace ptx
