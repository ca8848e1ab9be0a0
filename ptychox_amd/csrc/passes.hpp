// Generic row / column FFT passes with pluggable prologue and epilogue.
//
// Row pass: a 256-thread work-group transforms S = 256/T rows of length N
// (T = N/E threads per row), each thread loading and storing elements
// t + q*T of its row -- fully coalesced along the contiguous axis.
//
// Column pass: a work-group of T*C threads transforms C adjacent columns of
// length N; lane layout c = tid % C, t = tid / C, so every load/store
// instruction touches C*8 contiguous bytes per row (128 B for C = 16).  The
// whole C x N tile lives in LDS (single buffer, padded per column).
//
// An "op" is a plain struct of device pointers and sizes passed by value as
// the kernel argument, with
//   rows:    template <int N, int E> load(row, valid, t, v) / store(row, valid, t, v)
//   columns: template <int N, int E> load(frame, col, t, v) / store(frame, col, t, v)
// where v[q] is element t + q*(N/E) of the sequence.
#pragma once

#include "buffer_io.hpp"
#include "fft_core.hpp"
#include "ptx_internal.hpp"

namespace ptx {

// E = complex elements per thread (the FFT radix); the runtime tuning table
// (ptx_internal.hpp) picks it per pass kind and length among compiled variants.
template <int N, int E = elems_for<N>()>
struct RowCfg {
  static_assert(E <= N && N % E == 0, "E must divide N");
  static constexpr int T = N / E;
  static constexpr int THREADS = 256;
  static constexpr int S = THREADS / T;  // rows per work-group
};

// LDS stride between the columns of a column tile (float2 units).  With C >= 16
// columns the 16 lanes that share an LDS cycle all hold one row index and the
// padded sequence stride lds_len (= 1 mod 32) spreads them over the banks.
// With C = 8 (C = 4) two (four) row indices share those lanes, and an odd
// stride puts column c + 1 of row t on the banks of column c of row t + 1;
// an even stride (+1; +3 for C = 4) separates them -- from a bank model of
// every exchange of the column transforms, confirmed by SQ_LDS_BANK_CONFLICT.
template <int N, int C>
__host__ __device__ constexpr int col_stride() {
  return C >= 16 ? lds_len<N>() : C == 4 ? lds_len<N>() + 3 : lds_len<N>() + 1;
}

template <int N, int E = elems_for<N>()>
struct ColCfg {
  static_assert(E <= N && N % E == 0, "E must divide N");
  static constexpr int T = N / E;
  // LDS budget per work-group ~72 KB (two work-groups per CU), <= 1024 threads
  static constexpr int BY_LDS = 73728 / (8 * (lds_len<N>() + 3));
  static constexpr int BY_THREADS = 1024 / T;
  static constexpr int CM0 = BY_LDS < BY_THREADS ? BY_LDS : BY_THREADS;
  static constexpr int CMAX = CM0 >= 16 ? 16 : CM0 >= 8 ? 8 : CM0 >= 4 ? 4 : 2;
};

// Block -> tile order of the column passes.  A column tile of C columns reads
// C*8 bytes of each complex row (C*4 of a float row), which is half of a
// 128-byte line when C = 8 (complex) or C = 16 (float measurements); the other
// half belongs to the neighbouring tile.  Work-groups are dispatched to the 8
// XCDs round-robin (block b runs on XCD b % 8, each XCD with its own L2), so in
// every group of 16 blocks, blocks r and r + 8 -- same XCD, dispatched back to
// back -- take the adjacent tiles 2r' and 2r' + 1 (r' = r mod 8) and share the
// line through that L2 instead of both fetching it from HBM.  A bijection on
// each full group of 16; a grid that is not a multiple of 16 keeps the identity.
__device__ __forceinline__ int64_t xcd_paired_block(unsigned b, unsigned nblocks) {
  if (nblocks & 15u) return b;
  const unsigned r = b & 15u;
  return (int64_t)((b & ~15u) + 2u * (r & 7u) + (r >> 3));
}

// Frame and column tile of column-pass block blk.  Tiles per frame (frame
// width / columns per tile, both powers of two) is a power of two, so this is
// a shift and a mask instead of a 64-bit division.
struct TileIdx {
  int64_t frame, tile;
};
__device__ __forceinline__ TileIdx split_tile(int64_t blk, int64_t tiles_per_frame) {
  const int sh = __builtin_ctzll((unsigned long long)tiles_per_frame);
  return TileIdx{blk >> sh, blk & (tiles_per_frame - 1)};
}

template <int N, int E, int DIR, class Op>
__global__ void __launch_bounds__(256) k_rows(Op op, int64_t nrows) {
  using Cf = RowCfg<N, E>;
  constexpr int T = Cf::T, S = Cf::S;
  __shared__ float2 lds[S * lds_len<N>()];
  const int seq = threadIdx.x / T;
  const int t = threadIdx.x % T;
  const int64_t row = (int64_t)blockIdx.x * S + seq;
  const bool valid = row < nrows;
  float2 v[E];
  op.template load<N, E>(row, valid, t, v);
  fft_row<N, E, DIR>(v, t, &lds[seq * lds_len<N>()]);
  op.template store<N, E>(row, valid, t, v);
}

template <int N, int E, int C, int DIR, class Op>
__global__ void __launch_bounds__((ColCfg<N, E>::T * C)) k_cols(Op op, int64_t tiles_per_frame) {
  __shared__ float2 lds[C * col_stride<N, C>()];
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int64_t blk = xcd_paired_block(blockIdx.x, gridDim.x);
  const TileIdx ti = split_tile(blk, tiles_per_frame);
  const int64_t frame = ti.frame;
  const int64_t col = ti.tile * C + c;
  float2 v[E];
  op.template load<N, E>(frame, col, t, v);
  fft_seq<N, E, DIR>(v, t, &lds[c * col_stride<N, C>()]);
  op.template store<N, E>(frame, col, t, v);
}

// Launch helpers (grid sizes); return false if the launch configuration is impossible.
template <int N, int E, int DIR, class Op>
inline void launch_rows(const Op& op, int64_t nrows, hipStream_t stream) {
  using Cf = RowCfg<N, E>;
  const int64_t blocks = (nrows + Cf::S - 1) / Cf::S;
  k_rows<N, E, DIR, Op><<<dim3((unsigned)blocks), dim3(Cf::THREADS), 0, stream>>>(op, nrows);
}

template <int N, int E, int DIR, class Op>
inline void launch_cols(const Op& op, int64_t frames, int64_t width, hipStream_t stream) {
  using Cf = ColCfg<N, E>;
  if (width >= Cf::CMAX) {
    constexpr int C = Cf::CMAX;
    const int64_t tiles = width / C;
    k_cols<N, E, C, DIR, Op><<<dim3((unsigned)(frames * tiles)), dim3(Cf::T * C), 0, stream>>>(op, tiles);
  } else {
    // narrow frames (width 8 with CMAX 16): one tile of 8 columns per frame
    constexpr int C = 8;
    const int64_t tiles = width / C;
    k_cols<N, E, C, DIR, Op><<<dim3((unsigned)(frames * tiles)), dim3(Cf::T * C), 0, stream>>>(op, tiles);
  }
}

// ---------------------------------------------------------------- column-strip layout
// A [rows, width] frame stored as strips of 16 adjacent columns (8 when the
// width is 8), each strip row-major [rows][16]: element (y, x) at
// ((x/16) rows + y) 16 + x%16.  A strip row is one 128-byte line, so the row
// pass that writes the frame stores whole lines, and the column passes read
// an 8- or 16-column tile of it from one contiguous block (the two 8-column
// tiles of a strip run back to back on one XCD and share its lines in L2)
// instead of pieces of lines 8 KB apart, which HBM serves at a fraction of the
// streaming rate.
__host__ __device__ constexpr uint32_t strip_lg(int64_t width) { return width >= 16 ? 4u : 3u; }
__host__ __device__ __forceinline__ uint32_t strip_off(uint32_t y, uint32_t x, uint32_t rows, uint32_t lg) {
  return (((x >> lg) * rows + y) << lg) + (x & ((1u << lg) - 1u));
}

// Row-pass store of v[q] = element (y, t + q T) of one frame (N columns, `rows`
// rows) into its column-strip layout.  With T a multiple of the strip width
// the q steps are whole strips, a scalar offset: one address register per lane.
template <int N, int E>
__device__ __forceinline__ void store_row_strip(float2* frame, uint32_t rows, uint32_t y, int t,
                                                const float2 (&v)[E]) {
  constexpr int T = N / E;
  constexpr uint32_t lg = strip_lg(N);
  const auto r = buffer_rsrc(frame, (uint32_t)N * rows * 8u);
  if constexpr (T % (1 << lg) == 0) {
    const uint32_t base = strip_off(y, (uint32_t)t, rows, lg) * 8u;
#pragma unroll
    for (int q = 0; q < E; ++q) buf_store_f2(v[q], r, base, (uint32_t)q * (uint32_t)T * rows * 8u);
  } else {
#pragma unroll
    for (int q = 0; q < E; ++q) buf_store_f2(v[q], r, strip_off(y, (uint32_t)(t + q * T), rows, lg) * 8u, 0u);
  }
}

// ---------------------------------------------------------------- transfer-function multiply
// Applies F^-1_cols D F_cols to the columns of row-transformed frames: with the
// rows already (inverse-)transformed, this completes  F^-1 (D . F x)  up to the
// final row pass.  ``transfer`` is one [N, width] frame shared by every frame of
// the batch (CONJ multiplies by its conjugate, for the adjoint).
//
// Column-tile-major and software-pipelined: work-group w owns one column tile
// and a residue class of frames (f = group, group + ngroups, ...), so the
// tile's transfer values are loaded once and stay in registers -- no transfer
// traffic per frame (the 8 MB transfer of a 1024^2 frame does not fit an
// XCD's L2).  Two register sets alternate, one holding the frame being
// transformed and the other the loads of the next, in flight meanwhile.  A
// frame past the end gets a zero-length buffer (its loads return zero without
// touching memory), so every path through the loop issues the same memory
// instructions and the compiler's wait counts stay exact; the first frame is
// peeled for the same reason.  The exchanges alternate between two LDS
// regions with one barrier each; the inverse transform starts on the region
// after the forward's last, so consecutive exchanges never share a region and
// no other barrier is needed.
// STRIP: the frames are read from `src` in the column-strip layout and the
// results stored to `tb` in the plain layout (otherwise in place in `tb`).
// REG_TW: the kernel holds its stage twiddles in LoopTwiddles<N, E, REG_TW>
template <int N, int E, int C, bool REG_TW = true>
struct PipeColsShape {
  static constexpr int T = ColCfg<N, E>::T;
  static constexpr int THREADS = T * C;
  static constexpr int EXCHANGES = FftShape<N, E>::STAGES - 1;
  static constexpr size_t LDS_BYTES =
      (2 * (size_t)C * col_stride<N, C>() + LoopTwiddles<N, E, REG_TW>::TABLE) * sizeof(float2);
  static_assert(LDS_BYTES <= 160 * 1024, "LDS budget");
  // resident work-groups per CU: threads and LDS
  static constexpr int PER_CU = (1024 / THREADS) < (int)((160 * 1024) / LDS_BYTES) ? (1024 / THREADS)
                                                                                  : (int)((160 * 1024) / LDS_BYTES);
};

template <int N, int E, int C, bool CONJ, bool STRIP>
__global__ void __launch_bounds__((ColCfg<N, E>::T * C))
    k_transfer_cols(float2* tb, const float2* src, const float2* __restrict__ transfer, int64_t width,
                    int64_t tiles_per_frame, int64_t frames, int64_t ngroups) {
  using Sh = PipeColsShape<N, E, C>;
  constexpr int T = Sh::T;
  __shared__ float2 lds0[C * col_stride<N, C>()];
  __shared__ float2 lds1[C * col_stride<N, C>()];
  __shared__ float2 s_tw[LoopTwiddles<N, E>::TABLE];
  const int c = threadIdx.x % C;
  const int t = threadIdx.x / C;
  const int64_t fsz = (int64_t)N * width;               // elements per frame
  const uint32_t fbytes = (uint32_t)(fsz * 8);
  const uint32_t lane = (uint32_t)(t * (int)width + c);  // element offset inside a tile (row t, column c)
  const uint32_t rstep = (uint32_t)(T * width);          // elements from row t + q T to t + (q + 1) T
  const unsigned nf = (unsigned)frames, ng = (unsigned)ngroups;
  // adjacent column tiles on one XCD (they share 128-byte lines through its L2)
  const TileIdx ti = split_tile(xcd_paired_block(blockIdx.x, gridDim.x), tiles_per_frame);
  const uint32_t col0 = (uint32_t)(ti.tile * C);
  const unsigned group = (unsigned)ti.frame;  // < ngroups <= frames
  LoopTwiddles<N, E> tw;  // thread t holds the same rows of every frame
  tw.init(t, s_tw);
  float2 tr[E];
  {
    const auto rtr = buffer_rsrc(transfer, fbytes);
#pragma unroll
    for (int q = 0; q < E; ++q) tr[q] = buf_load_f2(rtr, lane * 8u, (q * rstep + col0) * 8u);
  }
  // strip layout: row t of column col0 + c, then T rows (T strip rows) further per q
  const uint32_t lg = strip_lg(width);
  const uint32_t sbase = strip_off(0u, col0, (uint32_t)N, lg);
  const uint32_t slane = strip_off((uint32_t)t, col0 + (uint32_t)c, (uint32_t)N, lg) - sbase;
  // rows q in [q0, q1) of frame f
  auto fetch = [&](unsigned f, float2 (&dv)[E], int q0, int q1) {
    const bool live = f < nf;
    const auto rt = buffer_rsrc((STRIP ? src : tb) + (int64_t)(live ? f : 0u) * fsz, live ? fbytes : 0u);
#pragma unroll
    for (int q = 0; q < E; ++q) {
      if (q < q0 || q >= q1) continue;
      if constexpr (STRIP)
        dv[q] = buf_load_f2(rt, slane * 8u, (sbase + ((uint32_t)(q * T) << lg)) * 8u);
      else
        dv[q] = buf_load_f2(rt, lane * 8u, (q * rstep + col0) * 8u);
    }
  };
  // frame f from v; the next frame's loads are issued in two halves, its
  // first E/2 rows before this call and the rest after this frame's forward
  // transform, so that waves do not stall on a full memory address queue at
  // the start of the transform
  auto process = [&](unsigned f, float2 (&v)[E], unsigned fnext, float2 (&vnext)[E]) {
    // the exchange addresses depend only on (t, c); recompute them per frame
    // from the thread index (one register kept live across the loop) instead
    // of letting them be hoisted out of the loop and held in registers
    unsigned tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int tt = (int)(tid / C), cc = (int)(tid % C);
    float2* l0 = &lds0[cc * col_stride<N, C>()];
    float2* l1 = &lds1[cc * col_stride<N, C>()];
    tw.template stages<-1>(v, tt, l0, l1);
    fetch(fnext, vnext, E / 2, E);
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = CONJ ? c_mulc(v[q], tr[q]) : c_mul(v[q], tr[q]);
    if constexpr (Sh::EXCHANGES % 2 == 1)
      tw.template stages<+1>(v, tt, l1, l0);
    else
      tw.template stages<+1>(v, tt, l0, l1);
    const auto rt = buffer_rsrc(tb + (int64_t)f * fsz, fbytes);
#pragma unroll
    for (int q = 0; q < E; ++q) buf_store_f2(v[q], rt, lane * 8u, (q * rstep + col0) * 8u);
  };
  float2 va[E], vb[E];
  unsigned f = group;
  fetch(f, va, 0, E);
  fetch(f + ng, vb, 0, E / 2);  // in flight during this frame's transforms
  process(f, va, f + ng, vb);
  f += ng;
  while (f < nf) {
    fetch(f + ng, va, 0, E / 2);
    process(f, vb, f + ng, va);
    f += ng;
    if (f >= nf) break;
    fetch(f + ng, vb, 0, E / 2);
    process(f, va, f + ng, vb);
    f += ng;
  }
}

// host launcher (check_launch is the caller's): tiles_per_frame x ngroups
// work-groups, ngroups = the frame groups that fill the resident slots (or the
// tuned work-group target of the column passes)
template <int N, int E, int C>
inline void launch_transfer_cols_c(float2* tb, const float2* src, const float2* transfer, int64_t frames,
                                   int64_t width, bool conj, hipStream_t s) {
  using Sh = PipeColsShape<N, E, C>;
  const int64_t tiles = width / C;
  // work-groups: the resident slots by default, or the tuned target of the column passes
  int64_t ngroups = (int64_t)tuned_groups(PASS_COL, N, device_cu_count() * Sh::PER_CU) / tiles;
  if (ngroups > frames) ngroups = frames;
  if (ngroups < 1) ngroups = 1;
  const dim3 grid((unsigned)(tiles * ngroups));
  const dim3 block(Sh::THREADS);
  if (src) {
    if (conj)
      k_transfer_cols<N, E, C, true, true><<<grid, block, 0, s>>>(tb, src, transfer, width, tiles, frames, ngroups);
    else
      k_transfer_cols<N, E, C, false, true><<<grid, block, 0, s>>>(tb, src, transfer, width, tiles, frames, ngroups);
  } else {
    if (conj)
      k_transfer_cols<N, E, C, true, false><<<grid, block, 0, s>>>(tb, tb, transfer, width, tiles, frames, ngroups);
    else
      k_transfer_cols<N, E, C, false, false><<<grid, block, 0, s>>>(tb, tb, transfer, width, tiles, frames, ngroups);
  }
}

// src == null: in place in tb; otherwise from src in the column-strip layout
// (width a multiple of 8) into tb
template <int N, int E>
inline void launch_transfer_cols(float2* tb, const float2* transfer, int64_t frames, int64_t width, bool conj,
                                 hipStream_t s, const float2* src = nullptr) {
  using Cf = ColCfg<N, E>;
  if constexpr (Cf::CMAX > 8) {
    if (width < Cf::CMAX) {  // narrow frames: 8 columns per tile
      launch_transfer_cols_c<N, E, 8>(tb, src, transfer, frames, width, conj, s);
      return;
    }
  }
  launch_transfer_cols_c<N, E, Cf::CMAX>(tb, src, transfer, frames, width, conj, s);
}

// ---------------------------------------------------------------- plain ops

// rows of a contiguous [nrows, N] complex array, in place, scale on store
struct RowsInPlace {
  float2* data;
  float scale;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const float2* p = data + row * N;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = valid ? p[t + q * T] : c_zero();
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t row, bool valid, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    if (!valid) return;
    float2* p = data + row * N;
#pragma unroll
    for (int q = 0; q < E; ++q) p[t + q * T] = c_scale(v[q], scale);
  }
};

// rows of a contiguous [nrows, N] complex array from `in` to `out`, scale on store
struct RowsOutOfPlace {
  const float2* in;
  float2* out;
  float scale;
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
    float2* p = out + row * N;
#pragma unroll
    for (int q = 0; q < E; ++q) p[t + q * T] = c_scale(v[q], scale);
  }
};

// columns of contiguous [frames, N, width] complex frames, in place, scale on store
struct ColsInPlace {
  float2* data;
  int64_t width;
  float scale;
  template <int N, int E>
  __device__ __forceinline__ void load(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    const float2* p = data + frame * N * width;
#pragma unroll
    for (int q = 0; q < E; ++q) v[q] = p[(t + q * T) * (int)width + (int)col];
  }
  template <int N, int E>
  __device__ __forceinline__ void store(int64_t frame, int64_t col, int t, float2 (&v)[E]) const {
    constexpr int T = N / E;
    float2* p = data + frame * N * width;
#pragma unroll
    for (int q = 0; q < E; ++q) p[(t + q * T) * (int)width + (int)col] = c_scale(v[q], scale);
  }
};

}  // namespace ptx
