// Host-side plumbing shared by the C-ABI translation units: status/error
// reporting, size checks and the compile-time dispatch over power-of-two
// transform lengths.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#include "ptx.h"
#include "fft_core.hpp"

namespace ptx {

void set_error(const char* fmt, ...);

// sorted-segment gather unless atomics are requested (and not overridden)
inline bool use_segments(unsigned flags) {
  return (flags & PTX_FLAG_DETERMINISTIC) || !(flags & PTX_FLAG_ATOMIC_SCATTER);
}

inline bool is_pow2(int64_t n) { return n > 0 && (n & (n - 1)) == 0; }
inline bool size_ok(int64_t n) { return is_pow2(n) && n >= PTX_MIN_N && n <= PTX_MAX_N; }

// Returns PTX_ERR_HIP (with the message recorded) if the last launch failed.
// Also records the next armed timing event (ptx_timing_events) after the
// kernel(s) named `what`.
ptx_status check_launch(const char* what, hipStream_t stream = nullptr);

// positions_out[b] = positions_all[order[b]]; an order entry outside [0, count)
// reads entry 0 instead and sets *bad when bad is non-null (defined in ptx_optim.hip).
ptx_status launch_gather_positions(const int64_t* order, const int64_t* positions_all, int64_t count, int64_t batch,
                                   int64_t* positions_out, int* bad, hipStream_t stream);

// Compute units of the current device (cached per device; 256 on MI355X).
int device_cu_count();

// Work-groups of a persistent pass over `ntiles` tiles with `per_cu` resident
// work-groups per compute unit: a multiple of 16 when the grid is at least
// 16 wide (so the XCD pairing of adjacent tiles carries over to the tile loop).
inline int64_t persistent_grid(int64_t ntiles, int per_cu) {
  int64_t g = (int64_t)device_cu_count() * per_cu;
  if (g > ntiles) g = ntiles;
  if (g >= 16) g &= ~(int64_t)15;
  return g < 1 ? 1 : g;
}

// Records timing event 0 when armed; called by each entry point before its first kernel.
void timing_begin(hipStream_t stream);

// Pass kinds with their own elements-per-thread setting in the tuning table.
// PASS_VIEW_ROWS: row passes that loop over views holding per-element state
// of several frames in registers (the coded mask-gradient pass).
// PASS_LOSS_ROWS: row passes that evaluate the amplitude loss between an
// inverse and a forward transform (the coded step's sensor-plane pass).
enum PassKind {
  PASS_ROW = 0,
  PASS_ROW_LOOP = 1,
  PASS_COL = 2,
  PASS_COL_LOOP = 3,
  PASS_VIEW_ROWS = 4,
  PASS_LOSS_ROWS = 5,
  PASS_KINDS = 6
};

// Elements per thread (8 or 16) the tuning table holds for (kind, n); n a
// supported power of two.  Defaults come from MI355X measurements: radix-8
// passes run ~15 % faster than radix-16 for the plain row / column passes at
// 256..1024 points, and the view-looping scatter rows (PASS_ROW_LOOP) at
// radix 8 with kRowLoopGroups work-groups took 58 us against 91 us at radix 16
// with 1024 (2048^2 object, 512^2 probe, 64 views), the autotuner's choice
// there and for the 128- and 256-point scatters of mixed workloads.
int tuned_elems(int kind, int64_t n);

// Work-group target of the passes that walk views (the gathers, the
// sorted-segment scatters and the coded view loops): how many work-groups the
// views of a launch are spread over, which fixes the chunk of views each
// work-group walks.  `def` when the table holds no setting for (kind, n).
int tuned_groups(int kind, int64_t n, int def);

// Built-in work-group target of the PASS_ROW_LOOP scatters.
constexpr int kRowLoopGroups = 512;

// Consistent view of the tuning tables for the duration of one entry-point
// call: the first scope on a thread copies both process-wide tables into a
// thread-local snapshot that tuned_elems / tuned_groups read until it ends,
// so a ptx_tune_set* on another thread cannot change the schedule between a
// call's workspace layout and its launches.
struct TuneScope {
  TuneScope();
  ~TuneScope();
  TuneScope(const TuneScope&) = delete;
  TuneScope& operator=(const TuneScope&) = delete;
};

struct AdamArgs;
// The stand-alone Adam update of `adam` with gradient `grad` of n complex
// elements (defined in ptx_optim.hip).
ptx_status launch_adam(const AdamArgs& adam, const float2* grad, int64_t n, hipStream_t stream);

// Body of the conventional fused step (ptx_base.hip), shared by its entry
// points: ws_bytes of caller scratch at workspace (see ptx_base_step);
// adam_obj / adam_probe (may be null) fuse those Adam updates.
ptx_status base_step_impl(const char* fn, const void* sample, int64_t H, int64_t W, const void* probe, int64_t h,
                          int64_t w, const int64_t* pos, int64_t count, const float* measured, const int64_t* order,
                          int64_t B, float loss_scale, float* loss_per_position, float* amplitude, void* grad_sample,
                          void* grad_probe, void* workspace, size_t ws_bytes, unsigned flags, hipStream_t s,
                          const AdamArgs* adam_obj = nullptr, const AdamArgs* adam_probe = nullptr);

// Call F<N, E>::run(args...) for the runtime power-of-two n and the tuned E.
template <template <int, int> class F, typename... Args>
ptx_status dispatch_ne(int kind, int64_t n, Args&&... args) {
  const bool e16 = tuned_elems(kind, n) == 16;
  switch (n) {
    case 8: return F<8, 8>::run(args...);
    case 16: return e16 ? F<16, 16>::run(args...) : F<16, 8>::run(args...);
    case 32: return e16 ? F<32, 16>::run(args...) : F<32, 8>::run(args...);
    case 64: return e16 ? F<64, 16>::run(args...) : F<64, 8>::run(args...);
    case 128: return e16 ? F<128, 16>::run(args...) : F<128, 8>::run(args...);
    case 256: return e16 ? F<256, 16>::run(args...) : F<256, 8>::run(args...);
    case 512: return e16 ? F<512, 16>::run(args...) : F<512, 8>::run(args...);
    case 1024: return e16 ? F<1024, 16>::run(args...) : F<1024, 8>::run(args...);
    case 2048: return e16 ? F<2048, 16>::run(args...) : F<2048, 8>::run(args...);
    default:
      set_error("transform length %lld is not a power of two in [%d, %d]", (long long)n, PTX_MIN_N, PTX_MAX_N);
      return PTX_ERR_UNSUPPORTED;
  }
}

}  // namespace ptx
