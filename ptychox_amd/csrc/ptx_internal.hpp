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

// Records timing event 0 when armed; called by each entry point before its first kernel.
void timing_begin(hipStream_t stream);

// Call F<N>::run(args...) for the runtime power-of-two n.
template <template <int> class F, typename... Args>
ptx_status dispatch_n(int64_t n, Args&&... args) {
  switch (n) {
    case 8: return F<8>::run(args...);
    case 16: return F<16>::run(args...);
    case 32: return F<32>::run(args...);
    case 64: return F<64>::run(args...);
    case 128: return F<128>::run(args...);
    case 256: return F<256>::run(args...);
    case 512: return F<512>::run(args...);
    case 1024: return F<1024>::run(args...);
    case 2048: return F<2048>::run(args...);
    default:
      set_error("transform length %lld is not a power of two in [%d, %d]", (long long)n, PTX_MIN_N, PTX_MAX_N);
      return PTX_ERR_UNSUPPORTED;
  }
}

}  // namespace ptx
