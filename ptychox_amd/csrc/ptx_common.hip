// Error reporting, version queries and the standalone batched 2-D FFT entry point.
#include <atomic>
#include <cstdint>
#include <cstdarg>
#include <cstdio>

#include "passes.hpp"
#include "ptx_internal.hpp"

namespace ptx {

static thread_local char g_last_error[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------- tuning table
static std::atomic<int> g_tune[PASS_KINDS][16];  // 0 = built-in default

static int default_elems(int /*kind*/, int64_t n) {
  if (n <= 8) return 8;
  if (n == 16) return 16;
  return 8;  // every pass kind (PASS_VIEW_ROWS needs it: five register arrays per thread)
}

static std::atomic<int> g_groups[PASS_KINDS][16];  // 0 = built-in default

// per-thread snapshot of both tables, live while a TuneScope is open
static thread_local int t_snap_elems[PASS_KINDS][16];
static thread_local int t_snap_groups[PASS_KINDS][16];
static thread_local int t_scope_depth = 0;

TuneScope::TuneScope() {
  if (t_scope_depth++ > 0) return;
  for (int k = 0; k < PASS_KINDS; ++k)
    for (int l = 0; l < 16; ++l) {
      t_snap_elems[k][l] = g_tune[k][l].load(std::memory_order_relaxed);
      t_snap_groups[k][l] = g_groups[k][l].load(std::memory_order_relaxed);
    }
}
TuneScope::~TuneScope() { --t_scope_depth; }

int tuned_groups(int kind, int64_t n, int def) {
  if (kind < 0 || kind >= PASS_KINDS || !is_pow2(n)) return def;
  const int l = ilog2((int)n);
  if (l < 0 || l >= 16) return def;
  const int v = t_scope_depth > 0 ? t_snap_groups[kind][l] : g_groups[kind][l].load(std::memory_order_relaxed);
  return v ? v : def;
}

int tuned_elems(int kind, int64_t n) {
  if (kind < 0 || kind >= PASS_KINDS || !is_pow2(n)) return default_elems(kind, n);
  const int l = ilog2((int)n);
  if (l < 0 || l >= 16) return default_elems(kind, n);
  const int v = t_scope_depth > 0 ? t_snap_elems[kind][l] : g_tune[kind][l].load(std::memory_order_relaxed);
  return v ? v : default_elems(kind, n);
}

static thread_local hipEvent_t* t_events = nullptr;
static thread_local int t_nevents = 0;
static thread_local int t_used = 0;
static thread_local const char* t_labels[256];

static void timing_record(hipStream_t stream, const char* label) {
  if (t_used < t_nevents && t_used < 256) {
    hipEventRecord(t_events[t_used], stream);
    t_labels[t_used] = label;
    ++t_used;
  }
}

void timing_begin(hipStream_t stream) {
  if (t_nevents > 0) timing_record(stream, "begin");
}

ptx_status check_launch(const char* what, hipStream_t stream) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("%s: HIP launch failed: %s", what, hipGetErrorString(e));
    return PTX_ERR_HIP;
  }
  if (t_nevents > 0) timing_record(stream, what);
  return PTX_OK;
}

int device_cu_count() {
  constexpr int kMaxDevices = 64;
  static std::atomic<int> cache[kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 256;
  int n = cache[dev].load(std::memory_order_relaxed);
  if (n > 0) return n;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
  cache[dev].store(n, std::memory_order_relaxed);
  return n;
}

// ---------------------------------------------------------------- streaming copy
// Grid-stride copy of n 16-byte elements: thread t of chunk c moves elements
// c * 1024 + t + 256 u (u < 4), all four loads issued before the stores; the
// chunk where a thread's last element passes n is finished with bounds checks.
__global__ void __launch_bounds__(256) k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b,
                                                     size_t n) {
  constexpr int U = 4;
  const size_t stride = (size_t)gridDim.x * 256 * U;
  size_t i = (size_t)blockIdx.x * 256 * U + threadIdx.x;
  for (; i + 256 * (U - 1) < n; i += stride) {
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) v[u] = a[i + 256 * u];
#pragma unroll
    for (int u = 0; u < U; ++u) b[i + 256 * u] = v[u];
  }
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (i + 256 * u < n) b[i + 256 * u] = a[i + 256 * u];
}

template <int N, int E>
struct FftRowsRun {
  static ptx_status run(float2* data, int64_t nrows, int inverse, hipStream_t s) {
    RowsInPlace op{data, 1.0f};
    if (inverse)
      launch_rows<N, E, +1>(op, nrows, s);
    else
      launch_rows<N, E, -1>(op, nrows, s);
    return check_launch("fft2d rows", s);
  }
};

template <int N, int E>
struct FftColsRun {
  static ptx_status run(float2* data, int64_t frames, int64_t width, int inverse, hipStream_t s) {
    ColsInPlace op{data, width, 1.0f};
    if (inverse)
      launch_cols<N, E, +1>(op, frames, width, s);
    else
      launch_cols<N, E, -1>(op, frames, width, s);
    return check_launch("fft2d cols", s);
  }
};

}  // namespace ptx

using namespace ptx;

extern "C" {

const char* ptx_last_error(void) { return g_last_error; }

ptx_status ptx_timing_events(void* const* events, int n) {
  if (n < 0 || (n > 0 && events == nullptr)) {
    set_error("ptx_timing_events: invalid event array");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  t_events = n > 0 ? (hipEvent_t*)events : nullptr;
  t_nevents = n > 256 ? 256 : n;
  t_used = 0;
  return PTX_OK;
}

int ptx_timing_count(void) { return t_used; }

ptx_status ptx_tune_set(int kind, int64_t n, int elems) {
  if (kind < 0 || kind >= PASS_KINDS) {
    set_error("ptx_tune_set: pass kind %d out of range [0, %d)", kind, (int)PASS_KINDS);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (!size_ok(n)) {
    set_error("ptx_tune_set: length %lld unsupported", (long long)n);
    return PTX_ERR_UNSUPPORTED;
  }
  if (elems != 0 && elems != 8 && elems != 16) {
    set_error("ptx_tune_set: elements per thread must be 8 or 16 (0 = default), got %d", elems);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (n == 8 && elems == 16) {
    set_error("ptx_tune_set: length 8 takes 8 elements per thread");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  g_tune[kind][ilog2((int)n)].store(elems, std::memory_order_relaxed);
  return PTX_OK;
}

int ptx_tune_get(int kind, int64_t n) { return size_ok(n) ? tuned_elems(kind, n) : 0; }

ptx_status ptx_tune_set_groups(int kind, int64_t n, int groups) {
  if (kind < 0 || kind >= PASS_KINDS) {
    set_error("ptx_tune_set_groups: pass kind %d out of range [0, %d)", kind, (int)PASS_KINDS);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (!size_ok(n)) {
    set_error("ptx_tune_set_groups: length %lld unsupported", (long long)n);
    return PTX_ERR_UNSUPPORTED;
  }
  if (groups < 0 || groups > 65536) {
    set_error("ptx_tune_set_groups: work-group target must be in [1, 65536] (0 = default), got %d", groups);
    return PTX_ERR_INVALID_ARGUMENT;
  }
  g_groups[kind][ilog2((int)n)].store(groups, std::memory_order_relaxed);
  return PTX_OK;
}

int ptx_tune_get_groups(int kind, int64_t n) { return size_ok(n) ? tuned_groups(kind, n, 0) : 0; }

const char* ptx_timing_label(int k) { return (k >= 0 && k < t_used) ? t_labels[k] : ""; }

ptx_status ptx_stream_copy(const void* src, void* dst, size_t bytes, void* stream) {
  if (bytes == 0) return PTX_OK;
  if (!src || !dst || bytes % 16 != 0 || ((uintptr_t)src | (uintptr_t)dst) % 16 != 0) {
    set_error("ptx_stream_copy: pointers must be non-null and 16-byte aligned, bytes a multiple of 16");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  const uintptr_t s0 = (uintptr_t)src, d0 = (uintptr_t)dst;
  if (s0 < d0 + bytes && d0 < s0 + bytes) {
    set_error("ptx_stream_copy: source and destination overlap");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  const size_t n = bytes / 16;
  size_t grid = (n + 1023) / 1024;
  if (grid > 16384) grid = 16384;
  hipStream_t s = (hipStream_t)stream;
  k_stream_copy<<<dim3((unsigned)grid), dim3(256), 0, s>>>((const float4*)src, (float4*)dst, n);
  return check_launch("stream copy", s);
}

const char* ptx_version(void) { return "ptychox_hip 0.1.0 (gfx950)"; }

int ptx_abi_version(void) { return PTX_ABI_VERSION; }

ptx_status ptx_fft2d(void* data, int64_t batch, int64_t h, int64_t w, int inverse, void* stream) {
  TuneScope tune_scope;
  g_last_error[0] = 0;
  if (batch < 0 || (batch > 0 && data == nullptr)) {
    set_error("ptx_fft2d: invalid data pointer or batch");
    return PTX_ERR_INVALID_ARGUMENT;
  }
  if (!size_ok(h) || !size_ok(w)) {
    set_error("ptx_fft2d: frame %lld x %lld unsupported (powers of two in [%d, %d])", (long long)h, (long long)w,
              PTX_MIN_N, PTX_MAX_N);
    return PTX_ERR_UNSUPPORTED;
  }
  if (batch == 0) return PTX_OK;
  hipStream_t s = (hipStream_t)stream;
  timing_begin(s);
  float2* d = (float2*)data;
  ptx_status st = dispatch_ne<FftRowsRun>(PASS_ROW, w, d, batch * h, inverse, s);
  if (st != PTX_OK) return st;
  return dispatch_ne<FftColsRun>(PASS_COL, h, d, batch, w, inverse, s);
}

}  // extern "C"
