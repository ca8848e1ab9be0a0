// Allocation-free variant of the conventional fused step, for callers that
// cannot allocate device memory at all (code under stream capture, tight
// serving loops, harnesses that forbid hipMalloc): its scratch space is a
// file-scope device array of a fixed maximum size instead of a caller
// workspace.  Each GPU holds its own copy of the array from the first call on
// that device until the process ends.
//
// The array's name is a host-side stand-in in host code; its device address
// is looked up once per device with hipGetSymbolAddress (with that device
// current) and only that address is ever handed to kernels.
#include <mutex>

#include "ptx_internal.hpp"

namespace ptx {
namespace {

constexpr size_t kScratchBytes = PTX_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device); PTX_ERR_HIP with a message if the lookup fails.
ptx_status scratch_on_current_device(const char* fn, void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    set_error("%s: no usable current device (%s)", fn, e != hipSuccess ? hipGetErrorString(e) : "index out of range");
    return PTX_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error("%s: hipGetSymbolAddress(scratch) failed on device %d: %s", fn, dev, hipGetErrorString(e));
      return PTX_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return PTX_OK;
}

}  // namespace
}  // namespace ptx

using namespace ptx;

extern "C" {

ptx_status ptx_base_step_noalloc(const void* sample, int64_t H, int64_t W, const void* probe, int64_t h, int64_t w,
                                 const int64_t* pos, int64_t B, const float* measured, float loss_scale,
                                 float* loss_per_position, float* amplitude, void* grad_sample, void* grad_probe,
                                 unsigned flags, void* stream) {
  TuneScope tune_scope;  // one schedule for the size check and the launches
  const char* fn = "ptx_base_step_noalloc";
  const size_t need = ptx_base_step_workspace(H, W, h, w, B, flags);
  if (need == 0) return base_step_impl(fn, sample, H, W, probe, h, w, pos, B, measured, nullptr, B, loss_scale,
                                       loss_per_position, amplitude, grad_sample, grad_probe, nullptr, 0, flags,
                                       (hipStream_t)stream);  // reports the shape error
  if (need > kScratchBytes) {
    set_error("%s: this shape needs %zu bytes of scratch, more than the static maximum of %zu "
              "(use ptx_base_step with a workspace)", fn, need, kScratchBytes);
    return PTX_ERR_WORKSPACE;
  }
  void* ws = nullptr;
  if (B > 0) {
    const ptx_status st = scratch_on_current_device(fn, &ws);
    if (st != PTX_OK) return st;
  }
  return base_step_impl(fn, sample, H, W, probe, h, w, pos, B, measured, nullptr, B, loss_scale, loss_per_position,
                        amplitude, grad_sample, grad_probe, ws, kScratchBytes, flags, (hipStream_t)stream);
}

}  // extern "C"
