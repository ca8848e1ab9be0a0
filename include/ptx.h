/*
 * ptx.h -- C ABI of the MI355X-native ptychography forward/adjoint library
 * (libptychox_hip.so, HIP for gfx950).
 *
 * It replaces the pybind11 modules ``ptychoX.{base,fpm,coded}._C`` that the
 * reference's operators call (python/ptychoX/csrc/{base,fpm,coded}_extension.cpp) with plain
 * C entry points: device pointers, sizes and a stream, no torch types.
 *
 * Conventions
 *  - complex64 buffers are interleaved (re, im) float pairs, row-major,
 *    contiguous, on the device;  float buffers likewise.
 *  - positions / LED indices are device int64 [batch][2] = (x, y), zero-based,
 *    x the column (python/ptychoX/base/ops.py:61-62).
 *  - every call is stream-ordered on ``stream`` (a hipStream_t; NULL = the
 *    default stream), does not synchronise and allocates no device memory:
 *    scratch space is the caller's ``workspace`` of at least the size the
 *    matching *_workspace() query returns.
 *  - transform lengths (probe/pupil/frame sides) are powers of two in
 *    [PTX_MIN_N, PTX_MAX_N]; other sizes return PTX_ERR_UNSUPPORTED.
 *  - a non-zero status leaves a message for ptx_last_error() (thread-local).
 *  - outputs are overwritten unless PTX_FLAG_ACCUMULATE is given.
 *  - calls are thread-safe with one exception: the tuning tables set by
 *    ptx_tune_set / ptx_tune_set_groups are process-wide.  Each call reads
 *    them once, into a per-call snapshot used for both its workspace layout
 *    and its launches; but a workspace size queried before a ptx_tune_set*
 *    can be too small for a call made after it (the call then fails with
 *    PTX_ERR_WORKSPACE).  Tune before the work starts, not beside it.  Other
 *    state: the thread-local error string and timing hook.
 */
#ifndef PTX_H_
#define PTX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTX_ABI_VERSION 1
#define PTX_MIN_N 8
#define PTX_MAX_N 2048

typedef enum ptx_status {
  PTX_OK = 0,
  PTX_ERR_INVALID_ARGUMENT = 1,
  PTX_ERR_UNSUPPORTED = 2,
  PTX_ERR_OUT_OF_RANGE = 3,
  PTX_ERR_WORKSPACE = 4,
  PTX_ERR_HIP = 5
} ptx_status;

/* Object / spectrum gradient (conventional, near-field, FPM).  By default it is
 * a sorted-segment gather: every view writes its window contribution once and
 * each output tile then sums the windows that cover it in ascending view
 * order -- bitwise reproducible, the summation order of the reference torch
 * loop (python/ptychoX/base/ops.py:100-105), and faster than float atomics on
 * MI355X.  PTX_FLAG_ATOMIC_SCATTER switches to float atomic adds (the
 * reference CUDA kernels' method, base_kernel.cu:139-169); PTX_FLAG_DETERMINISTIC
 * forces the sorted-segment gather even when PTX_FLAG_ATOMIC_SCATTER is set.
 * The coded operator is deterministic in every mode. */
#define PTX_FLAG_DETERMINISTIC 1u
/* Add into the gradient outputs instead of overwriting them. */
#define PTX_FLAG_ACCUMULATE 2u
#define PTX_FLAG_ATOMIC_SCATTER 4u
/* ptx_coded_backward only: the gradient convention of the MATLAB near-field
 * operator adPtychoX_nearf (matlab/PtychoX_updated/scripts/backward/bwd_adPtychoX_nearf.cu):
 * the detector step is gZ = Z G_up/(A_up + 0.001) without the 1/ds^2 of the
 * exact derivative (kernels_nearf.cuh:507-530), and both gradients are returned
 * conjugated (its mask sum is conj(gY) X, kernels_nearf.cuh:394-436, and its
 * object gradient is the forward DFT of the conjugated spectrum sum,
 * kernels_nearf.cuh:186-240).  With transfer_1 = NULL this is that operator. */
#define PTX_FLAG_MATLAB_NEARF 8u
/* ptx_base_step / _indexed / _noalloc: the path of square frames.  Frames
 * of side 32 run by default as one fused kernel per chunk of views with the
 * whole frame in LDS; every other shape takes the row / column passes
 * through the workspace.  PTX_FLAG_THREE_PASS selects the passes for 32^2 as
 * well; PTX_FLAG_WHOLE_FRAME selects the whole-frame kernel for sides 32, 64
 * and 128 (measured slower than the passes at 128^2 on MI355X).  Results
 * agree to FP32 rounding. */
#define PTX_FLAG_THREE_PASS 16u
#define PTX_FLAG_WHOLE_FRAME 32u

/* Message of the last failing call on this thread ("" if none). */
const char* ptx_last_error(void);
/* Library version string and ABI version (PTX_ABI_VERSION). */
const char* ptx_version(void);
int ptx_abi_version(void);

/* ------------------------------------------------------------------ profiling hook
 * No counterpart in the reference (it only times whole epochs with
 * torch.cuda.Event, python/train_fpm.py:237-239,275-277).  Arms per-kernel
 * timing for the following calls on this thread: each entry point records the
 * next event of ``events`` on its stream before its first kernel (label
 * "begin") and one after each of its kernels (label = the kernel), until n
 * events are used; ptx_timing_label(k) names the interval that ends at
 * events[k].  events are hipEvent_t handles owned by the caller and must stay
 * valid while armed; n = 0 disarms.  Costs nothing when disarmed. */
ptx_status ptx_timing_events(void* const* events, int n);
int ptx_timing_count(void);
const char* ptx_timing_label(int k);

/* Streaming copy of ``bytes`` bytes from src to dst (device pointers, 16-byte
 * aligned, bytes a multiple of 16, the ranges not overlapping) on ``stream``:
 * 16-byte loads and stores, four per thread in flight, a grid of up to 16384
 * work-groups striding over the buffer.  No counterpart in the reference: it
 * measures the HBM rate a plain read + write stream reaches on this device,
 * the practical ceiling the fused passes' mixed traffic is compared with. */
ptx_status ptx_stream_copy(const void* src, void* dst, size_t bytes, void* stream);

/* ------------------------------------------------------------------ tuning table
 * No counterpart in the reference (cuFFT chose its own kernels).  Every FFT pass
 * is compiled for 8 and for 16 complex elements per thread (the radix of its
 * in-register butterflies); this table picks one per pass kind and transform
 * length.  kind: 0 row passes, 1 row passes that loop over a chunk of views
 * (object / probe accumulation), 2 column passes, 3 column passes that
 * loop over views, 4 row passes that loop over views keeping several frames'
 * values per thread (coded mask gradient), 5 row passes that evaluate the
 * amplitude loss between two transforms (coded step).  elems: 8 or 16, 0
 * restores the built-in default (8 above 16 points).
 * Process-wide; set it before launching work (the autotuner in
 * ptychox_amd/autotune.py measures and fills it). */
ptx_status ptx_tune_set(int kind, int64_t n, int elems);
int ptx_tune_get(int kind, int64_t n);

/* Work-group target of the passes that walk a chunk of views per work-group,
 * per pass kind and transform length: kind 0 the forward gathers (default
 * 1024), 1 the sorted-segment scatters (512), 2 the transfer-function
 * column passes (one work-group per CU) and the coded spectrum column pass
 * (256), 3 the coded spectrum-gradient column pass (256), 4 the coded
 * mask-gradient row pass (2048); kind 5 has no view chunks (the setting is
 * kept but unused).  The chunk of views per work-group follows
 * from it; for kinds 1, 3 and 4 so does the number of partial slices, hence
 * the workspace size (query the *_workspace functions after changing it).
 * groups: 1..65536 (1: every work-group walks all views of its rows or
 * columns), 0 restores the default.  No reference counterpart. */
ptx_status ptx_tune_set_groups(int kind, int64_t n, int groups);
int ptx_tune_get_groups(int kind, int64_t n); /* 0: the built-in default */

/* ------------------------------------------------------------------ FFT
 * Batched in-place 2-D C2C transform of ``batch`` frames of h x w,
 * unnormalised in both directions (inverse != 0 uses exp(+i)).
 * Replaces the cuFFT plan cache + execution of
 * python/ptychoX/csrc/base_kernel.cu:43-83,225-235 (and the fpm/coded copies).
 * Plans are compile-time, so there is nothing to cache or clear. */
ptx_status ptx_fft2d(void* data, int64_t batch, int64_t h, int64_t w, int inverse, void* stream);

/* ------------------------------------------------------------------ conventional
 * Forward of ptychoX_base: far_field[b] = fftshift(fft2(sample[y_b:y_b+h, x_b:x_b+w] * probe)),
 * amplitude[b] = |far_field[b]|.
 * Replaces ptychox_base_forward_cuda (python/ptychoX/csrc/base_kernel.cu:302-329,
 * bound at base_extension.cpp:21-32) and the torch branch base/ops.py:60-66.
 * The reference also returns the patch stack; this library re-gathers patches
 * from ``sample`` in the backward instead of storing them.
 * Positions are device data and are not checked here: a window that leaves the
 * sample gives a zero frame (and no gradient in the adjoint and step entry
 * points) without reading or writing outside the sample.  The Python operator
 * checks them on the host first and raises ValueError.
 * far_field: [batch, h, w] complex64 out; amplitude: [batch, h, w] float out. */
ptx_status ptx_base_forward(const void* sample, int64_t sample_h, int64_t sample_w,
                            const void* probe, int64_t probe_h, int64_t probe_w,
                            const int64_t* positions_xy, int64_t batch,
                            void* far_field, float* amplitude, void* stream);

/* Workspace bytes ptx_base_backward needs. */
size_t ptx_base_backward_workspace(int64_t sample_h, int64_t sample_w, int64_t probe_h, int64_t probe_w,
                                   int64_t batch, unsigned flags);

/* Backward of ptychoX_base: with Gamma = g * Phi / max(|Phi|, FLT_EPSILON) and
 * chi_b = N * ifft2(ifftshift(Gamma_b)):
 *   grad_sample[y_b:y_b+h, x_b:x_b+w] += chi_b * conj(probe),
 *   grad_probe += sum_b chi_b * conj(patch_b).
 * Replaces ptychox_base_backward_cuda (python/ptychoX/csrc/base_kernel.cu:331-357,
 * bound at base_extension.cpp:34-60) and the torch branch base/ops.py:93-105.
 * grad_output: [batch, h, w] float; far_field: forward's output;
 * grad_sample: [sample_h, sample_w] complex64; grad_probe: [h, w] complex64. */
ptx_status ptx_base_backward(const float* grad_output, const void* far_field,
                             const void* sample, int64_t sample_h, int64_t sample_w,
                             const void* probe, int64_t probe_h, int64_t probe_w,
                             const int64_t* positions_xy, int64_t batch,
                             void* grad_sample, void* grad_probe,
                             void* workspace, size_t workspace_bytes, unsigned flags, void* stream);

/* Workspace bytes ptx_base_step needs. */
size_t ptx_base_step_workspace(int64_t sample_h, int64_t sample_w, int64_t probe_h, int64_t probe_w,
                               int64_t batch, unsigned flags);

/* Fused training step of the conventional operator with the amplitude loss of
 * the reference driver (python/train_ptychography.py:549-558: ptychoX_base ->
 * F.mse_loss -> backward), in three passes over the batch:
 *   L = loss_scale * sum_{b,i} (A_bi - Y_bi)^2   (loss_scale = 1/(batch*h*w) is mse_loss "mean"),
 *   loss_per_position[b] = mean_i (A_bi - Y_bi)^2,
 *   grad_sample, grad_probe = dL/d(sample, probe) in the PyTorch convention.
 * measured: [batch, h, w] float;  amplitude: optional [batch, h, w] float out (NULL to skip). */
ptx_status ptx_base_step(const void* sample, int64_t sample_h, int64_t sample_w,
                         const void* probe, int64_t probe_h, int64_t probe_w,
                         const int64_t* positions_xy, int64_t batch,
                         const float* measured, float loss_scale,
                         float* loss_per_position, float* amplitude,
                         void* grad_sample, void* grad_probe,
                         void* workspace, size_t workspace_bytes, unsigned flags, void* stream);

/* Same step for a permuted mini-batch, without copying data: view b of the
 * batch is entry order[b] of the full tables positions_all [count][2] and
 * measured_all [count, h, w] (python/train_ptychography.py:532-553 draws
 * torch.randperm each epoch and indexes both tables).  An order entry outside
 * [0, count) reads entry 0 instead.  count must lie in [1, 2^31 - 1] when
 * batch > 0 (PTX_ERR_INVALID_ARGUMENT otherwise).  Workspace:
 * ptx_base_step_workspace. */
ptx_status ptx_base_step_indexed(const void* sample, int64_t sample_h, int64_t sample_w,
                                 const void* probe, int64_t probe_h, int64_t probe_w,
                                 const int64_t* positions_all, const float* measured_all, int64_t count,
                                 const int64_t* order, int64_t batch, float loss_scale,
                                 float* loss_per_position, float* amplitude,
                                 void* grad_sample, void* grad_probe,
                                 void* workspace, size_t workspace_bytes, unsigned flags, void* stream);

/* Allocation-free variant of ptx_base_step.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory at all
 * (code under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc), which therefore cannot provide a workspace either.  Same
 * arguments and results as ptx_base_step without the workspace: the scratch
 * space is a static device array of PTX_NOALLOC_SCRATCH_BYTES (192 MiB) that
 * each GPU holds from the first call on it until the process ends.  A shape
 * whose ptx_base_step_workspace exceeds that returns PTX_ERR_WORKSPACE before
 * anything is launched (object 2048^2 / probe 512^2 / batch 64 needs 162 MiB
 * with the built-in schedule).  The call allocates nothing.  Every call on a
 * GPU uses that GPU's one array: calls that overlap on one GPU (other streams,
 * other threads) race and must be serialised by the caller; calls in one
 * stream are ordered.  Nothing in the array needs to start at zero. */
#define PTX_NOALLOC_SCRATCH_BYTES ((size_t)192 << 20)
ptx_status ptx_base_step_noalloc(const void* sample, int64_t sample_h, int64_t sample_w,
                                 const void* probe, int64_t probe_h, int64_t probe_w,
                                 const int64_t* positions_xy, int64_t batch,
                                 const float* measured, float loss_scale,
                                 float* loss_per_position, float* amplitude,
                                 void* grad_sample, void* grad_probe, unsigned flags, void* stream);

/* ------------------------------------------------------------------ optimiser
 * One Adam update with torch.optim.Adam's arithmetic (foreach variant) as the
 * reference drivers configure it (python/train_ptychography.py:500-507,
 * python/train_fpm.py, python/train_coded.py): on n floats -- a complex64
 * tensor is 2n floats updated per real component, as torch does for complex
 * parameters.  hyper: device float[3] = {lr / (1 - beta1^t), sqrt(1 - beta2^t),
 * eps}, read when the kernel runs, so a captured graph of a whole epoch is
 * replayed with new step sizes by rewriting that buffer.
 * PTX_ADAM_PROJECT_REAL_NONNEG (param is complex64, n even) follows the update
 * with the coded driver's constraint on the coded surface,
 * param = complex(max(Re param, 0), 0) (python/train_coded.py:240-241); the
 * moments of both components are updated as torch does before that copy. */
#define PTX_ADAM_PROJECT_REAL_NONNEG 1u
ptx_status ptx_adam_step(float* param, const float* grad, float* exp_avg, float* exp_avg_sq, int64_t n,
                         float beta1, float beta2, const float* hyper, unsigned flags, void* stream);

/* Adam state of one complex64 unknown for ptx_base_train_step: moments of
 * 2 floats per element and the device step-size table of ptx_adam_step. */
typedef struct ptx_adam {
  float* exp_avg;
  float* exp_avg_sq;
  const float* hyper; /* device float[3] = {lr / (1 - beta1^t), sqrt(1 - beta2^t), eps} */
  float beta1, beta2;
} ptx_adam;

/* One mini-batch of the conventional driver's training loop
 * (python/train_ptychography.py:549-559: ptychoX_base -> F.mse_loss ->
 * backward -> optimizer.step()) as ptx_base_step_indexed followed by
 * ptx_adam_step on the sample and on the probe, with the Adam updates applied
 * where the gradients are finished (the object-gradient gather's and the
 * probe-partial sum's epilogues) instead of storing the gradients and reading
 * them back.  Results equal those of that sequence of calls.  sample and probe
 * are updated in place.  adam_sample / adam_probe NULL: that unknown is not
 * updated and its gradient is written to grad_sample / grad_probe as by
 * ptx_base_step_indexed; otherwise that buffer is scratch of the same shape
 * whose contents afterwards are unspecified.  order NULL: the batch is the
 * tables' first `batch` entries.  PTX_FLAG_ACCUMULATE is rejected when an
 * update is requested.  Workspace: ptx_base_step_workspace. */
ptx_status ptx_base_train_step(void* sample, int64_t sample_h, int64_t sample_w,
                               void* probe, int64_t probe_h, int64_t probe_w,
                               const int64_t* positions_all, const float* measured_all, int64_t count,
                               const int64_t* order, int64_t batch, float loss_scale,
                               float* loss_per_position, float* amplitude,
                               void* grad_sample, void* grad_probe,
                               const ptx_adam* adam_sample, const ptx_adam* adam_probe,
                               void* workspace, size_t workspace_bytes, unsigned flags, void* stream);

/* positions_out[b] = positions_all[order[b]] (int64 pairs); entries of order
 * outside [0, count) read entry 0 and set *bad_index_flag to 1 (the caller
 * zeroes the flag).  An empty table (count 0) with batch > 0 is
 * PTX_ERR_INVALID_ARGUMENT. */
ptx_status ptx_gather_positions(const int64_t* order, const int64_t* positions_all, int64_t count, int64_t batch,
                                int64_t* positions_out, int* bad_index_flag, void* stream);

/* ------------------------------------------------------------------ Fourier ptychography
 * Forward of ptychoX_fourier: object_fft = fft2(object_field);
 * sub_b[r,c] = object_fft[(y_b + r + H/2) mod H, (x_b + c + W/2) mod W];
 * image_field[b] = ifft2(fftshift(sub_b * pupil)); amplitude = |image_field|.
 * led_indices_xy may be negative (wrap-around, python/ptychoX/fpm/ops.py:122-123).
 * Replaces ptychox_fpm_forward_cuda (python/ptychoX/csrc/fpm_kernel.cu:233-273, bound at
 * fpm_extension.cpp:21-32) and the torch branch fpm/ops.py:56-59.  The reference
 * returns the gathered sub-spectra; this library re-gathers them from object_fft.
 * object_fft: [H, W] complex64 out; image_field: [batch, h, w] complex64 out;
 * amplitude: [batch, h, w] float out. */
ptx_status ptx_fpm_forward(const void* object_field, int64_t object_h, int64_t object_w,
                           const void* pupil, int64_t pupil_h, int64_t pupil_w,
                           const int64_t* led_indices_xy, int64_t batch,
                           void* object_fft, void* image_field, float* amplitude, void* stream);

size_t ptx_fpm_backward_workspace(int64_t object_h, int64_t object_w, int64_t pupil_h, int64_t pupil_w,
                                  int64_t batch, unsigned flags);

/* Backward of ptychoX_fourier: g_b = ifftshift(fft2(G E/max(|E|, FLT_EPSILON)))/(h w);
 * grad_pupil = sum_b g_b conj(sub_b); grad_object = H W ifft2(index_add(g_b conj(pupil))).
 * Replaces ptychox_fpm_backward_cuda (python/ptychoX/csrc/fpm_kernel.cu:275-317, bound at
 * fpm_extension.cpp:34-60) and the torch branch fpm/ops.py:85-97.
 * grad_object: [H, W] complex64; grad_pupil: [h, w] complex64. */
ptx_status ptx_fpm_backward(const float* grad_output, const void* image_field, const void* object_fft,
                            int64_t object_h, int64_t object_w,
                            const void* pupil, int64_t pupil_h, int64_t pupil_w,
                            const int64_t* led_indices_xy, int64_t batch,
                            void* grad_object, void* grad_pupil,
                            void* workspace, size_t workspace_bytes, unsigned flags, void* stream);

size_t ptx_fpm_step_workspace(int64_t object_h, int64_t object_w, int64_t pupil_h, int64_t pupil_w,
                              int64_t batch, unsigned flags);

/* Fused FPM training step with the amplitude MSE loss of python/train_fpm.py:257-263
 * (same loss conventions as ptx_base_step). */
ptx_status ptx_fpm_step(const void* object_field, int64_t object_h, int64_t object_w,
                        const void* pupil, int64_t pupil_h, int64_t pupil_w,
                        const int64_t* led_indices_xy, int64_t batch,
                        const float* measured, float loss_scale,
                        float* loss_per_position, float* amplitude,
                        void* grad_object, void* grad_pupil,
                        void* workspace, size_t workspace_bytes, unsigned flags, void* stream);

/* ------------------------------------------------------------------ near-field
 * Forward of ptychoX_nearfield: field[b] = ifft2(transfer * fft2(sample[y_b:y_b+h, x_b:x_b+w] * probe)),
 * amplitude = |field|; transfer is [h, w] complex64 with DC at [0, 0] (angular
 * spectrum exp(i pi lambda z (fx^2 + fy^2)) for example).
 * Replaces the MATLAB near-field operator's propagation chain
 * (matlab/PtychoX_updated/scripts/forward/fwd_adPtychoX_nearf.cu:58-114), applied
 * per scan window; the reference has no Python near-field operator. */
ptx_status ptx_nearfield_forward(const void* sample, int64_t sample_h, int64_t sample_w,
                                 const void* probe, int64_t probe_h, int64_t probe_w, const void* transfer,
                                 const int64_t* positions_xy, int64_t batch,
                                 void* field, float* amplitude, void* stream);

size_t ptx_nearfield_backward_workspace(int64_t sample_h, int64_t sample_w, int64_t probe_h, int64_t probe_w,
                                        int64_t batch, unsigned flags);

/* Backward: chi_b = ifft2(conj(transfer) fft2(G field/max(|field|, FLT_EPSILON)));
 * grad_sample[win_b] += chi_b conj(probe); grad_probe += sum_b chi_b conj(patch_b).
 * Replaces matlab/PtychoX_updated/scripts/backward/bwd_adPtychoX_nearf.cu:81-144
 * (whose gradient is the heuristic dl/(out + 0.001), kernels_nearf.cuh:525;
 * this one is the exact adjoint). */
ptx_status ptx_nearfield_backward(const float* grad_output, const void* field,
                                  const void* sample, int64_t sample_h, int64_t sample_w,
                                  const void* probe, int64_t probe_h, int64_t probe_w, const void* transfer,
                                  const int64_t* positions_xy, int64_t batch,
                                  void* grad_sample, void* grad_probe,
                                  void* workspace, size_t workspace_bytes, unsigned flags, void* stream);

size_t ptx_nearfield_step_workspace(int64_t sample_h, int64_t sample_w, int64_t probe_h, int64_t probe_w,
                                    int64_t batch, unsigned flags);

/* Fused near-field training step with the amplitude MSE loss (conventions of ptx_base_step). */
ptx_status ptx_nearfield_step(const void* sample, int64_t sample_h, int64_t sample_w,
                              const void* probe, int64_t probe_h, int64_t probe_w, const void* transfer,
                              const int64_t* positions_xy, int64_t batch,
                              const float* measured, float loss_scale,
                              float* loss_per_position, float* amplitude,
                              void* grad_sample, void* grad_probe,
                              void* workspace, size_t workspace_bytes, unsigned flags, void* stream);

/* ------------------------------------------------------------------ coded ptychography
 * Workspace bytes of the coded entry points: step = 0 for ptx_coded_forward and
 * ptx_coded_backward, step != 0 for ptx_coded_step. */
size_t ptx_coded_workspace(int64_t height, int64_t width, int64_t batch, int64_t downsample, int step);

/* Forward of ptychoX_coded on H x W frames:
 *   object_spectrum = fft2(object_field);  R_b = exp(i 2 pi (dx_b fx + dy_b fy)) (signed integer f);
 *   x_forward[b] = ifft2(object_spectrum transfer_1 R_b);
 *   sensor_field[b] = ifft2(fft2(x_forward[b] coded_surface) transfer_2);
 *   amplitude[b] = sqrt(max(avgpool_ds |sensor_field[b]|^2, 0))  -> [batch, H/ds, W/ds].
 * shifts_xy: device float [batch][2] normalised shifts (dx, dy).
 * transfer_1 may be NULL for H1 = 1 (every coded entry point): the full-frame
 * near-field form of the MATLAB operator, whose first propagator is disabled
 * (matlab/PtychoX_updated/scripts/forward/fwd_adPtychoX_nearf.cu:63-114,
 * kernels_nearf.cuh:34-80).
 * Replaces ptychox_coded_forward_cuda (python/ptychoX/csrc/coded_kernel.cu:283-345, bound at
 * coded_extension.cpp:26-56) and the torch path coded/ops.py:136-154.  x_forward and
 * sensor_field are normalised like the torch path (the CUDA extension keeps them
 * unnormalised, coded_kernel.cu:318,341). */
ptx_status ptx_coded_forward(const void* object_field, int64_t height, int64_t width, const void* coded_surface,
                             const void* transfer_1, const void* transfer_2, const float* shifts_xy, int64_t batch,
                             int64_t downsample, void* object_spectrum, void* x_forward, void* sensor_field,
                             float* amplitude, void* workspace, size_t workspace_bytes, void* stream);

/* Backward of ptychoX_coded w.r.t. object_field and coded_surface (transfer
 * functions and shifts get no gradient, coded/ops.py:133):
 *   gZ = Z G_up/(ds^2 max(A_up, 1e-7)); gY = ifft2(conj(H2) fft2(gZ));
 *   grad_coded = sum_b gY conj(X_b); grad_object = N ifft2(sum_b conj(H1 R_b) fft2(gY conj(C))/N).
 * Replaces ptychox_coded_backward_cuda (python/ptychoX/csrc/coded_kernel.cu:347-424, bound at
 * coded_extension.cpp:58-88).  Always deterministic: both shared gradients are
 * accumulated over fixed chunks of views and summed in a fixed order.
 * PTX_FLAG_MATLAB_NEARF selects the MATLAB near-field gradient convention
 * (see the flag); PTX_FLAG_ACCUMULATE adds into the outputs. */
ptx_status ptx_coded_backward(const float* grad_output, const void* object_spectrum, const void* x_forward,
                              const void* sensor_field, const void* coded_surface, const void* transfer_1,
                              const void* transfer_2, const float* shifts_xy, int64_t batch, int64_t height,
                              int64_t width, int64_t downsample, void* grad_object, void* grad_coded,
                              void* workspace, size_t workspace_bytes, unsigned flags, void* stream);

/* Fused coded training step with the amplitude MSE loss of python/train_coded.py:231-238
 * (loss conventions of ptx_base_step; loss_per_position averages over the H/ds x W/ds pixels). */
ptx_status ptx_coded_step(const void* object_field, int64_t height, int64_t width, const void* coded_surface,
                          const void* transfer_1, const void* transfer_2, const float* shifts_xy, int64_t batch,
                          int64_t downsample, const float* measured, float loss_scale, float* loss_per_position,
                          float* amplitude, void* grad_object, void* grad_coded, void* workspace,
                          size_t workspace_bytes, unsigned flags, void* stream);

/* ------------------------------------------------------------------ alternative losses
 * Stencil losses on stacks of frames [batch][h][w], with their gradients.
 * With d = pred - target (or the field), forward differences with wrap-around
 * dx(i,j) = d(i,j+1) - d(i,j), dy(i,j) = d(i+1,j) - d(i,j) (torch.roll(d, -1),
 * MATLAB d([2:end,1])) and a per-site penalty phi:
 *   PTX_LOSS_L2           phi = |d|^2              (+loss_fun/l2_loss.m; mse_loss "mean" with scale 1/numel)
 *   PTX_LOSS_FD_ISO       phi = sqrt(|dx|^2 + |dy|^2 + eps)
 *                         (python/train_coded.py:101-106, eps 1e-5; +loss_fun/tv_loss.m 'isotropic' with eps 0)
 *   PTX_LOSS_FD_ANISO     phi = |dx| + |dy|        (python/train_coded.py:107; +loss_fun/fd_loss.m, tv_loss.m)
 *   PTX_LOSS_GRAD_L2      phi = |dx|^2 + |dy|^2    (+loss_fun/tv_loss.m, sparse_loss.m 'L2')
 *   PTX_LOSS_SPARSE_ISO   phi = 1 - exp(-8 sqrt(|dx|^2 + |dy|^2 + eps))   (+loss_fun/sparse_loss.m, eps 1e-5)
 *   PTX_LOSS_SPARSE_ANISO phi = 1 - exp(-8 (|dx| + |dy|))                 (+loss_fun/sparse_loss.m)
 * loss_per_frame[b] = scale * sum_ij phi; grad = scale * dL/dpred in PyTorch's
 * convention (real: dL/dx; complex: dL/dRe + i dL/dIm).  |.| has the zero
 * subgradient at 0 (torch.abs backward); with eps = 0 a flat site contributes
 * a zero gradient.  Sums are deterministic (fixed-order work-group and frame
 * reductions).  workspace: ptx_loss_workspace bytes. */
#define PTX_LOSS_L2 0
#define PTX_LOSS_FD_ISO 1
#define PTX_LOSS_FD_ANISO 2
#define PTX_LOSS_GRAD_L2 3
#define PTX_LOSS_SPARSE_ISO 4
#define PTX_LOSS_SPARSE_ANISO 5

size_t ptx_loss_workspace(int64_t batch, int64_t h, int64_t w);

/* Residual loss of predicted amplitudes (pred, target: device float [batch][h][w];
 * target may be NULL for d = pred).  The loss step of the reference training
 * drivers (python/train_coded.py:234-237 with --loss fd-isotropic / fd-anisotropic),
 * the upstream gradient of ptx_*_backward. */
ptx_status ptx_amplitude_loss(int kind, const float* pred, const float* target, int64_t batch, int64_t h, int64_t w,
                              float scale, float eps, float* loss_per_frame, float* grad, void* workspace,
                              size_t workspace_bytes, void* stream);

/* Regulariser of complex fields (device complex64 [batch][h][w]), the MATLAB
 * examples' tv_loss / sparse_loss of an object (matlab/Examples/Ptychography/+loss_fun). */
ptx_status ptx_field_regularizer(int kind, const void* field, int64_t batch, int64_t h, int64_t w, float scale,
                                 float eps, float* loss_per_frame, void* grad, void* workspace,
                                 size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PTX_H_ */
