/* fourierconvolution_amd.h -- additive C ABI of libFourierConvolutionCUDALib
 * (MI355X build).  None of these symbols exists in the reference; each one
 * names the reference behaviour it extends.  Dims use the {x, y, z} order of
 * convolution3DfftCUDAInPlace (x contiguous).
 *
 * Return codes: 0 success, 1 invalid argument, 2 unsupported size,
 * 3 out of device memory, 4 HIP runtime error, 5 no device, 6 internal.
 */
#ifndef FOURIERCONVOLUTION_AMD_H
#define FOURIERCONVOLUTION_AMD_H

#include "FourierConvolutionCUDALib_Export.h"

#ifdef __cplusplus
#define FCAMD_API extern "C" FourierConvolutionCUDALib_EXPORT
#else
#define FCAMD_API extern FourierConvolutionCUDALib_EXPORT
#endif

/* Operation selector for the convolution entry points. */
#define FCAMD_CONVOLVE 0  /* reference semantics (src/convolution3Dfft.cu:41-62) */
#define FCAMD_CORRELATE 1 /* adjoint w.r.t. the image: conjugated kernel spectrum */

/* Error state of the calling thread (the reference prints and exits instead,
 * src/book.h:22-30). */
FCAMD_API int fcamd_last_error(void);
FCAMD_API const char* fcamd_last_error_message(void);

/* delete[] for the buffers convolution3DfftCUDA / _test return, for callers
 * that cannot call C++ delete[] (src/convolution3Dfft.cu:240,327). */
FCAMD_API void fcamd_free_result(float* p);

/* convolution3DfftCUDAInPlace with a return code and an operation selector. */
FCAMD_API int fcamd_convolve3d_host(float* im, const int* imDim, const float* kernel,
                                    const int* kernelDim, int op, int devCUDA);

/* Device memory, in bytes, that convolution3DfftCUDAInPlace /
 * fcamd_convolve3d_host allocate on device devCUDA for these dims (the
 * reference reports only cuFFT's work area, src/convolution3Dfft.cu:587-603).
 * Host-pointer calls of 32 MiB or more, and any call whose single-shot
 * buffers exceed the device budget, stream z-slabs through a ring of slab
 * buffers (no whole volume on the device).  The budget is
 * FCAMD_DEVICE_BUDGET_MB when set (read on every call; then no device is
 * needed for this query), else the device's free memory.  -1 on error. */
FCAMD_API long long fcamd_host_device_bytes(const int* imDim, const int* kernelDim, int devCUDA);

/* Device-resident form of the hot path: every pointer is device memory on
 * the current device; d_out may alias d_image; `stream` is a hipStream_t
 * (NULL = the null stream).  d_workspace may be NULL (a per-device pooled
 * buffer is used; calls on one device are then serialised) or must hold
 * fcamd_workspace_bytes() bytes.  Asynchronous with respect to the host. */
FCAMD_API long long fcamd_workspace_bytes(const int* imDim, const int* kernelDim);
FCAMD_API int fcamd_convolve3d_device(float* d_out, const float* d_image, const int* imDim,
                                      const float* d_kernel, const int* kernelDim, int op,
                                      void* d_workspace, long long workspace_bytes, void* stream);

/* fcamd_convolve3d_device without allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory -- code under
 * graph capture, tight serving loops, harnesses that forbid allocation.  The
 * workspace and the twiddle tables live in one device array of the library,
 * FCAMD_NOALLOC_MAX_WORKSPACE bytes of workspace plus 1 MiB of tables, that
 * each device holds until the process ends; a call whose
 * fcamd_workspace_bytes() exceeds FCAMD_NOALLOC_MAX_WORKSPACE fails with code
 * 2 (unsupported size).  The call allocates nothing, so it may be recorded
 * into a graph, also for lengths the process has not used before.  Every
 * call on one device shares that array: calls that overlap on one device
 * (other threads, other streams) race; order them.  Same results as
 * fcamd_convolve3d_device; d_out may alias d_image. */
#define FCAMD_NOALLOC_MAX_WORKSPACE (128LL << 20)
FCAMD_API int fcamd_convolve3d_device_noalloc(float* d_out, const float* d_image,
                                              const int* imDim, const float* d_kernel,
                                              const int* kernelDim, int op, void* stream);

/* Adjoint with respect to the kernel (no counterpart in the reference; the
 * gradient half of the differentiable form).  For out = conv(im, k) and an
 * upstream gradient g (image-sized), writes
 *   dk[j] = sum_i g[i] * im[(i - (j - floor(k/2))) mod n]
 * for every kernel tap j (kernelDim-sized, device memory).  The gradient
 * with respect to the image is fcamd_convolve3d_device(..., FCAMD_CORRELATE,
 * ...) applied to g.  Workspace as for fcamd_convolve3d_device. */
FCAMD_API long long fcamd_grad_kernel_workspace_bytes(const int* imDim, const int* kernelDim);
FCAMD_API int fcamd_conv_grad_kernel_device(float* d_dkernel, const float* d_grad_out,
                                            const float* d_image, const int* imDim,
                                            const int* kernelDim, void* d_workspace,
                                            long long workspace_bytes, void* stream);

/* Boundary modes (no counterpart in the reference; its header's TODO at
 * src/convolution3Dfft.h:39,54 asks for padding "to impose the boundary
 * conditions that you want: mirror, zero, etc.").  The drop-in entry points
 * are circular (FCAMD_BOUNDARY_PERIODIC).  The other modes give the linear
 * convolution of the image extended beyond its edges by
 *   ZERO     zeros                               0 0 | a b c d | 0 0
 *   NEAREST  the edge value                      a a | a b c d | d d
 *   REFLECT  half-sample symmetric (edge twice)  b a | a b c d | d c
 *   MIRROR   whole-sample symmetric              c b | a b c d | c b
 * (scipy.ndimage's 'constant', 'nearest', 'reflect' and 'mirror'); the
 * kernel is centred at k/2 as in convolution3DfftCUDAInPlace and may exceed
 * the image.  Each axis with k > 1 is transformed at a fast length
 * P >= n + k - 1 (fcamd_padded_extents); the padding is virtual (read
 * through the boundary map in the first pass, cropped in the last).
 * op FCAMD_CORRELATE correlates under the same extension (the adjoint for
 * PERIODIC and ZERO). */
#define FCAMD_BOUNDARY_PERIODIC 0
#define FCAMD_BOUNDARY_ZERO 1
#define FCAMD_BOUNDARY_NEAREST 2
#define FCAMD_BOUNDARY_REFLECT 3
#define FCAMD_BOUNDARY_MIRROR 4
FCAMD_API long long fcamd_boundary_workspace_bytes(const int* imDim, const int* kernelDim,
                                                   int boundary);
FCAMD_API int fcamd_padded_extents(const int* imDim, const int* kernelDim, int boundary,
                                   int* padded);
FCAMD_API int fcamd_convolve3d_boundary_device(float* d_out, const float* d_image,
                                               const int* imDim, const float* d_kernel,
                                               const int* kernelDim, int op, int boundary,
                                               void* d_workspace, long long workspace_bytes,
                                               void* stream);
/* host-pointer form of the above, in place like convolution3DfftCUDAInPlace */
FCAMD_API int fcamd_convolve3d_boundary_host(float* im, const int* imDim, const float* kernel,
                                             const int* kernelDim, int op, int boundary,
                                             int devCUDA);

/* Convolution plans (the kernel-spectrum reuse the reference leaves as a
 * TODO, src/convolution3Dfft.cu:265,363 "reuse the results").  create
 * validates the shapes, takes a device workspace (NULL: the plan allocates
 * fcamd_workspace_bytes()), and transforms the kernel's rows and planes
 * once on `stream`; execute runs only the image's passes (same result as
 * fcamd_convolve3d_device with this kernel and op; d_out may alias
 * d_image); set_kernel replaces the kernel (same kernel dims).  Calls on one
 * plan share its workspace and must be ordered (one stream, or
 * synchronised).  create returns NULL on error. */
FCAMD_API void* fcamd_conv_plan_create(const int* imDim, const int* kernelDim,
                                       const float* d_kernel, int op, void* d_workspace,
                                       long long workspace_bytes, void* stream);
FCAMD_API int fcamd_conv_plan_set_kernel(void* plan, const float* d_kernel, void* stream);
FCAMD_API int fcamd_conv_plan_execute(void* plan, float* d_out, const float* d_image,
                                      void* stream);
FCAMD_API void fcamd_conv_plan_destroy(void* plan);

/* Batched uneven blocks (no counterpart in the reference): `nblocks`
 * convolutions, block b with dims imDims[3b..3b+2] ({x,y,z}), kernel dims
 * kernelDims[3b..3b+2], device buffers d_images[b], d_kernels[b] and
 * d_outs[b] (may alias the image).  fcamd_batch_execute enqueues the
 * sequence on `stream`, fcamd_batch_graph_launch captures it into hipGraphs
 * on the first call (one per lane, or one for the length groups) and
 * replays them afterwards, forked from and joined back into `stream`.  The buffers are
 * bound at creation.  Two schedules, chosen at creation by the environment:
 *   FCAMD_BATCH_LANES=N (unset = 4): the blocks are dealt to N streams
 *     ("lanes"), largest estimated cost first to the least-loaded lane;
 *     each lane runs its blocks' pass sequences one after another in its
 *     own workspace slot (the workspace is one slot per lane).
 *   FCAMD_BATCH_LANES=0: the blocks of every pass are grouped by transform
 *     length (one launch per length, length groups on forked streams), in
 *     waves of consecutive blocks whose workspaces add up to at most
 *     FCAMD_BATCH_WAVE_MB MiB (unset or 0 = a single wave); waves share one
 *     workspace region.
 * d_workspace may be NULL (the plan allocates it) or hold
 * fcamd_batch_workspace_bytes() bytes (computed under the same settings).
 * Every run of one plan uses that plan's workspace, side streams and events,
 * so runs of one plan must be ordered, e.g. all on one stream.  Runs on
 * different streams that overlap race.  create returns NULL on error. */
FCAMD_API long long fcamd_batch_workspace_bytes(int nblocks, const int* imDims,
                                                const int* kernelDims);
FCAMD_API void* fcamd_batch_create(int nblocks, const int* imDims, const int* kernelDims,
                                   float* const* d_outs, const float* const* d_images,
                                   const float* const* d_kernels, int op, void* d_workspace,
                                   long long workspace_bytes);
FCAMD_API int fcamd_batch_execute(void* plan, void* stream);
FCAMD_API int fcamd_batch_graph_launch(void* plan, void* stream);
FCAMD_API void fcamd_batch_destroy(void* plan);

/* Multiview Richardson-Lucy deconvolution (no counterpart in the
 * reference; the iterative caller of the hot path):
 *   psi <- psi * (1/V) sum_v corr(phi_v / max(conv(psi, P_v), eps), P_v)
 * with conv / corr the convolution semantics of convolution3DfftCUDAInPlace
 * and its adjoint.  V = nviews (1..8) views d_views[v] share the image dims
 * imDim ({x,y,z}); view v's PSF d_kernels[v] has dims kernelDims[3v..3v+2].
 * init_psi = 1 starts from the mean of the views, else from d_psi.  fused = 1
 * uses the fused pass sequence when every axis length has compile-time
 * kernels (fcamd_rl_fused_available), else the per-view form.  All pointers
 * are device memory; the pointer arrays are host arrays. */
FCAMD_API long long fcamd_rl_workspace_bytes(int nviews, const int* imDim, const int* kernelDims,
                                             int fused);
FCAMD_API int fcamd_rl_fused_available(int nviews, const int* imDim, const int* kernelDims);
FCAMD_API int fcamd_rl_multiview_device(float* d_psi, const float* const* d_views, int nviews,
                                        const int* imDim, const float* const* d_kernels,
                                        const int* kernelDims, int iterations, float eps,
                                        int init_psi, int fused, void* d_workspace,
                                        long long workspace_bytes, void* stream);

/* Per-pass HIP-event timing of the convolution passes (diagnostics and the
 * benchmark's roofline figures).  While enabled every pass launch is
 * bracketed by two events on its stream.  fcamd_pass_times waits for them,
 * writes summed milliseconds and launch counts per pass (up to `cap`
 * passes), resets, and returns the pass count (negative error code on
 * failure).  Passes: 0 kernel rows x-FFT, 1 kernel planes y-FFT,
 * 2 image x-R2C, 3 image y-FFT, 4 fused z pass, 5 inverse y-FFT,
 * 6 x-C2R. */
FCAMD_API void fcamd_pass_timing(int enable);
FCAMD_API int fcamd_pass_times(double* ms, int* counts, int cap);
FCAMD_API const char* fcamd_pass_name(int pass);

/* Unfused pointwise spectrum product on device arrays of n complex values:
 * dst <- scale * dst * src (conj(src) when conj != 0); the stand-alone form
 * of src/convolution3Dfft.cu:47-62. */
FCAMD_API int fcamd_modulate_device(void* d_dst, const void* d_src, long long n, float scale,
                                    int conj, void* stream);

/* Plain streaming copy of `bytes` (multiple of 16) between device buffers,
 * used to calibrate HBM byte counters and measure achievable bandwidth. */
FCAMD_API int fcamd_stream_copy_device(void* d_dst, const void* d_src, long long bytes,
                                       void* stream);

/* FFT radix plan the library uses for a transform length (for diagnostics):
 * writes up to `cap` radices into `radices`, returns the stage count. */
FCAMD_API int fcamd_radix_plan(int length, int* radices, int cap);

/* 1 when a compile-time (register-resident) kernel serves this axis length
 * for the given pass (0 = x forward, 1 = x inverse, 2 = y forward,
 * 3 = fused z, 4 = y inverse). */
FCAMD_API int fcamd_has_static_kernel(int pass, int length);

/* Lists the lengths with compile-time kernels (up to `cap`); returns how many
 * there are. */
FCAMD_API int fcamd_static_lengths(int* lengths, int cap);

/* Variants of the compile-time kernels (pass: 0 x forward, 1 x inverse,
 * 2 y column forward, 3 fused z, 4 y column inverse; 2 and 4 offer the same
 * variants and are chosen separately).  Names describe the tile shape and method; the
 * selection is per (pass, length) and process-wide; fcamd_autotune sets it
 * from timings.  select returns 0 or an error code. */
FCAMD_API int fcamd_static_variant_count(int pass, int length);
FCAMD_API const char* fcamd_static_variant_name(int pass, int length, int index);
FCAMD_API int fcamd_select_static_variant(int pass, int length, int index);
FCAMD_API int fcamd_selected_static_variant(int pass, int length);

/* Plan-time autotuner: for a convolution of imDim / kernelDim on the current
 * device, times every variant of every pass (reps timed runs each, on
 * `stream`, with scratch buffers it allocates and frees) and keeps the
 * fastest per (pass, length).  Do not run convolutions concurrently on other
 * threads while it runs.  With FCAMD_TUNE_CACHE set, choices are appended to
 * that file; fcamd_autotune_load(path or NULL for the variable) re-applies
 * them. */
FCAMD_API int fcamd_autotune(const int* imDim, const int* kernelDim, int reps, void* stream);
FCAMD_API int fcamd_autotune_load(const char* path);

/* enable = 0 routes every pass through the runtime-length kernels (for
 * cross-checks and comparisons); 1 restores the default.  Process-wide. */
FCAMD_API void fcamd_use_static_kernels(int enable);

#endif /* FOURIERCONVOLUTION_AMD_H */
