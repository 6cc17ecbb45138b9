/* libhtamd -- MI355X (gfx950) operator stack under the holotomography
 * Bilinear-Hessian solver: Fresnel propagator D/DT, cubic B-spline shift
 * family, USFFT Radon transform and the fused BH gradient.
 *
 * Conventions
 *   - All sizes are int64_t.  Every entry point returns HTAMD_OK (0) on
 *     success and a negative code otherwise; htamd_last_error() returns a
 *     thread-local message for the last failure.  No exception crosses the ABI.
 *   - Buffers are device pointers owned by the caller, C-contiguous, complex
 *     data interleaved (re, im) float32 ("complex64").
 *   - Plans own constant tables (generated in FP64, stored FP32) on the device
 *     that was current when the plan was created; using a plan while another
 *     device is current is an error.
 *   - Work is enqueued on the given HIP stream (NULL = default stream) and is
 *     stream-ordered; entry points do not synchronise.  A plan holds no
 *     mutable state, so calls with distinct buffers may run concurrently on
 *     different streams.
 */
#ifndef HTAMD_H
#define HTAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HTAMD_ABI_VERSION 1

#define HTAMD_OK 0
#define HTAMD_ERR_ARG (-1)
#define HTAMD_ERR_UNSUPPORTED (-2)
#define HTAMD_ERR_HIP (-3)

typedef struct htamd_prop_plan htamd_prop_plan;

const char* htamd_last_error(void);
int htamd_abi_version(void);

/* ---- Fresnel propagator ---------------------------------------------------
 * Replaces the reference's propagator construction and its native
 * convolution ABI: Propagation.__init__ (src/holotomocupy/propagation.py:14-44),
 * conv2d_create (src/holotomocupy/cuda/conv2d.cu:117-129).
 * n, nz: detector width / height (powers of two in [8, 2048]); distance[ndist]
 * in the same length unit as wavelength and voxelsize.  The transfer function
 * exp(-i pi wavelength distance (fx^2+fy^2)), fx = fftfreq(2n, voxelsize), is
 * evaluated in FP64. */
int htamd_prop_plan_create(int64_t n, int64_t nz, int64_t ndist, double wavelength,
                           double voxelsize, const double* distance, htamd_prop_plan** plan);

/* out[batch, nz, n] = D_j(in) (adjoint = 0) or DT_j(in) (adjoint = 1).
 * Replaces Propagation.D / Propagation.DT (src/holotomocupy/propagation.py:70-112),
 * the padding kernels pad_fwd / pad_adj (src/holotomocupy/cuda_kernels.py:78-127)
 * and conv2d_run (src/holotomocupy/cuda/conv2d.cu:131-155).  in may equal out. */
int htamd_prop_apply(const htamd_prop_plan* plan, const void* in, void* out, int64_t batch,
                     int64_t j, int adjoint, void* stream);

/* Replaces conv2d_destroy (src/holotomocupy/cuda/conv2d.cu:157-159). */
int htamd_prop_plan_destroy(htamd_prop_plan* plan);

#ifdef __cplusplus
}
#endif

#endif /* HTAMD_H */
