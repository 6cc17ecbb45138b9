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

/* ---- cubic B-spline shift family ------------------------------------------
 * Geometry of Shift(n, npsi, nz, nzpsi, obj_dtype) (src/holotomocupy/shift.py:18-48):
 * coefficient grids [batch, nzpsi, npsi], detector grids [batch, nz, n];
 * is_complex selects complex64 (else float32) data.  Shifts r are float32
 * (y, x) pairs at r[b*r_stride + 0/1], magnifications m[b*m_stride]
 * (device pointers; strides in floats, so r = pos + 2*j with
 * r_stride = 2*ndist addresses pos[:, j] of a [batch, ndist, 2] array).
 * Output of every entry point is fully overwritten. */
typedef struct htamd_shift_plan htamd_shift_plan;

/* Replaces Shift.__init__ (src/holotomocupy/shift.py:18-48); builds the FP64
 * prefilter taps. */
int htamd_shift_plan_create(int64_t n, int64_t npsi, int64_t nz, int64_t nzpsi, int is_complex,
                            htamd_shift_plan** plan);
int htamd_shift_plan_destroy(htamd_shift_plan* plan);

/* out = coeff(in): periodic cubic B-spline prefilter over [batch, nzpsi, npsi]
 * (npsi, nzpsi even; out != in).  Replaces Shift.coeff (shift.py:74-82). */
int htamd_bspline_prefilter(const htamd_shift_plan* plan, const void* in, void* out, int64_t batch,
                            void* stream);

/* out[batch, nz, n] = S(c, r, m).  Replaces Shift.S / Shift.curlySc
 * (shift.py:102-111, 159-168) and the s / sf kernels with dir = 0
 * (src/holotomocupy/cuda_kernels.py:161-290). */
int htamd_bspline_interp(const htamd_shift_plan* plan, const void* c, const float* r, int64_t r_stride,
                         const float* m, int64_t m_stride, void* out, int64_t batch, void* stream);

/* c_out[batch, nzpsi, npsi] = Sadj(g, r, m) (gather form, no atomics).
 * Replaces Shift.Sadj (shift.py:113-122) and s / sf with dir = 1. */
int htamd_bspline_interp_adj(const htamd_shift_plan* plan, const void* g, const float* r,
                             int64_t r_stride, const float* m, int64_t m_stride, void* c_out,
                             int64_t batch, void* stream);

/* out = dS(c, r, m)(c1, dr).  Replaces Shift.dcurlySc (shift.py:170-183) and
 * the ds / dsf kernels (cuda_kernels.py:547-687). */
int htamd_bspline_d1(const htamd_shift_plan* plan, const void* c, const void* c1, const float* r,
                     int64_t r_stride, const float* m, int64_t m_stride, const float* dr,
                     int64_t dr_stride, void* out, int64_t batch, void* stream);

/* out = d2S(c, r, m)((c1, dr1), (c2, dr2)) with the reference's pairing.
 * Replaces Shift.d2curlySc (shift.py:206-221) and d2s / d2sf
 * (cuda_kernels.py:377-542). */
int htamd_bspline_d2(const htamd_shift_plan* plan, const void* c, const void* c1, const void* c2,
                     const float* r, int64_t r_stride, const float* m, int64_t m_stride,
                     const float* dr1, int64_t dr1_stride, const float* dr2, int64_t dr2_stride,
                     void* out, int64_t batch, void* stream);

/* Bytes of device workspace htamd_bspline_d1_adj needs for `batch` angles. */
int htamd_bspline_d1_adj_workspace_size(const htamd_shift_plan* plan, int64_t batch, int64_t* bytes);

/* Adjoint of the differential: out1[batch, nzpsi, npsi] = Sadj(g), and
 * out2[batch, 2] = (Re<g, dt1>, Re<g, dt2>), reduced deterministically.
 * Replaces Shift.dcurlySadjc (shift.py:185-204) and dsadj / dsadjf
 * (cuda_kernels.py:690-843) plus its redot reductions. */
int htamd_bspline_d1_adj(const htamd_shift_plan* plan, const void* c, const void* g, const float* r,
                         int64_t r_stride, const float* m, int64_t m_stride, void* out1, float* out2,
                         void* workspace, int64_t workspace_bytes, int64_t batch, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HTAMD_H */
