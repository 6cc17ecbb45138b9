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

/* Library housekeeping, no counterpart in the reference (which raises Python
 * exceptions from CuPy): the message of this thread's last failed call, and
 * the ABI version the loader checks. */
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

/* As htamd_prop_apply with field b read at in + b*in_stride and written at
 * out + b*out_stride (strides in complex elements, >= nz*n), e.g. the slice
 * x[:, j] of a [batch, ndist, nz, n] array (stride ndist*nz*n).  Replaces the
 * per-distance D / DT calls on array slices of the reference cascade
 * (src/holotomocupy/rec_mpi.py:578-580, 593-598, 638-639). */
int htamd_prop_apply_strided(const htamd_prop_plan* plan, const void* in, int64_t in_stride, void* out,
                             int64_t out_stride, int64_t batch, int64_t j, int adjoint, void* stream);

/* One axis of D_j / DT_j: axis 0 = the row (x) pass in -> out, axis 1 = the
 * column (y) pass, in place (in must equal out).  D_j = pass(axis 1) after
 * pass(axis 0); the two passes commute.  Replaces the row / column stages of
 * conv2d_run (src/holotomocupy/cuda/conv2d.cu:144-154). */
int htamd_prop_pass(const htamd_prop_plan* plan, const void* in, int64_t in_stride, void* out,
                    int64_t out_stride, int64_t batch, int64_t j, int adjoint, int axis, void* stream);

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

/* Bytes of device workspace htamd_bspline_d1_adj needs for `batch` angles
 * (the per-workgroup partial sums of dcurlySadjc's out2, shift.py:185-204; the
 * reference allocates its temporaries inside the call). */
int htamd_bspline_d1_adj_workspace_size(const htamd_shift_plan* plan, int64_t batch, int64_t* bytes);

/* Adjoint of the differential: out1[batch, nzpsi, npsi] = Sadj(g), and
 * out2[batch, 2] = (Re<g, dt1>, Re<g, dt2>), reduced deterministically.
 * Replaces Shift.dcurlySadjc (shift.py:185-204) and dsadj / dsadjf
 * (cuda_kernels.py:690-843) plus its redot reductions. */
int htamd_bspline_d1_adj(const htamd_shift_plan* plan, const void* c, const void* g, const float* r,
                         int64_t r_stride, const float* m, int64_t m_stride, void* out1, float* out2,
                         void* workspace, int64_t workspace_bytes, int64_t batch, void* stream);

/* Allocation-free form of htamd_bspline_d1_adj.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The partial sums live in a fixed device array of the library, one per GPU,
 * of htamd_bspline_d1_adj_noalloc_max_bytes() bytes (4 MiB: 128 angles of a
 * 2048^2 detector); a batch that needs more returns HTAMD_ERR_UNSUPPORTED.  It
 * calls no hipMalloc / hipFree.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (different streams or threads) race.  Results
 * equal htamd_bspline_d1_adj's bit for bit. */
int htamd_bspline_d1_adj_noalloc_max_bytes(int64_t* bytes);
int htamd_bspline_d1_adj_noalloc(const htamd_shift_plan* plan, const void* c, const void* g, const float* r,
                                 int64_t r_stride, const float* m, int64_t m_stride, void* out1, float* out2,
                                 int64_t batch, void* stream);

/* Magnified back-shift of the Paganin initial guess (complex64 only):
 * c_out[batch, nzpsi, npsi] = Sback(g, r, m), g [batch, nz, n], the stretched
 * spline phi(t/m) of support |t| < 2m, computed as a deterministic gather.
 * Replaces Shift.Sback (src/holotomocupy/shift.py:138-149) and the sback kernel
 * (src/holotomocupy/cuda_kernels.py:294-368). */
int htamd_bspline_sback(const htamd_shift_plan* plan, const void* g, const float* r, int64_t r_stride,
                        const float* m, int64_t m_stride, void* c_out, int64_t batch, void* stream);

/* out = IFFT2(FFT2(in) / (D_y (x) D_x)), D(f) = phi(0) + 2 sum_{k=1..4} phi(k/m)
 * cos(2 pi k f), in / out [batch, nzpsi, npsi] complex64 (may alias); npsi and
 * nzpsi powers of two in [8, 2048].  Replaces Shift.coeffback
 * (src/holotomocupy/shift.py:84-96). */
int htamd_bspline_coeffback(const htamd_shift_plan* plan, const void* in, void* out, int64_t batch, double m,
                            void* stream);

/* ---- pointwise cascade stages and reductions -------------------------------
 * Complex arrays are complex64; per-angle fields are [batch, ndist, npix]
 * (npix = nz*n), probe-like arrays [ndist, npix]. */

/* out[b,j,i] = scale * sum_t p_t[j,i] q_t[b,j,i] over the non-NULL pairs.
 * Replaces the probe products of F1 / dF1 / d2F_dF1
 * (src/holotomocupy/rec_mpi.py:572-623). */
int htamd_cas_probe_combo(void* out, const void* p0, const void* q0, const void* p1, const void* q1,
                          const void* p2, const void* q2, const void* p3, const void* q3,
                          int64_t batch, int64_t ndist, int64_t npix, float scale, void* stream);

/* Pointwise exp stages of F2 (rec_mpi.py:645-724: _F2_fused, _dF2_fused,
 * _d2F_dF2_fused, _gF2_fused) on `count` elements; x22 / y / z / w are
 * complex64 if x_complex else float32 (in mode 3 y is the complex y12):
 *   mode 0: out1 = exp(i x22)
 *   mode 1: out1 = exp(i x22), out2 = out1 * i * y        (out2 may be NULL)
 *   mode 2: out1 = exp(i x22) * (-y z) [+ exp(i x22) * i * w if w != NULL]
 *   mode 3: out1 = -i * y * conj(exp(i x22)); real_out keeps the real part
 *           (float32 output) as for a float32 object. */
int htamd_cas_expi(int mode, const void* x22, const void* y, const void* z, const void* w, void* out1,
                   void* out2, int64_t count, int x_complex, int real_out, void* stream);

/* out = scale * (x - d x/|x|): the data-term gradient _gF0_fused
 * (rec_mpi.py:555-569); d is float32. */
int htamd_cas_residual(const void* x, const float* d, void* out, int64_t count, float scale,
                       void* stream);

/* Tail of gF1 (rec_mpi.py:626-643): y11[j] (+)= sum_b y12[b,j] conj(x12[b,j]),
 * then y12[b,j] *= conj(x11[j]) in place; accumulate != 0 adds to y11. */
int htamd_cas_probe_grad(void* y12, const void* x12, const void* x11, void* y11, int64_t batch,
                         int64_t ndist, int64_t npix, int accumulate, void* stream);

/* Deterministic FP64 reductions over `count` complex elements into result[0]
 * (device double): kind 0: sum (|x|-d)^2 (F0, rec_mpi.py:517-527);
 * kind 1: sum Re((x - d x/|x|) conj y) (dF0, rec_mpi.py:529-537);
 * kind 2: sum (1-d/|x|) Re(y conj z) + (d/|x|) Re(l conj y) Re(l conj z)
 *         [+ Re((x - d l) conj w)], l = x/|x| (d2F_dF0, rec_mpi.py:539-553). */
int htamd_cas_reduce_workspace_size(int64_t count, int64_t* bytes);
int htamd_cas_reduce(int kind, const void* x, const float* d, const void* y, const void* z,
                     const void* w, int64_t count, double* result, void* workspace,
                     int64_t workspace_bytes, void* stream);

/* ---- fused BH gradient ------------------------------------------------------
 * One BH gradient evaluation of the data term for `nangles` angles, i.e. what
 * Rec.gradients_cascade computes for a chunk (src/holotomocupy/rec_mpi.py:327-353
 * with the cascade 517-813), in three FFT passes per (angle, distance) and no
 * per-operator intermediates.  The plan binds a propagator plan (n, nz, ndist
 * distances) and a shift plan (object grid nzpsi x npsi, object dtype); both
 * must outlive it.  Inputs (device):
 *   proj [nangles, nzpsi, npsi] object dtype, pos [nangles, ndist, 2] (y, x),
 *   mag [nangles, ndist] (eff_demagnifications), prb [nprobes, ndist, nz, n],
 *   probe_index [nangles] int32 (NULL: every angle uses probe 0; angles of one
 *   probe must be contiguous unless accumulate_prb), data [nangles, ndist, nz, n]
 *   float32 = sqrt(intensity), angle_scales [nangles, 4] float32 =
 *   (2/data_size, rho_obj^2, rho_pos^2, 0), rho_prb [nprobes] = rho_prb^2.
 * Outputs: gproj like proj, gpos like pos, gprb like prb (overwritten, or
 * added to when accumulate_prb != 0), objective (nullable, device double) =
 * sum (|x0| - d)^2 over the chunk.  Deterministic (no atomics). */
typedef struct htamd_bh_plan htamd_bh_plan;
int htamd_bh_plan_create(const htamd_prop_plan* prop, const htamd_shift_plan* shift, htamd_bh_plan** plan);
int htamd_bh_plan_destroy(htamd_bh_plan* plan);
int htamd_bh_workspace_size(const htamd_bh_plan* plan, int64_t nangles, int64_t* bytes);
/* Profiling hook (no counterpart in the reference): events[2k] and
 * events[2k+1] (hipEvent_t, caller-created; NULL entries skipped) are recorded
 * on the launch stream before and after stage k of every later
 * htamd_bh_gradient call on this plan: 0 prefilter, 1 rows forward, 2 columns,
 * 3 rows adjoint, 4 Sadj, 5 prefilter of the object gradient, 6 reductions.
 * count = 0 switches it off.  Not for concurrent use of one plan. */
int htamd_bh_set_stage_events(htamd_bh_plan* plan, void** events, int count);
/* Replaces Rec.gradients_cascade / _gradients_cascade and the gF0..gF3 chain
 * (src/holotomocupy/rec_mpi.py:327-353, 555-813). */
int htamd_bh_gradient(const htamd_bh_plan* plan, int64_t nangles, const void* proj, const float* pos,
                      const float* mag, const void* prb, const int* probe_index, const float* data,
                      const float* angle_scales, const float* rho_prb, int accumulate_prb, void* gproj,
                      float* gpos, void* gprb, double* objective, void* workspace,
                      int64_t workspace_bytes, void* stream);

/* As htamd_bh_gradient with the per-request arrays addressed through a table instead of packed:
 * tables is a device array of 3 * nangles + 2 * nprobes device pointers,
 *   [proj of angle 0 .. nangles-1 | data of each angle | gproj of each angle |
 *    prb of probe slot 0 .. nprobes-1 | gprb of each probe slot],
 * each pointing at that angle's [nzpsi, npsi] plane ([ndist, nz, n] for data) or that slot's
 * [ndist, nz, n] probe, so a batch of requests is read from and written to the requests' own
 * buffers; pos, mag, probe_index, angle_scales, rho_prb and gpos stay packed.  No reference
 * counterpart (the reference evaluates one chunk of one array, rec_mpi.py:327-353); it exists for
 * the serving loop. */
int htamd_bh_gradient_tab(const htamd_bh_plan* plan, int64_t nangles, int64_t nprobes, const void* const* tables,
                          const float* pos, const float* mag, const int* probe_index, const float* angle_scales,
                          const float* rho_prb, int accumulate_prb, float* gpos, double* objective, void* workspace,
                          int64_t workspace_bytes, void* stream);

/* Fused BH bilinear form of the data term.  Replaces Rec.hessian_cascade / _hessian_cascade and the
 * dF1..dF3, d2F_dF0..d2F_dF3 chain (src/holotomocupy/rec_mpi.py:250-303, 517-813) for one chunk of
 * angles: result (+)= scale * sum [(1 - d/|x0|) Re(conj(y0) z0) + d/|x0|^3 Re(conj(x0) y0) Re(conj(x0) z0)
 * + (1 - d/|x0|) Re(conj(x0) w0)], with x0, y0, z0, w0 the propagated fields of the point and of the
 * first- and second-order perturbations along directions y = (prb_y, proj_y, pos_y) and
 * z = (prb_z, proj_z, pos_z); scale = 2/data_size.  proj_z, pos_z, prb_z all NULL means z = y (the
 * reference's y_is_z branch: three propagations instead of four).  The second-order shift term keeps
 * the reference's d2s pairing (cuda_kernels.py:377-450).  Layouts as htamd_bh_gradient (one probe per
 * distance); result is one device double.  Deterministic (fixed-order FP64 partial sums).  The
 * workspace holds the prefiltered projections and the propagated fields of the chunk; its size
 * takes same_directions = 0 (general), 1 (z = y) or 2 (the pair form below). */
int htamd_bh_hessian_workspace_size(const htamd_bh_plan* plan, int64_t nangles, int same_directions,
                                    int64_t* bytes);
int htamd_bh_hessian(const htamd_bh_plan* plan, int64_t nangles, const void* proj, const float* pos,
                     const float* mag, const void* prb, const void* proj_y, const float* pos_y,
                     const void* prb_y, const void* proj_z, const float* pos_z, const void* prb_z,
                     const float* data, double scale, double* result, int accumulate, void* workspace,
                     int64_t workspace_bytes, void* stream);
/* The two bilinear forms of the BH loop's beta step, H(y, z) and H(z, z) (rec_mpi.py:186-188: top =
 * hessian(vars, grads, etas), bottom = hessian(vars, etas, etas)), in one pass: result[0] (+)= H(y, z),
 * result[1] (+)= H(z, z); five propagated fields serve both instead of seven.  No reference
 * counterpart as one call.  Workspace: htamd_bh_hessian_workspace_size with same_directions = 2. */
int htamd_bh_hessian_pair(const htamd_bh_plan* plan, int64_t nangles, const void* proj, const float* pos,
                          const float* mag, const void* prb, const void* proj_y, const float* pos_y,
                          const void* prb_y, const void* proj_z, const float* pos_z, const void* prb_z,
                          const float* data, double scale, double* result, int accumulate, void* workspace,
                          int64_t workspace_bytes, void* stream);

/* ---- BLAS-1 helpers of the BH iteration ----------------------------------------
 * Arrays are float streams of n floats (a complex64 array of m elements is
 * n = 2m); 16-byte aligned pointers take a vectorised path.  Reductions are deterministic
 * (fixed FP64 partial sums) into a device double; accumulate != 0 adds to it. */
/* Replaces linear_batch / mulc_batch on device arrays (src/holotomocupy/chunking.py:251-290):
 * out = a x + b y (y NULL: out = a x); out may equal x or y. */
int htamd_axpby(const void* x, const void* y, float a, float b, void* out, int64_t n, void* stream);
int htamd_blas1_workspace_size(int64_t* bytes);
/* Replaces redot / redot_batch (src/holotomocupy/utils.py:96-103, chunking.py:234-247):
 * result (+)= Re<x, y> = sum x_k y_k. */
int htamd_redot(const void* x, const void* y, int64_t n, double* result, int accumulate, void* workspace,
                int64_t workspace_bytes, void* stream);
/* Replaces linear_redot_batch (src/holotomocupy/chunking.py:264-278): x = a x + b y,
 * then result (+)= Re<y, x>, in one pass. */
int htamd_axpby_redot(void* x, const void* y, float a, float b, int64_t n, double* result, int accumulate,
                      void* workspace, int64_t workspace_bytes, void* stream);
/* Device-coefficient forms for a BH loop that keeps its step sizes on the device (rec_mpi.py:188-222
 * compute alpha and beta on the host after a synchronising .get()): ab points to two device doubles
 * (a, b), rounded to float as the host forms round their float arguments.
 * htamd_axpby_dc replaces linear_batch (chunking.py:249-262); htamd_axpby_redot_dc replaces
 * linear_redot_batch (chunking.py:264-278) and adds scale * Re<y, x> to result (with accumulate) so
 * that the loop's weighted sum  top -= Re<g_v, eta_v> / rho_v^2  (rec_mpi.py:198-205) stays on the
 * device. */
int htamd_axpby_dc(const void* x, const void* y, const double* ab, void* out, int64_t n, void* stream);
int htamd_axpby_redot_dc(void* x, const void* y, const double* ab, int64_t n, double* result, double scale,
                         int accumulate, void* workspace, int64_t workspace_bytes, void* stream);

/* ---- 3-D Laplacian of the regulariser -----------------------------------------
 * Replaces `_lap` and the stencil steps of gradient_laplacian / hessian_laplacian / _lap_energy_local
 * (src/holotomocupy/rec_mpi.py:418-494): out[k] = add[k] + scale * lap(in)[z0 + k] for
 * k < nz_out, with the 7-point stencil periodic in y and x and zero outside slices [0, nz_in) in z.
 * in: [nz_in, ny, nx] float32 (or complex64 if is_complex); add may be NULL or alias out; out must
 * not alias in. */
int htamd_laplacian3d(const void* in, int64_t nz_in, int64_t ny, int64_t nx, int is_complex, int64_t z0,
                      int64_t nz_out, const void* add, float scale, void* out, void* stream);

/* ---- USFFT Radon transform ---------------------------------------------------
 * Geometry of Tomo(n, nz, theta, mask_r) (src/holotomocupy/tomo.py:12-54): slices
 * n x n, n a power of two in [8, 1024], ntheta angles (radians; rounded to
 * float32 as the reference stores them), mask [n, n] float64 support mask
 * (NULL = 1; the smooth circular mask of tomo.py:34-44 is built by the caller).
 * eps = 1e-3, mu and the gridding half-width m follow tomo.py:14-16. */
typedef struct htamd_tomo_plan htamd_tomo_plan;
int htamd_tomo_plan_create(int64_t n, int64_t ntheta, const double* theta, const double* mask,
                           htamd_tomo_plan** plan);
int htamd_tomo_plan_destroy(htamd_tomo_plan* plan);
/* Device scratch for nz slices: nz (6 n^2 + ntheta n) complex64 (the 2n x 2n
 * frequency grid, one half-transformed n x 2n stage and the transformed
 * sinogram of the adjoint). */
int htamd_tomo_workspace_size(const htamd_tomo_plan* plan, int64_t nz, int64_t* bytes);
/* Replaces Tomo.R (src/holotomocupy/tomo.py:56-86) with the gather kernel
 * (src/holotomocupy/cuda_kernels.py:3-76, dir 0): obj [nz, n, n] -> out
 * [ntheta, nz, n], float32 (real part kept, as for a float32 object) or
 * complex64 when is_complex. */
int htamd_tomo_R(const htamd_tomo_plan* plan, const void* obj, int is_complex, void* out, int64_t nz,
                 void* workspace, int64_t workspace_bytes, void* stream);
/* Replaces Tomo.RT (src/holotomocupy/tomo.py:88-116) with the scatter of
 * cuda_kernels.py:3-76 (dir 1) computed as a deterministic gather over the
 * transposed taps: data [ntheta, nz, n] -> out [nz, n, n]; dtypes as for
 * htamd_tomo_R. */
int htamd_tomo_RT(const htamd_tomo_plan* plan, const void* data, int is_complex, void* out, int64_t nz,
                  void* workspace, int64_t workspace_bytes, void* stream);

/* Detector-axis filter of filtered back-projection: out = IFFT(h FFT(data)) along the last axis of
 * data [ntheta, nz, n] (float32: real part kept, as the reference does for a float32 sinogram), with
 * h on fftfreq(n): filter 0 = ramp 4n|f|, 1 = Shepp-Logan 4n|f| sinc(f), 2 = Parzen 4n|f| w(u) with the
 * reference's u = 2 (4n|f|).  h is built in FP64 per plan.  out may equal data; no workspace.
 * Replaces Tomo._filter_sino (src/holotomocupy/tomo.py:118-162). */
int htamd_tomo_filter(const htamd_tomo_plan* plan, const void* data, int is_complex, int filter, void* out,
                      int64_t nz, void* stream);
/* Filtered back-projection out [nz, n, n] = RT(filter(data)) sqrt(n / ntheta) in the passes of RT alone:
 * the filter multiplies the bins of RT's own line transform (shifted by n/2 for its c1d modulation), so
 * the sinogram is transformed once.  Workspace as htamd_tomo_RT.
 * Replaces Tomo.fbp (src/holotomocupy/tomo.py:164-186). */
int htamd_tomo_fbp(const htamd_tomo_plan* plan, const void* data, int is_complex, int filter, void* out,
                   int64_t nz, void* workspace, int64_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* HTAMD_H */
