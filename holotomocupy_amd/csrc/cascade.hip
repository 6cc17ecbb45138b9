// Pointwise stages and reductions of the BH cascade (src/holotomocupy/rec_mpi.py:517-813):
// the reference's cp.fuse kernels (_F0_fused ... _gF2_fused), the probe products of
// F1 / dF1 / d2F_dF1, the angle sum of gF1 and the redot / sum reductions.
// Reductions are deterministic: fixed grid, fixed per-thread order, FP64 partials.
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/htamd.h"
#include "htamd_bspline.hpp"
#include "htamd_internal.hpp"

namespace htamd {

constexpr int kEwThreads = 256;

static unsigned ew_grid(int64_t count) {
  int64_t g = (count + kEwThreads - 1) / kEwThreads;
  if (g > 65536) g = 65536;
  return (unsigned)(g < 1 ? 1 : g);
}

// out[b, j, i] = scale * sum_t p_t[j, i] q_t[b, j, i]
__global__ __launch_bounds__(kEwThreads) void cas_probe_combo_k(
    float2* __restrict__ out, const float2* __restrict__ p0, const float2* __restrict__ q0,
    const float2* __restrict__ p1, const float2* __restrict__ q1, const float2* __restrict__ p2,
    const float2* __restrict__ q2, const float2* __restrict__ p3, const float2* __restrict__ q3,
    int64_t batch, int64_t ndist, int64_t npix, float scale) {
  const int64_t total = batch * ndist * npix;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int64_t pk = k % (ndist * npix);
    float2 acc = make_float2(0.f, 0.f);
    if (p0) acc = cadd(acc, cmul(p0[pk], q0[k]));
    if (p1) acc = cadd(acc, cmul(p1[pk], q1[k]));
    if (p2) acc = cadd(acc, cmul(p2[pk], q2[k]));
    if (p3) acc = cadd(acc, cmul(p3[pk], q3[k]));
    out[k] = cscale(acc, scale);
  }
}

template <typename T> __device__ __forceinline__ float2 as_c(T v);
template <> __device__ __forceinline__ float2 as_c<float>(float v) { return make_float2(v, 0.f); }
template <> __device__ __forceinline__ float2 as_c<float2>(float2 v) { return v; }


// mode 0: out1 = e^{ix}; 1: out1 = e^{ix}, out2 = e^{ix} i y; 2: out1 = e^{ix}(-y z) [+ e^{ix} i w];
// 3: out1 = -i y12 conj(e^{ix}) (TO = float keeps the real part)
template <typename TX, typename TO>
__global__ __launch_bounds__(kEwThreads) void cas_expi_k(int mode, const TX* __restrict__ x,
                                                         const void* __restrict__ yv,
                                                         const TX* __restrict__ z,
                                                         const TX* __restrict__ w,
                                                         TO* __restrict__ out1,
                                                         float2* __restrict__ out2, int64_t count) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x) {
    const float2 e = expi(as_c(x[k]));
    float2 r;
    if (mode == 0) {
      r = e;
    } else if (mode == 1) {
      const float2 y = as_c(((const TX*)yv)[k]);
      if (out2) out2[k] = cmul(e, make_float2(-y.y, y.x));
      r = e;
    } else if (mode == 2) {
      const float2 y = as_c(((const TX*)yv)[k]);
      const float2 zz = as_c(z[k]);
      const float2 yz = cmul(y, zz);
      r = cmul(e, make_float2(-yz.x, -yz.y));
      if (w) {
        const float2 ww = as_c(w[k]);
        r = cadd(r, cmul(e, make_float2(-ww.y, ww.x)));
      }
    } else {
      const float2 y12 = ((const float2*)yv)[k];
      const float2 t = cmulc(y12, e);  // y12 conj(e)
      r = make_float2(t.y, -t.x);      // * (-i)
    }
    if constexpr (sizeof(TO) == sizeof(float)) {
      out1[k] = r.x;
    } else {
      out1[k] = r;
    }
  }
}

// y0 = scale (x - d x/|x|)
__global__ __launch_bounds__(kEwThreads) void cas_residual_k(const float2* __restrict__ x,
                                                             const float* __restrict__ d,
                                                             float2* __restrict__ out, int64_t count,
                                                             float scale) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = x[k];
    const float a = sqrtf(v.x * v.x + v.y * v.y);
    const float f = 1.0f - d[k] / a;
    out[k] = make_float2(scale * v.x * f, scale * v.y * f);
  }
}

// y11[j, i] (+)= sum_b y12[b, j, i] conj(x12[b, j, i]);  y12[b, j, i] *= conj(x11[j, i])
__global__ __launch_bounds__(kEwThreads) void cas_probe_grad_k(float2* __restrict__ y12,
                                                               const float2* __restrict__ x12,
                                                               const float2* __restrict__ x11,
                                                               float2* __restrict__ y11,
                                                               int64_t batch, int64_t ndist,
                                                               int64_t npix, int accumulate) {
  const int64_t plane = ndist * npix;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < plane;
       k += (int64_t)gridDim.x * blockDim.x) {
    const float2 p = x11[k];
    float2 acc = accumulate ? y11[k] : make_float2(0.f, 0.f);
    for (int64_t b = 0; b < batch; ++b) {
      const float2 v = y12[b * plane + k];
      acc = cadd(acc, cmulc(v, x12[b * plane + k]));
      y12[b * plane + k] = cmulc(v, p);
    }
    y11[k] = acc;
  }
}

// ---- deterministic reductions ----------------------------------------------
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double cas_term(int kind, const float2* x, const float* d, const float2* y,
                                           const float2* z, const float2* w, int64_t k) {
  const float2 v = x[k];
  const float a = sqrtf(v.x * v.x + v.y * v.y);
  if (kind == 0) {
    const float t = a - d[k];
    return (double)(t * t);
  }
  const float2 l = make_float2(v.x / a, v.y / a);
  if (kind == 1) {
    const float2 g = make_float2(v.x - d[k] * l.x, v.y - d[k] * l.y);
    return (double)vredot(g, y[k]);
  }
  const float d0 = d[k] / a;
  const float2 yy = y[k], zz = z[k];
  float t = (1.0f - d0) * vredot(yy, zz) + d0 * vredot(l, yy) * vredot(l, zz);
  if (w) t += vredot(make_float2(v.x - d[k] * l.x, v.y - d[k] * l.y), w[k]);
  return (double)t;
}

__global__ __launch_bounds__(kEwThreads) void cas_reduce_partial_k(int kind, const float2* x,
                                                                   const float* d, const float2* y,
                                                                   const float2* z, const float2* w,
                                                                   int64_t count,
                                                                   double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x)
    s += cas_term(kind, x, d, y, z, w, k);
  __shared__ double red[kEwThreads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kEwThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kEwThreads) void cas_reduce_final_k(const double* __restrict__ partial,
                                                                 int nparts, double* __restrict__ out) {
  __shared__ double red[kEwThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kEwThreads) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kEwThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

static int launch_status(const char* what) {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, what);
}

}  // namespace htamd

using namespace htamd;

extern "C" {

int htamd_cas_probe_combo(void* out, const void* p0, const void* q0, const void* p1, const void* q1,
                          const void* p2, const void* q2, const void* p3, const void* q3,
                          int64_t batch, int64_t ndist, int64_t npix, float scale, void* stream) {
  if (!out || batch < 0 || ndist < 1 || npix < 1)
    return fail(HTAMD_ERR_ARG, "cas_probe_combo: bad arguments");
  if ((!p0) != (!q0) || (!p1) != (!q1) || (!p2) != (!q2) || (!p3) != (!q3))
    return fail(HTAMD_ERR_ARG, "cas_probe_combo: each term needs both factors");
  const int64_t total = batch * ndist * npix;
  if (total == 0) return HTAMD_OK;
  hipLaunchKernelGGL(cas_probe_combo_k, dim3(ew_grid(total)), dim3(kEwThreads), 0, (hipStream_t)stream,
                     (float2*)out, (const float2*)p0, (const float2*)q0, (const float2*)p1,
                     (const float2*)q1, (const float2*)p2, (const float2*)q2, (const float2*)p3,
                     (const float2*)q3, batch, ndist, npix, scale);
  return launch_status("cas_probe_combo");
}

int htamd_cas_expi(int mode, const void* x22, const void* y, const void* z, const void* w, void* out1,
                   void* out2, int64_t count, int x_complex, int real_out, void* stream) {
  if (mode < 0 || mode > 3) return fail(HTAMD_ERR_ARG, "cas_expi: mode must be 0..3");
  if (count < 0 || !x22 || !out1) return fail(HTAMD_ERR_ARG, "cas_expi: bad arguments");
  if ((mode >= 1 && !y) || (mode == 2 && !z)) return fail(HTAMD_ERR_ARG, "cas_expi: missing operand");
  if (real_out && mode != 3) return fail(HTAMD_ERR_ARG, "cas_expi: real output only for mode 3");
  if (count == 0) return HTAMD_OK;
  hipStream_t st = (hipStream_t)stream;
  const unsigned g = ew_grid(count);
  if (x_complex) {
    hipLaunchKernelGGL((cas_expi_k<float2, float2>), dim3(g), dim3(kEwThreads), 0, st, mode,
                       (const float2*)x22, y, (const float2*)z, (const float2*)w, (float2*)out1,
                       (float2*)out2, count);
  } else if (real_out) {
    hipLaunchKernelGGL((cas_expi_k<float, float>), dim3(g), dim3(kEwThreads), 0, st, mode,
                       (const float*)x22, y, (const float*)z, (const float*)w, (float*)out1,
                       (float2*)out2, count);
  } else {
    hipLaunchKernelGGL((cas_expi_k<float, float2>), dim3(g), dim3(kEwThreads), 0, st, mode,
                       (const float*)x22, y, (const float*)z, (const float*)w, (float2*)out1,
                       (float2*)out2, count);
  }
  return launch_status("cas_expi");
}

int htamd_cas_residual(const void* x, const float* d, void* out, int64_t count, float scale,
                       void* stream) {
  if (count < 0 || !x || !d || !out) return fail(HTAMD_ERR_ARG, "cas_residual: bad arguments");
  if (count == 0) return HTAMD_OK;
  hipLaunchKernelGGL(cas_residual_k, dim3(ew_grid(count)), dim3(kEwThreads), 0, (hipStream_t)stream,
                     (const float2*)x, d, (float2*)out, count, scale);
  return launch_status("cas_residual");
}

int htamd_cas_probe_grad(void* y12, const void* x12, const void* x11, void* y11, int64_t batch,
                         int64_t ndist, int64_t npix, int accumulate, void* stream) {
  if (!y12 || !x12 || !x11 || !y11 || batch < 0 || ndist < 1 || npix < 1)
    return fail(HTAMD_ERR_ARG, "cas_probe_grad: bad arguments");
  hipLaunchKernelGGL(cas_probe_grad_k, dim3(ew_grid(ndist * npix)), dim3(kEwThreads), 0,
                     (hipStream_t)stream, (float2*)y12, (const float2*)x12, (const float2*)x11,
                     (float2*)y11, batch, ndist, npix, accumulate);
  return launch_status("cas_probe_grad");
}

int htamd_cas_reduce_workspace_size(int64_t count, int64_t* bytes) {
  if (!bytes) return fail(HTAMD_ERR_ARG, "cas_reduce_workspace_size: NULL");
  (void)count;
  *bytes = (int64_t)kRedBlocks * (int64_t)sizeof(double);
  return HTAMD_OK;
}

int htamd_cas_reduce(int kind, const void* x, const float* d, const void* y, const void* z,
                     const void* w, int64_t count, double* result, void* workspace,
                     int64_t workspace_bytes, void* stream) {
  if (kind < 0 || kind > 2) return fail(HTAMD_ERR_ARG, "cas_reduce: kind must be 0..2");
  if (!x || !d || !result || count < 0 || (kind >= 1 && !y) || (kind == 2 && !z))
    return fail(HTAMD_ERR_ARG, "cas_reduce: bad arguments");
  if (!workspace || workspace_bytes < (int64_t)kRedBlocks * (int64_t)sizeof(double))
    return fail(HTAMD_ERR_ARG, "cas_reduce: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  int64_t g = (count + kEwThreads - 1) / kEwThreads;
  g = g < 1 ? 1 : (g > kRedBlocks ? kRedBlocks : g);
  hipLaunchKernelGGL(cas_reduce_partial_k, dim3((unsigned)g), dim3(kEwThreads), 0, st, kind,
                     (const float2*)x, d, (const float2*)y, (const float2*)z, (const float2*)w, count,
                     (double*)workspace);
  int s = launch_status("cas_reduce");
  if (s) return s;
  hipLaunchKernelGGL(cas_reduce_final_k, dim3(1), dim3(kEwThreads), 0, st, (const double*)workspace,
                     (int)g, result);
  return launch_status("cas_reduce");
}

}  // extern "C"
