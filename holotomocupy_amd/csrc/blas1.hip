// Streaming BLAS-1 kernels of the BH iteration helpers (redot_batch,
// linear_batch, linear_redot_batch, mulc_batch; src/holotomocupy/chunking.py:234-290,
// utils.py redot).  Arrays are handled as float streams: a complex64 array of
// n elements is 2n floats, and Re<x, y> = sum x_k y_k over those floats.
// Reductions are deterministic: fixed per-block partial sums in FP64, then one
// block sums the partials in order.
#include <string>

#include "../../include/htamd.h"
#include "htamd_common.hpp"
#include "htamd_internal.hpp"

namespace htamd {

constexpr int kB1Threads = 256;
constexpr int kB1Blocks = 1024;

static unsigned b1_grid(int64_t n4) {
  int64_t g = (n4 + kB1Threads - 1) / kB1Threads;
  return (unsigned)(g < 1 ? 1 : (g > 8192 ? 8192 : g));
}

// out = a x + b y (y may be NULL: out = a x); n floats, float4 body + tail
// (VEC = false: every pointer handled as plain floats, for unaligned views).
// ab (device, may be NULL): the coefficients are read there instead, as
// doubles rounded to float, so a solver can keep them on the device.
template <bool VEC>
__global__ __launch_bounds__(kB1Threads) void b1_axpby_k(const float* __restrict__ x,
                                                          const float* __restrict__ y, float a, float b,
                                                          const double* __restrict__ ab,
                                                          float* __restrict__ out, int64_t n) {
  if (ab) {
    a = (float)ab[0];
    b = (float)ab[1];
  }
  if (!VEC) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
      float r = a * x[k];
      if (y) r = fmaf(b, y[k], r);
      out[k] = r;
    }
    return;
  }
  const int64_t n4 = n / 4;
  const float4* x4 = (const float4*)x;
  const float4* y4 = (const float4*)y;
  float4* o4 = (float4*)out;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n4; k += (int64_t)gridDim.x * blockDim.x) {
    const float4 u = x4[k];
    float4 r = make_float4(a * u.x, a * u.y, a * u.z, a * u.w);
    if (y) {
      const float4 v = y4[k];
      r = make_float4(fmaf(b, v.x, r.x), fmaf(b, v.y, r.y), fmaf(b, v.z, r.z), fmaf(b, v.w, r.w));
    }
    o4[k] = r;
  }
  const int64_t k = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0 && k < n) {
    float r = a * x[k];
    if (y) r = fmaf(b, y[k], r);
    out[k] = r;
  }
}

// partial[block] = sum over the block's share of x_k * y_k, where x is
// (if AXPBY) first replaced by a x + b y
template <bool AXPBY, bool VEC>
__global__ __launch_bounds__(kB1Threads) void b1_dot_partial_k(float* __restrict__ x, const float* __restrict__ y,
                                                                float a, float b, const double* __restrict__ ab,
                                                                int64_t n, double* __restrict__ partial) {
  if (AXPBY && ab) {
    a = (float)ab[0];
    b = (float)ab[1];
  }
  double s = 0.0;
  const int64_t n4 = VEC ? n / 4 : 0;
  float4* x4 = (float4*)x;
  const float4* y4 = (const float4*)y;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n4; k += (int64_t)gridDim.x * blockDim.x) {
    float4 u = x4[k];
    const float4 v = y4[k];
    if (AXPBY) {
      u = make_float4(fmaf(b, v.x, a * u.x), fmaf(b, v.y, a * u.y), fmaf(b, v.z, a * u.z), fmaf(b, v.w, a * u.w));
      x4[k] = u;
    }
    s += (double)(u.x * v.x + u.y * v.y) + (double)(u.z * v.z + u.w * v.w);
  }
  // tail (or, unvectorised, everything) as plain floats
  for (int64_t k = 4 * n4 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x) {
    float u = x[k];
    if (AXPBY) {
      u = fmaf(b, y[k], a * u);
      x[k] = u;
    }
    s += (double)(u * y[k]);
  }
  __shared__ double red[kB1Threads];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kB1Threads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// out (+)= scale * sum of the partials, in a fixed order
__global__ __launch_bounds__(kB1Threads) void b1_dot_final_k(const double* __restrict__ partial, int nparts,
                                                              int accumulate, double scale,
                                                              double* __restrict__ out) {
  __shared__ double red[kB1Threads];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kB1Threads) s += partial[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kB1Threads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = accumulate ? out[0] + scale * red[0] : scale * red[0];
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

static int b1_dot(bool axpby, void* x, const void* y, float a, float b, const double* ab, int64_t n,
                  double* result, int accumulate, double scale, void* ws, int64_t ws_bytes, void* stream,
                  const char* what) {
  if (n < 0 || !result || (n > 0 && (!x || !y))) return fail(HTAMD_ERR_ARG, std::string(what) + ": bad arguments");
  if (!ws || ws_bytes < (int64_t)kB1Blocks * (int64_t)sizeof(double))
    return fail(HTAMD_ERR_ARG, std::string(what) + ": workspace too small");
  hipStream_t st = (hipStream_t)stream;
  const bool vec = aligned16(x) && aligned16(y);
  int64_t g = ((vec ? n / 4 : n) + kB1Threads - 1) / kB1Threads;
  g = g < 1 ? 1 : (g > kB1Blocks ? kB1Blocks : g);
#define HTAMD_B1DOT(AX, V)                                                                                   \
  hipLaunchKernelGGL((b1_dot_partial_k<AX, V>), dim3((unsigned)g), dim3(kB1Threads), 0, st, (float*)x, \
                     (const float*)y, a, b, ab, n, (double*)ws)
  if (axpby) {
    if (vec) HTAMD_B1DOT(true, true); else HTAMD_B1DOT(true, false);
  } else {
    if (vec) HTAMD_B1DOT(false, true); else HTAMD_B1DOT(false, false);
  }
#undef HTAMD_B1DOT
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, what);
  hipLaunchKernelGGL(b1_dot_final_k, dim3(1), dim3(kB1Threads), 0, st, (const double*)ws, (int)g, accumulate,
                     scale, result);
  e = hipGetLastError();
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, what);
}

}  // namespace htamd

using namespace htamd;

extern "C" {

static int b1_axpby(const void* x, const void* y, float a, float b, const double* ab, void* out, int64_t n,
                    void* stream, const char* what) {
  if (n < 0 || (n > 0 && (!x || !out))) return fail(HTAMD_ERR_ARG, std::string(what) + ": bad arguments");
  if (n == 0) return HTAMD_OK;
  if (aligned16(x) && aligned16(out) && (!y || aligned16(y)))
    hipLaunchKernelGGL(b1_axpby_k<true>, dim3(b1_grid(n / 4)), dim3(kB1Threads), 0, (hipStream_t)stream,
                       (const float*)x, (const float*)y, a, b, ab, (float*)out, n);
  else
    hipLaunchKernelGGL(b1_axpby_k<false>, dim3(b1_grid(n)), dim3(kB1Threads), 0, (hipStream_t)stream,
                       (const float*)x, (const float*)y, a, b, ab, (float*)out, n);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, what);
}

int htamd_axpby(const void* x, const void* y, float a, float b, void* out, int64_t n, void* stream) {
  return b1_axpby(x, y, a, b, nullptr, out, n, stream, "axpby");
}

int htamd_axpby_dc(const void* x, const void* y, const double* ab, void* out, int64_t n, void* stream) {
  if (!ab) return fail(HTAMD_ERR_ARG, "axpby_dc: NULL coefficients");
  return b1_axpby(x, y, 0.0f, 0.0f, ab, out, n, stream, "axpby_dc");
}

int htamd_blas1_workspace_size(int64_t* bytes) {
  if (!bytes) return fail(HTAMD_ERR_ARG, "blas1_workspace_size: NULL");
  *bytes = (int64_t)kB1Blocks * (int64_t)sizeof(double);
  return HTAMD_OK;
}

int htamd_redot(const void* x, const void* y, int64_t n, double* result, int accumulate, void* workspace,
                int64_t workspace_bytes, void* stream) {
  return b1_dot(false, (void*)x, y, 1.0f, 0.0f, nullptr, n, result, accumulate, 1.0, workspace, workspace_bytes,
                stream, "redot");
}

int htamd_axpby_redot(void* x, const void* y, float a, float b, int64_t n, double* result, int accumulate,
                      void* workspace, int64_t workspace_bytes, void* stream) {
  return b1_dot(true, x, y, a, b, nullptr, n, result, accumulate, 1.0, workspace, workspace_bytes, stream,
                "axpby_redot");
}

int htamd_axpby_redot_dc(void* x, const void* y, const double* ab, int64_t n, double* result, double scale,
                         int accumulate, void* workspace, int64_t workspace_bytes, void* stream) {
  if (!ab) return fail(HTAMD_ERR_ARG, "axpby_redot_dc: NULL coefficients");
  return b1_dot(true, x, y, 0.0f, 0.0f, ab, n, result, accumulate, scale, workspace, workspace_bytes, stream,
                "axpby_redot_dc");
}

}  // extern "C"
