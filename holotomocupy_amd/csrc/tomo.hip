// USFFT Radon transform R / RT for gfx950 (src/holotomocupy/tomo.py:12-116,
// gridding kernel src/holotomocupy/cuda_kernels.py:3-76).
//
//   R(obj)  = c1d IFFT1(c1d gather(c2d FFT2(c2d embed(obj phi))) / n) / 4
//   RT(d)   = crop(c2d IFFT2(c2d scatter(c1d FFT1(c1d d)))) phi
//
// c1d[k] = (-1)^(k+1), c2d[i,j] = (-1)^(i+j) on the 2n x 2n grid.  The 2-D
// transforms run as two LDS FFT passes (rows, then columns; only the n
// non-zero rows of the embedded slice are transformed along x), with the sign
// modulations, zero-embedding, cropping and deapodisation folded into the
// loads and stores.  The polar gather / scatter works one (angle, slice) line
// per row group with the 9x9 Gaussian taps evaluated in FP64; the line's 1-D
// FFT is fused.  The scatter uses float atomics on the grid, as the reference
// kernel does (its summation order is not fixed).
#include <cmath>
#include <string>
#include <vector>

#include "../../include/htamd.h"
#include "fft_lds.hpp"
#include "htamd_bspline.hpp"
#include "htamd_internal.hpp"

namespace htamd {

struct TomoArgs {
  int n = 0, nz = 0, ntheta = 0, m = 0;
  double mu = 0.0;
  const double* cs = nullptr;      // [ntheta] cos(theta), sin at cs[ntheta + k]
  const float* phi = nullptr;      // [n, n] deapodisation (mask and scale included)
  const float2* tw2 = nullptr;     // twiddles of length 2n
  const float2* tw1 = nullptr;     // twiddles of length n
  float2* W = nullptr;             // [nz, n, 2n]
  float2* G = nullptr;             // [nz, 2n, 2n]
};

__device__ __forceinline__ float sgnf(int k) { return (k & 1) ? -1.0f : 1.0f; }

template <typename T> __device__ __forceinline__ float2 ld_c(const T* p, int64_t i);
template <> __device__ __forceinline__ float2 ld_c<float>(const float* p, int64_t i) { return make_float2(p[i], 0.f); }
template <> __device__ __forceinline__ float2 ld_c<float2>(const float2* p, int64_t i) { return p[i]; }
template <typename T> __device__ __forceinline__ void st_c(T* p, int64_t i, float2 v);
template <> __device__ __forceinline__ void st_c<float>(float* p, int64_t i, float2 v) { p[i] = v.x; }
template <> __device__ __forceinline__ void st_c<float2>(float2* p, int64_t i, float2 v) { p[i] = v; }

// ---- forward pass A: rows of the embedded slice (length 2n) ----------------
template <int N2, int E, typename TI>
__global__ __launch_bounds__(256) void tomo_rows_fwd_k(TomoArgs a, const TI* __restrict__ obj) {
  constexpr int T = N2 / E, ROWS = 256 / T;
  __shared__ float2 lds[ROWS * fft_buf<N2>()];
  const int n = N2 / 2, h = N2 / 4;
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int64_t row = (int64_t)blockIdx.x * ROWS + r;  // (z, i'), i' < n
  const bool valid = row < (int64_t)a.nz * n;
  const int64_t z = valid ? row / n : 0;
  const int ip = valid ? (int)(row % n) : 0;
  float2* s = lds + r * fft_buf<N2>();
  FftTw<N2, E> twr;
  fft_load_twiddles<N2, E>(twr, t, a.tw2);
  const TI* src = obj + (z * n + ip) * (int64_t)n;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int p = t + e * T;
    float2 v = make_float2(0.f, 0.f);
    if (p >= h && p < h + n) {
      const int jp = p - h;
      v = cscale(ld_c<TI>(src, jp), a.phi[ip * n + jp] * sgnf(ip + h + p));
    }
    s[fpad(p)] = v;
  }
  __syncthreads();
  fft_lds<N2, E, false>(s, t, twr);
  if (valid) {
    float2* dst = a.W + (z * n + ip) * (int64_t)N2;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + e * T;
      dst[k] = cscale(s[fpad(k)], sgnf(k));
    }
  }
}

// ---- forward pass B: columns (n non-zero rows embedded in 2n) ---------------
template <int N2, int E, int C>
__global__ __launch_bounds__(C*(N2 / E)) void tomo_cols_fwd_k(TomoArgs a) {
  constexpr int T = N2 / E, NT = C * T, LS = col_stride<N2, C>();
  __shared__ float2 lds[C * LS];
  const int n = N2 / 2, h = N2 / 4;
  const int tiles = N2 / C;
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t z = tile / tiles;
  const int k0 = (int)(tile % tiles) * C;
  const int cc = threadIdx.x / T, t = threadIdx.x % T;
  FftTw<N2, E> twr;
  fft_load_twiddles<N2, E>(twr, t, a.tw2);
  const float2* src = a.W + z * (int64_t)n * N2 + k0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int idx = threadIdx.x + e * NT;
    const int c = idx % C, i = idx / C;
    lds[c * LS + fpad(i)] = (i >= h && i < h + n) ? src[(i - h) * N2 + c] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<N2, E, false>(lds + cc * LS, t, twr);
  float2* dst = a.G + z * (int64_t)N2 * N2 + k0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int idx = threadIdx.x + e * NT;
    const int c = idx % C, k = idx / C;
    dst[k * N2 + c] = cscale(lds[c * LS + fpad(k)], sgnf(k));
  }
}

// 9x9 (2m+1)^2 Gaussian taps of polar sample t of angle `th`: rows / cols on
// the 2n grid and separable FP64 weights (cuda_kernels.py:17-51).
__device__ __forceinline__ void tomo_taps(const TomoArgs& a, int th, int t, int& by, int& bx,
                                          double& x0, double& y0) {
  const int n = a.n;
  const double c = a.cs[th], sn = a.cs[a.ntheta + th];
  x0 = (double)(t - n / 2) / n * c;
  y0 = -(double)(t - n / 2) / n * sn;
  bx = (int)floor(2.0 * n * x0) - a.m;
  by = (int)floor(2.0 * n * y0) - a.m;
}

// ---- forward pass C: gather + line IFFT ------------------------------------
template <int N, int E, typename TO>
__global__ __launch_bounds__(256) void tomo_gather_k(TomoArgs a, TO* __restrict__ out) {
  constexpr int T = N / E, ROWS = 256 / T;
  __shared__ float2 lds[ROWS * fft_buf<N>()];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int64_t line = (int64_t)blockIdx.x * ROWS + r;  // (theta, z)
  const bool valid = line < (int64_t)a.ntheta * a.nz;
  const int th = valid ? (int)(line / a.nz) : 0;
  const int z = valid ? (int)(line % a.nz) : 0;
  float2* s = lds + r * fft_buf<N>();
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, a.tw1);
  const int n2 = 2 * N;
  const double cy = M_PI / a.mu, cx = -M_PI * M_PI / a.mu, inv2n = 1.0 / n2;
  const float2* g = a.G + (int64_t)z * n2 * n2;
  const int L = 2 * a.m + 1;
  for (int e = 0; e < E; ++e) {
    const int tt = t + e * T;
    int by, bx;
    double x0, y0;
    tomo_taps(a, th, tt, by, bx, x0, y0);
    float wx[16];
    for (int i0 = 0; i0 < L; ++i0) {
      const double w0 = (bx + i0) * inv2n - x0;
      wx[i0] = (float)exp(cx * w0 * w0);
    }
    float2 acc = make_float2(0.f, 0.f);
    for (int i1 = 0; i1 < L; ++i1) {
      const double w1 = (by + i1) * inv2n - y0;
      const float wy = (float)(cy * exp(cx * w1 * w1));
      const int row = (N + by + i1 + n2) % n2;
      float2 racc = make_float2(0.f, 0.f);
      for (int i0 = 0; i0 < L; ++i0) {
        const int col = (N + bx + i0 + n2) % n2;
        racc = vfma(wx[i0], g[(int64_t)row * n2 + col], racc);
      }
      acc = vfma(wy, racc, acc);
    }
    // c1d . (gather / n), c1d[k] = (-1)^(k+1)
    s[fpad(tt)] = cscale(acc, -sgnf(tt) / (float)N);
  }
  __syncthreads();
  fft_lds<N, E, true>(s, t, twr);
  if (valid) {
    TO* dst = out + ((int64_t)th * a.nz + z) * N;
    for (int e = 0; e < E; ++e) {
      const int tt = t + e * T;
      st_c<TO>(dst, tt, cscale(s[fpad(tt)], -sgnf(tt) / (4.0f * N)));
    }
  }
}

// ---- adjoint pass C': line FFT + scatter -----------------------------------
template <int N, int E, typename TI>
__global__ __launch_bounds__(256) void tomo_scatter_k(TomoArgs a, const TI* __restrict__ data) {
  constexpr int T = N / E, ROWS = 256 / T;
  __shared__ float2 lds[ROWS * fft_buf<N>()];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int64_t line = (int64_t)blockIdx.x * ROWS + r;
  const bool valid = line < (int64_t)a.ntheta * a.nz;
  const int th = valid ? (int)(line / a.nz) : 0;
  const int z = valid ? (int)(line % a.nz) : 0;
  float2* s = lds + r * fft_buf<N>();
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, a.tw1);
  const TI* src = data + ((int64_t)th * a.nz + z) * N;
  for (int e = 0; e < E; ++e) {
    const int tt = t + e * T;
    s[fpad(tt)] = valid ? cscale(ld_c<TI>(src, tt), -sgnf(tt)) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<N, E, false>(s, t, twr);
  if (!valid) return;  // no barriers below
  const int n2 = 2 * N;
  const double cy = M_PI / a.mu, cx = -M_PI * M_PI / a.mu, inv2n = 1.0 / n2;
  float2* g = a.G + (int64_t)z * n2 * n2;
  const int L = 2 * a.m + 1;
  for (int e = 0; e < E; ++e) {
    const int tt = t + e * T;
    const float2 v = cscale(s[fpad(tt)], -sgnf(tt));
    int by, bx;
    double x0, y0;
    tomo_taps(a, th, tt, by, bx, x0, y0);
    float wx[16];
    for (int i0 = 0; i0 < L; ++i0) {
      const double w0 = (bx + i0) * inv2n - x0;
      wx[i0] = (float)exp(cx * w0 * w0);
    }
    for (int i1 = 0; i1 < L; ++i1) {
      const double w1 = (by + i1) * inv2n - y0;
      const float wy = (float)(cy * exp(cx * w1 * w1));
      const int row = (N + by + i1 + n2) % n2;
      for (int i0 = 0; i0 < L; ++i0) {
        const int col = (N + bx + i0 + n2) % n2;
        const float w = wy * wx[i0];
        float2* p = g + (int64_t)row * n2 + col;
        atomicAdd(&p->x, w * v.x);
        atomicAdd(&p->y, w * v.y);
      }
    }
  }
}

// ---- adjoint pass B': columns of the grid, keep the n central rows ----------
template <int N2, int E, int C>
__global__ __launch_bounds__(C*(N2 / E)) void tomo_cols_adj_k(TomoArgs a) {
  constexpr int T = N2 / E, NT = C * T, LS = col_stride<N2, C>();
  __shared__ float2 lds[C * LS];
  const int n = N2 / 2, h = N2 / 4;
  const int tiles = N2 / C;
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t z = tile / tiles;
  const int k0 = (int)(tile % tiles) * C;
  const int cc = threadIdx.x / T, t = threadIdx.x % T;
  FftTw<N2, E> twr;
  fft_load_twiddles<N2, E>(twr, t, a.tw2);
  const float2* src = a.G + z * (int64_t)N2 * N2 + k0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int idx = threadIdx.x + e * NT;
    const int c = idx % C, k = idx / C;
    lds[c * LS + fpad(k)] = cscale(src[k * N2 + c], sgnf(k + k0 + c));
  }
  __syncthreads();
  fft_lds<N2, E, true>(lds + cc * LS, t, twr);
  float2* dst = a.W + z * (int64_t)n * N2 + k0;
  const float sc = 1.0f / N2;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int idx = threadIdx.x + e * NT;
    const int c = idx % C, i = idx / C;
    if (i >= h && i < h + n) dst[(i - h) * N2 + c] = cscale(lds[c * LS + fpad(i)], sc * sgnf(i));
  }
}

// ---- adjoint pass A': rows, crop and deapodise -------------------------------
template <int N2, int E, typename TO>
__global__ __launch_bounds__(256) void tomo_rows_adj_k(TomoArgs a, TO* __restrict__ out) {
  constexpr int T = N2 / E, ROWS = 256 / T;
  __shared__ float2 lds[ROWS * fft_buf<N2>()];
  const int n = N2 / 2, h = N2 / 4;
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int64_t row = (int64_t)blockIdx.x * ROWS + r;
  const bool valid = row < (int64_t)a.nz * n;
  const int64_t z = valid ? row / n : 0;
  const int ip = valid ? (int)(row % n) : 0;
  float2* s = lds + r * fft_buf<N2>();
  FftTw<N2, E> twr;
  fft_load_twiddles<N2, E>(twr, t, a.tw2);
  const float2* src = a.W + (z * n + ip) * (int64_t)N2;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + e * T;
    s[fpad(k)] = src[k];
  }
  __syncthreads();
  fft_lds<N2, E, true>(s, t, twr);
  if (valid) {
    TO* dst = out + (z * n + ip) * (int64_t)n;
    const float sc = 1.0f / N2;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int p = t + e * T;
      if (p >= h && p < h + n) {
        const int jp = p - h;
        st_c<TO>(dst, jp, cscale(s[fpad(p)], sc * sgnf(p) * a.phi[ip * n + jp]));
      }
    }
  }
}

__global__ void tomo_zero_k(float2* __restrict__ p, int64_t count) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * blockDim.x)
    p[k] = make_float2(0.f, 0.f);
}

// ---- host dispatch -----------------------------------------------------------
// columns per workgroup of the column passes: at most 16, at most about 70 KiB
// of LDS (8192 / N2 padded columns) and at most 1024 threads
template <int N2>
constexpr int tomo_cols_c() {
  constexpr int T = N2 / 8;
  constexpr int CT = 1024 / T;
  constexpr int CL = 8192 / N2 > 0 ? 8192 / N2 : 1;
  constexpr int C0 = CT < CL ? CT : CL;
  return C0 > 16 ? 16 : C0;
}

template <int N2, typename TIO>
static hipError_t tomo_fwd_n(const TomoArgs& a, const TIO* obj, TIO* out, hipStream_t st) {
  constexpr int E = 8, N = N2 / 2;
  constexpr int C = tomo_cols_c<N2>();
  const int64_t rows = (int64_t)a.nz * N;
  hipLaunchKernelGGL((tomo_rows_fwd_k<N2, E, TIO>), dim3((unsigned)((rows + 256 / (N2 / E) - 1) / (256 / (N2 / E)))),
                     dim3(256), 0, st, a, obj);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((tomo_cols_fwd_k<N2, E, C>), dim3((unsigned)(a.nz * (N2 / C))), dim3(C * (N2 / E)), 0, st, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  constexpr int EL = N >= 8 ? 8 : N;
  constexpr int RL = 256 / (N / EL);
  const int64_t lines = (int64_t)a.ntheta * a.nz;
  hipLaunchKernelGGL((tomo_gather_k<N, EL, TIO>), dim3((unsigned)((lines + RL - 1) / RL)), dim3(256), 0, st, a, out);
  return hipGetLastError();
}

template <int N2, typename TIO>
static hipError_t tomo_adj_n(const TomoArgs& a, const TIO* data, TIO* out, hipStream_t st) {
  constexpr int E = 8, N = N2 / 2;
  constexpr int C = tomo_cols_c<N2>();
  const int64_t gcount = (int64_t)a.nz * N2 * N2;
  int64_t zg = (gcount + 255) / 256;
  hipLaunchKernelGGL(tomo_zero_k, dim3((unsigned)(zg > 65536 ? 65536 : zg)), dim3(256), 0, st, a.G, gcount);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  constexpr int EL = N >= 8 ? 8 : N;
  constexpr int RL = 256 / (N / EL);
  const int64_t lines = (int64_t)a.ntheta * a.nz;
  hipLaunchKernelGGL((tomo_scatter_k<N, EL, TIO>), dim3((unsigned)((lines + RL - 1) / RL)), dim3(256), 0, st, a,
                     data);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL((tomo_cols_adj_k<N2, E, C>), dim3((unsigned)(a.nz * (N2 / C))), dim3(C * (N2 / E)), 0, st, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int64_t rows = (int64_t)a.nz * N;
  hipLaunchKernelGGL((tomo_rows_adj_k<N2, E, TIO>), dim3((unsigned)((rows + 256 / (N2 / E) - 1) / (256 / (N2 / E)))),
                     dim3(256), 0, st, a, out);
  return hipGetLastError();
}

template <typename TIO>
static hipError_t tomo_dispatch(bool adjoint, const TomoArgs& a, const void* in, void* out, hipStream_t st) {
#define HTAMD_TOMO_CASE(NN)                                                               \
  case NN:                                                                                \
    return adjoint ? tomo_adj_n<2 * NN, TIO>(a, (const TIO*)in, (TIO*)out, st)            \
                   : tomo_fwd_n<2 * NN, TIO>(a, (const TIO*)in, (TIO*)out, st);
  switch (a.n) {
    HTAMD_TOMO_CASE(8) HTAMD_TOMO_CASE(16) HTAMD_TOMO_CASE(32) HTAMD_TOMO_CASE(64) HTAMD_TOMO_CASE(128)
    HTAMD_TOMO_CASE(256) HTAMD_TOMO_CASE(512) HTAMD_TOMO_CASE(1024)
    default: return hipErrorInvalidValue;
  }
#undef HTAMD_TOMO_CASE
}

}  // namespace htamd

using namespace htamd;

struct htamd_tomo_plan {
  int64_t n = 0, ntheta = 0;
  int m = 0;
  double mu = 0.0;
  int device = -1;
  double* cs = nullptr;
  float* phi = nullptr;
  float2* tw2 = nullptr;
  float2* tw1 = nullptr;
};

extern "C" {

int htamd_tomo_plan_create(int64_t n, int64_t ntheta, const double* theta, const double* mask,
                           htamd_tomo_plan** plan) {
  if (!plan || !theta) return fail(HTAMD_ERR_ARG, "tomo_plan_create: NULL argument");
  *plan = nullptr;
  if (n < 8 || n > 1024 || (n & (n - 1)))
    return fail(HTAMD_ERR_UNSUPPORTED, "tomo_plan_create: n must be a power of two in [8, 1024]");
  if (ntheta < 1) return fail(HTAMD_ERR_ARG, "tomo_plan_create: need ntheta >= 1");
  auto* p = new htamd_tomo_plan();
  p->n = n;
  p->ntheta = ntheta;
  // eps, mu, m (tomo.py:14-16)
  const double eps = 1e-3;
  p->mu = -std::log(eps) / (2.0 * (double)n * (double)n);
  p->m = (int)std::ceil(2.0 * n / M_PI * std::sqrt(-p->mu * std::log(eps) + (p->mu * n) * (p->mu * n) / 4.0));
  if (p->m > 7) {
    delete p;
    return fail(HTAMD_ERR_UNSUPPORTED, "tomo_plan_create: gridding half-width above 7");
  }
  hipError_t e = hipGetDevice(&p->device);
  // cos / sin of the angles as the reference rounds them (float32 theta)
  std::vector<double> cs((size_t)(2 * ntheta));
  for (int64_t k = 0; k < ntheta; ++k) {
    const double th = (double)(float)theta[k];
    cs[(size_t)k] = std::cos(th);
    cs[(size_t)(ntheta + k)] = std::sin(th);
  }
  // deapodiser phi (tomo.py:19-21, 46-47) with the mask folded in
  std::vector<float> phi((size_t)(n * n));
  for (int64_t i = 0; i < n; ++i) {
    const double dy = -0.5 + (double)i / (double)n;
    for (int64_t j = 0; j < n; ++j) {
      const double dx = -0.5 + (double)j / (double)n;
      double v = std::exp(p->mu * (double)(n * n) * (dx * dx + dy * dy)) * (double)(1 - n % 4);
      v *= (mask ? mask[i * n + j] : 1.0) / ((double)n * std::sqrt((double)n * (double)ntheta));
      phi[(size_t)(i * n + j)] = (float)v;
    }
  }
  auto tw = [](int64_t N) {
    std::vector<float2> t((size_t)N);
    for (int64_t k = 0; k < N; ++k) {
      const double a = -2.0 * M_PI * (double)k / (double)N;
      t[(size_t)k] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    return t;
  };
  auto up = [&](auto** dst, const auto& src) {
    hipError_t r = hipMalloc((void**)dst, src.size() * sizeof(src[0]));
    if (r == hipSuccess) r = hipMemcpy(*dst, src.data(), src.size() * sizeof(src[0]), hipMemcpyHostToDevice);
    return r;
  };
  if (e != hipSuccess || (e = up(&p->cs, cs)) != hipSuccess || (e = up(&p->phi, phi)) != hipSuccess ||
      (e = up(&p->tw2, tw(2 * n))) != hipSuccess || (e = up(&p->tw1, tw(n))) != hipSuccess) {
    htamd_tomo_plan_destroy(p);
    return hip_fail(e, "tomo_plan_create");
  }
  *plan = p;
  return HTAMD_OK;
}

int htamd_tomo_plan_destroy(htamd_tomo_plan* p) {
  if (!p) return HTAMD_OK;
  (void)hipFree(p->cs);
  (void)hipFree(p->phi);
  (void)hipFree(p->tw2);
  (void)hipFree(p->tw1);
  delete p;
  return HTAMD_OK;
}

int htamd_tomo_workspace_size(const htamd_tomo_plan* p, int64_t nz, int64_t* bytes) {
  if (!p || !bytes || nz < 0) return fail(HTAMD_ERR_ARG, "tomo_workspace_size: bad arguments");
  *bytes = nz * (p->n * 2 * p->n + 4 * p->n * p->n) * 8;
  return HTAMD_OK;
}

static int tomo_run(const htamd_tomo_plan* p, bool adjoint, const void* in, int is_complex, void* out,
                    int64_t nz, void* ws, int64_t ws_bytes, void* stream, const char* what) {
  if (!p) return fail(HTAMD_ERR_ARG, std::string(what) + ": plan is NULL");
  if (nz < 0) return fail(HTAMD_ERR_ARG, std::string(what) + ": negative slice count");
  if (nz == 0) return HTAMD_OK;
  if (!in || !out) return fail(HTAMD_ERR_ARG, std::string(what) + ": NULL buffer");
  int64_t need = 0;
  htamd_tomo_workspace_size(p, nz, &need);
  if (!ws || ws_bytes < need)
    return fail(HTAMD_ERR_ARG, std::string(what) + ": workspace of " + std::to_string(need) + " bytes required");
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != p->device) return fail(HTAMD_ERR_ARG, std::string(what) + ": plan belongs to another device");
  TomoArgs a;
  a.n = (int)p->n;
  a.nz = (int)nz;
  a.ntheta = (int)p->ntheta;
  a.m = p->m;
  a.mu = p->mu;
  a.cs = p->cs;
  a.phi = p->phi;
  a.tw2 = p->tw2;
  a.tw1 = p->tw1;
  a.W = (float2*)ws;
  a.G = (float2*)ws + nz * p->n * 2 * p->n;
  hipError_t e = is_complex ? tomo_dispatch<float2>(adjoint, a, in, out, (hipStream_t)stream)
                            : tomo_dispatch<float>(adjoint, a, in, out, (hipStream_t)stream);
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, what);
}

int htamd_tomo_R(const htamd_tomo_plan* p, const void* obj, int is_complex, void* out, int64_t nz,
                 void* workspace, int64_t workspace_bytes, void* stream) {
  return tomo_run(p, false, obj, is_complex, out, nz, workspace, workspace_bytes, stream, "tomo_R");
}

int htamd_tomo_RT(const htamd_tomo_plan* p, const void* data, int is_complex, void* out, int64_t nz,
                  void* workspace, int64_t workspace_bytes, void* stream) {
  return tomo_run(p, true, data, is_complex, out, nz, workspace, workspace_bytes, stream, "tomo_RT");
}

}  // extern "C"
