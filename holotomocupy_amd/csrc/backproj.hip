// Magnified back-shift for the Paganin initial guess: Shift.Sback,
// Shift.coeffback and Shift.curlySback (src/holotomocupy/shift.py:84-153, the
// sback kernel src/holotomocupy/cuda_kernels.py:294-368).
//
// Sback spreads every detector pixel over the object grid with the stretched
// spline phi(t / m) (support |t| < 2m).  Here it runs as a gather: each object
// pixel collects the detector pixels whose window reaches it, re-evaluating the
// reference's float expressions (sample point, floor, window bounds) so the
// same terms are included; the sum order is fixed (the reference's atomics
// have none).  coeffback divides the 2-D spectrum by the separable
// D(f) = phi(0) + 2 sum_{k=1..4} phi(k/m) cos(2 pi k f) per axis; it runs as a
// row pass and a column pass of the LDS FFT with the filter applied between
// the forward and inverse transforms.
#include <cmath>
#include <string>

#include "../../include/htamd.h"
#include "fft_lds.hpp"
#include "htamd_bspline.hpp"
#include "htamd_internal.hpp"

namespace htamd {

// phi(t / m) with phi = 6 beta3, as the sback kernel evaluates it
__device__ __forceinline__ float phi_back(float t, float m) {
  t /= m;
  if (-2.0f < t && t <= -1.0f) return (t + 2.0f) * (t + 2.0f) * (t + 2.0f);
  if (-1.0f < t && t <= 1.0f) return 4.0f - 6.0f * t * t + 3.0f * fabsf(t) * t * t;
  if (1.0f < t && t <= 2.0f) return (2.0f - t) * (2.0f - t) * (2.0f - t);
  return 0.0f;
}

// sample point of detector pixel t (cuda_kernels.py:320-321)
__device__ __forceinline__ float sback_coord(int t, int npix, int npsi, float mag, float r) {
  const float half = (mag - 1.0f) / 2.0f;
  return (mag * (float)(t - npix / 2) - r + half) + (float)(npsi / 2);
}

// weight with which detector pixel t reaches object index X (0 if it does not):
// the reference loops j in [ceil(d - span), (int)(d + span)) around floor(x)
__device__ __forceinline__ float sback_weight(float x, int X, float mag) {
  const int ix = (int)floorf(x);
  const float d = x - (float)ix;
  const float span = 2.0f * mag;
  const int j = X - ix;
  if (j < (int)ceilf(d - span) || j >= (int)(d + span)) return 0.0f;
  return phi_back(d - (float)j, mag);
}


// One thread per object pixel (b, Y, X).
__global__ __launch_bounds__(256) void sback_k(const float2* __restrict__ g, const float* __restrict__ r,
                                               int64_t rs, const float* __restrict__ m, int64_t ms,
                                               float2* __restrict__ out, int n, int nz, int npsi, int nzpsi,
                                               int64_t batch) {
  const int64_t plane = (int64_t)npsi * nzpsi;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= plane * batch) return;
  const int64_t b = idx / plane;
  const int Y = (int)((idx % plane) / npsi), X = (int)(idx % npsi);
  const float mag = m[b * ms], ry = r[b * rs + 0], rx = r[b * rs + 1];
  // candidates: |x(t) - X| < 2 mag + 1 with x(t) ~ mag (t - n/2) + off
  const float offx = sback_coord(0, n, npsi, mag, rx) + mag * (float)(n / 2);
  const float offy = sback_coord(0, nz, nzpsi, mag, ry) + mag * (float)(nz / 2);
  const float reach = 2.0f * mag + 2.0f;
  int xlo = (int)floorf(((float)X - reach - offx) / mag) + n / 2 - 1;
  int xhi = (int)ceilf(((float)X + reach - offx) / mag) + n / 2 + 1;
  int ylo = (int)floorf(((float)Y - reach - offy) / mag) + nz / 2 - 1;
  int yhi = (int)ceilf(((float)Y + reach - offy) / mag) + nz / 2 + 1;
  xlo = max(xlo, 0);
  ylo = max(ylo, 0);
  xhi = min(xhi, n - 1);
  yhi = min(yhi, nz - 1);
  const float2* gb = g + b * (int64_t)n * nz;
  float2 acc = make_float2(0.f, 0.f);
  for (int ty = ylo; ty <= yhi; ++ty) {
    const float wy = sback_weight(sback_coord(ty, nz, nzpsi, mag, ry), Y, mag);
    if (wy == 0.0f) continue;
    float2 racc = make_float2(0.f, 0.f);
    for (int tx = xlo; tx <= xhi; ++tx) {
      const float wx = sback_weight(sback_coord(tx, n, npsi, mag, rx), X, mag);
      if (wx != 0.0f) {
        const float2 v = gb[(int64_t)ty * n + tx];
        racc = make_float2(fmaf(wx, v.x, racc.x), fmaf(wx, v.y, racc.y));
      }
    }
    acc = make_float2(fmaf(wy, racc.x, acc.x), fmaf(wy, racc.y, acc.y));
  }
  out[idx] = acc;
}

// 1 / D(f_k) of the coeffback filter at FFT bin k of length N; the reference
// builds D on x = linspace(-1/2, 1/2 - 1/N, N) and fftshifts it, so bin k
// sits at x = (k + N/2) mod N / N - 1/2.
__device__ __forceinline__ float coeffback_gain(int k, int N, const float* phik) {
  const double x = (double)((k + N / 2) % N) / (double)N - 0.5;
  double d = (double)phik[0];
  for (int q = 1; q <= 4; ++q) d += 2.0 * (double)phik[q] * cos(2.0 * M_PI * q * x);
  return (float)(1.0 / d);
}

struct CbTables {
  float phik[5];  // phi(k / m), k = 0..4
};

// rows of length N: x <- IFFT(FFT(x) / D) (normalised)
template <int N, int E>
__global__ __launch_bounds__(256) void coeffback_rows_k(const float2* __restrict__ in, float2* __restrict__ out,
                                                        int64_t nrows, const float2* __restrict__ tw, CbTables tb) {
  constexpr int T = N / E, ROWS = 256 / T;
  __shared__ float2 lds[ROWS * fft_buf<N>()];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int64_t row = (int64_t)blockIdx.x * ROWS + r;
  const bool valid = row < nrows;
  float2* s = lds + r * fft_buf<N>();
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, tw);
  const float2* src = in + (valid ? row : 0) * N;
#pragma unroll
  for (int e = 0; e < E; ++e) s[fpad(t + e * T)] = valid ? src[t + e * T] : make_float2(0.f, 0.f);
  __syncthreads();
  fft_lds<N, E, false>(s, t, twr);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + e * T;
    s[fpad(k)] = cscale(s[fpad(k)], coeffback_gain(k, N, tb.phik) / (float)N);
  }
  __syncthreads();
  fft_lds<N, E, true>(s, t, twr);
  if (valid) {
    float2* dst = out + row * N;
#pragma unroll
    for (int e = 0; e < E; ++e) dst[t + e * T] = s[fpad(t + e * T)];
  }
}

// columns of length N of [batch, N, w] planes, C adjacent columns per
// workgroup, in place
template <int N, int E, int C>
__global__ __launch_bounds__(C*(N / E)) void coeffback_cols_k(float2* data, int w, int tiles_per_plane,
                                                              const float2* __restrict__ tw, CbTables tb) {
  constexpr int T = N / E, NT = C * T, LS = col_stride<N, C>();
  __shared__ float2 lds[C * LS];
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t b = tile / tiles_per_plane;
  const int x0 = (int)(tile % tiles_per_plane) * C;
  float2* base = data + b * (int64_t)N * w + x0;
  const int cc = threadIdx.x / T, t = threadIdx.x % T;
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, tw);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int idx = threadIdx.x + e * NT;
    const int c = idx % C, q = idx / C;
    lds[c * LS + fpad(q)] = base[(int64_t)q * w + c];
  }
  __syncthreads();
  float2* s = lds + cc * LS;
  fft_lds<N, E, false>(s, t, twr);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + e * T;
    s[fpad(k)] = cscale(s[fpad(k)], coeffback_gain(k, N, tb.phik) / (float)N);
  }
  __syncthreads();
  fft_lds<N, E, true>(s, t, twr);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int idx = threadIdx.x + e * NT;
    const int c = idx % C, q = idx / C;
    base[(int64_t)q * w + c] = lds[c * LS + fpad(q)];
  }
}

template <int N>
static hipError_t cb_rows(const float2* in, float2* out, int64_t nrows, const float2* tw, const CbTables& tb,
                          hipStream_t st) {
  constexpr int E = 8, ROWS = 256 / (N / E);
  hipLaunchKernelGGL((coeffback_rows_k<N, E>), dim3((unsigned)((nrows + ROWS - 1) / ROWS)), dim3(256), 0, st, in,
                     out, nrows, tw, tb);
  return hipGetLastError();
}

template <int N>
static hipError_t cb_cols(float2* data, int64_t batch, int w, const float2* tw, const CbTables& tb, hipStream_t st) {
  constexpr int E = 8, T = N / E;
  constexpr int CT = 256 / T > 0 ? 256 / T : 1;
  constexpr int CL = 8192 / N > 0 ? 8192 / N : 1;
  constexpr int C0 = CT < CL ? CT : CL;
  constexpr int C = C0 > 16 ? 16 : C0;
  if (w % C != 0) {
    hipLaunchKernelGGL((coeffback_cols_k<N, E, 1>), dim3((unsigned)(batch * w)), dim3(T), 0, st, data, w, w, tw, tb);
  } else {
    hipLaunchKernelGGL((coeffback_cols_k<N, E, C>), dim3((unsigned)(batch * (w / C))), dim3(C * T), 0, st, data, w,
                       w / C, tw, tb);
  }
  return hipGetLastError();
}

#define HTAMD_POW2_SWITCH(NN, CALL)                                                                         \
  switch (NN) {                                                                                            \
    case 8: return CALL(8);                                                                                \
    case 16: return CALL(16);                                                                              \
    case 32: return CALL(32);                                                                              \
    case 64: return CALL(64);                                                                              \
    case 128: return CALL(128);                                                                            \
    case 256: return CALL(256);                                                                            \
    case 512: return CALL(512);                                                                            \
    case 1024: return CALL(1024);                                                                          \
    case 2048: return CALL(2048);                                                                          \
    default: return hipErrorInvalidValue;                                                                  \
  }

static hipError_t coeffback_rows(int64_t N, const float2* in, float2* out, int64_t nrows, const float2* tw,
                                 const CbTables& tb, hipStream_t st) {
#define HTAMD_CBR(K) cb_rows<K>(in, out, nrows, tw, tb, st)
  HTAMD_POW2_SWITCH(N, HTAMD_CBR)
#undef HTAMD_CBR
}

static hipError_t coeffback_cols(int64_t N, float2* data, int64_t batch, int w, const float2* tw, const CbTables& tb,
                                 hipStream_t st) {
#define HTAMD_CBC(K) cb_cols<K>(data, batch, w, tw, tb, st)
  HTAMD_POW2_SWITCH(N, HTAMD_CBC)
#undef HTAMD_CBC
}

// phi = 6 beta3 (shift.py:54-59) in FP64
static double phi_host(double t) {
  const double a = std::fabs(t);
  if (a <= 1.0) return 4.0 - 6.0 * t * t + 3.0 * a * t * t;
  if (a < 2.0) return (2.0 - a) * (2.0 - a) * (2.0 - a);
  return 0.0;
}

}  // namespace htamd

using namespace htamd;

extern "C" {

int htamd_bspline_sback(const htamd_shift_plan* p, const void* g, const float* r, int64_t r_stride, const float* m,
                        int64_t m_stride, void* c_out, int64_t batch, void* stream) {
  if (!p) return fail(HTAMD_ERR_ARG, "bspline_sback: plan is NULL");
  if (batch < 0) return fail(HTAMD_ERR_ARG, "bspline_sback: negative batch");
  if (batch == 0) return HTAMD_OK;
  if (!g || !r || !m || !c_out) return fail(HTAMD_ERR_ARG, "bspline_sback: NULL buffer");
  const int64_t total = batch * p->npsi * p->nzpsi;
  hipLaunchKernelGGL(sback_k, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)g, r, r_stride, m, m_stride, (float2*)c_out, (int)p->n, (int)p->nz, (int)p->npsi,
                     (int)p->nzpsi, batch);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "bspline_sback");
}

int htamd_bspline_coeffback(const htamd_shift_plan* p, const void* in, void* out, int64_t batch, double m,
                            void* stream) {
  if (!p) return fail(HTAMD_ERR_ARG, "bspline_coeffback: plan is NULL");
  if (batch < 0 || !(m > 0.0)) return fail(HTAMD_ERR_ARG, "bspline_coeffback: bad batch or magnification");
  if (batch == 0) return HTAMD_OK;
  if (!in || !out) return fail(HTAMD_ERR_ARG, "bspline_coeffback: NULL buffer");
  if (!p->tw_x || !p->tw_y)
    return fail(HTAMD_ERR_UNSUPPORTED, "bspline_coeffback: object grid sizes must be powers of two in [8, 2048]");
  const float2* tw_x = p->tw_x;
  const float2* tw_y = p->tw_y;
  CbTables tb;
  for (int k = 0; k <= 4; ++k) tb.phik[k] = (float)phi_host((double)k / m);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = coeffback_rows(p->npsi, (const float2*)in, (float2*)out, batch * p->nzpsi, tw_x, tb, st);
  if (e == hipSuccess) e = coeffback_cols(p->nzpsi, (float2*)out, batch, (int)p->npsi, tw_y, tb, st);
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "bspline_coeffback");
}

}  // extern "C"
