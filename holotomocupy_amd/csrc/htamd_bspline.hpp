// Shared device helpers and launch interface of the cubic B-spline kernels.
#pragma once
#include "htamd_common.hpp"

namespace htamd {

// ---- value helpers for real (float) and complex (float2) data --------------
template <typename T> __device__ __forceinline__ T vzero();
template <> __device__ __forceinline__ float vzero<float>() { return 0.0f; }
template <> __device__ __forceinline__ float2 vzero<float2>() { return make_float2(0.0f, 0.0f); }

__device__ __forceinline__ float vfma(float w, float v, float acc) { return fmaf(w, v, acc); }
__device__ __forceinline__ float2 vfma(float w, float2 v, float2 acc) {
  return make_float2(fmaf(w, v.x, acc.x), fmaf(w, v.y, acc.y));
}
// Re(a conj(b)) = Re a Re b + Im a Im b
__device__ __forceinline__ float vredot(float a, float b) { return a * b; }
__device__ __forceinline__ float vredot(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }

// ---- cubic B-spline tap weights ---------------------------------------------
// phi = 6 * beta3 (phi(0) = 4, phi(+-1) = 1).  For a fractional offset d in
// [0, 1) the four taps j = -1, 0, 1, 2 sit at t = d - j, each on a fixed
// polynomial piece of phi; w[j + 1] = phi(d - j), dw = phi', d2w = phi''.
__device__ __forceinline__ void bs_w(float d, float* w) {
  const float u = 1.0f - d;
  w[0] = u * u * u;
  w[1] = 4.0f - 6.0f * d * d + 3.0f * d * d * d;
  w[2] = 4.0f - 6.0f * u * u + 3.0f * u * u * u;
  w[3] = d * d * d;
}
__device__ __forceinline__ void bs_dw(float d, float* w) {
  const float u = 1.0f - d;
  w[0] = -3.0f * u * u;
  w[1] = -12.0f * d + 9.0f * d * d;
  w[2] = 12.0f * u - 9.0f * u * u;
  w[3] = 3.0f * d * d;
}
__device__ __forceinline__ void bs_d2w(float d, float* w) {
  const float u = 1.0f - d;
  w[0] = 6.0f * u;
  w[1] = -12.0f + 18.0f * d;
  w[2] = -12.0f + 18.0f * u;
  w[3] = 6.0f * d;
}

// Sampling coordinate of detector index t (cuda_kernels.py:174-177), in FP32
// with the reference's operation order: used only to bracket candidate taps
// (every caller keeps a one-pixel margin).
__device__ __forceinline__ float bs_coord(int t, int npix, int npsi, float m, float r) {
  const float half = (m - 1.0f) / 2.0f;
  return (m * (float)(t - npix / 2) - r + half) + (float)(npsi / 2);
}

// The same coordinate split into its integer part and fractional offset, which
// is what the taps use: x = m (t - npix/2) - r + (m - 1)/2 + npsi/2 is formed in
// FP64 (exact for these operands), so the offset d in [0, 1) carries FP32
// precision.  An FP32 coordinate near 1000 is only good to 6e-5 of a pixel,
// the dominant error of S against FP64 at 2048^2 detectors.  Every tap weight
// (S, its derivatives and the gather adjoints) comes from this one function, so
// the adjoint pairs use identical weights.
struct BsPos {
  int i;    // floor(x)
  float d;  // x - floor(x)
};
__device__ __forceinline__ BsPos bs_pos(int t, int npix, int npsi, float m, float r) {
  const float half = (m - 1.0f) / 2.0f;
  const double x = (double)m * (double)(t - npix / 2) - (double)r + (double)half + (double)(npsi / 2);
  const double fx = floor(x);
  return BsPos{(int)fx, (float)(x - fx)};
}

// Sadj as a gather: candidate detector pixels [lo, hi] whose taps can reach
// coefficient index i, bracketed analytically with a one-pixel margin (each is
// then re-evaluated with the forward's exact float expression).
__device__ __forceinline__ void bs_candidates(int i, int npix, int npsi, float mb, float rr, int& lo,
                                              int& hi) {
  const float off = bs_coord(0, npix, npsi, mb, rr) + mb * (float)(npix / 2);
  const float a = ((float)i - 2.0f - off) / mb + (float)(npix / 2);
  const float bnd = ((float)i + 2.0f - off) / mb + (float)(npix / 2);
  lo = (int)floorf(fminf(a, bnd)) - 1;
  hi = (int)ceilf(fmaxf(a, bnd)) + 1;
  lo = lo < 0 ? 0 : lo;
  hi = hi > npix - 1 ? npix - 1 : hi;
}

// weight with which sampling position p reaches coefficient index i (0 if none)
__device__ __forceinline__ float bs_reach(BsPos p, int i) {
  const int j = i - p.i;  // tap offset, valid in [-1, 2]
  if (j < -1 || j > 2) return 0.0f;
  float w[4];
  bs_w(p.d, w);
  return w[j + 1];
}

// Block-wide sum of a float2 (fixed order: wave shuffles, then LDS).
template <int NT>
__device__ __forceinline__ float2 block_sum2(float2 v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  constexpr int NW = NT / 64;
  __shared__ float2 red[NW];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  float2 s = make_float2(0.0f, 0.0f);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      s.x += red[i].x;
      s.y += red[i].y;
    }
  }
  return s;
}

// ---- launch interface ------------------------------------------------------
template <typename T>
struct BsArgs {
  const T* c = nullptr;   // coefficients [B, nzpsi, npsi] (adjoint: input g [B, nz, n])
  const T* c1 = nullptr;
  const T* c2 = nullptr;
  const float* r = nullptr;    // shifts, (y, x) at r[b*rs + 0/1]
  const float* m = nullptr;    // magnification at m[b*ms]
  const float* dr1 = nullptr;  // differential shifts
  const float* dr2 = nullptr;
  T* out = nullptr;
  int64_t rs = 2, ms = 1, d1s = 2, d2s = 2;
  int64_t batch = 0;
  int n = 0, nz = 0, npsi = 0, nzpsi = 0;
  int tiles_x = 0, tiles_y = 0;
};

constexpr int kPfH = 12;  // prefilter taps |d| <= kPfH
struct PrefilterTaps {
  float wx[2 * kPfH + 1];
  float wy[2 * kPfH + 1];
};

template <typename T> hipError_t bs_launch_forward(BsArgs<T> a, int mode, hipStream_t st);
template <typename T> hipError_t bs_launch_adjoint(BsArgs<T> a, hipStream_t st);
template <typename T> int64_t bs_dadj_partials(const BsArgs<T>& a);
template <typename T>
hipError_t bs_launch_dadj_reduce(BsArgs<T> a, const T* g, float2* partial, float* out2, hipStream_t st);
// Optional per-plane addresses of the prefilter's input and output (device
// arrays of batch pointers); a NULL table means plane b at in / out + b * plane.
template <typename T>
struct PlaneTab {
  const T* const* in = nullptr;
  T* const* out = nullptr;
};
template <typename T>
hipError_t bs_launch_prefilter(const T* in, T* out, int64_t batch, int nzpsi, int npsi,
                               const PrefilterTaps& taps, hipStream_t st, PlaneTab<T> tab = PlaneTab<T>());

}  // namespace htamd
