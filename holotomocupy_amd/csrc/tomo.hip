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
// loads and stores.  The 9x9 Gaussian gridding taps of every polar sample are
// built once per plan in FP64 (stored FP32): the forward gathers them per
// (angle, slice) line with the line's 1-D FFT fused; the adjoint uses the same
// taps transposed (one CSR row per grid point) and gathers onto the grid from
// a slice-contiguous copy of the transformed sinogram -- the reference's atomic
// scatter in a fixed summation order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/htamd.h"
#include "fft_any.hpp"
#include "fft_lds.hpp"
#include "htamd_bspline.hpp"
#include "htamd_internal.hpp"

namespace htamd {

struct TomoArgs {
  int n = 0, nz = 0, ntheta = 0, m = 0;
  const float* phi = nullptr;      // [n, n] deapodisation (mask and scale included)
  const float2* tw2 = nullptr;     // twiddles of length 2n
  const float2* tw1 = nullptr;     // twiddles of length n
  // gridding taps of polar sample s = theta n + t (built in FP64 on the host):
  const int2* tb = nullptr;        // [ntheta n] first tap (bx, by) on the 2n grid, before wrapping
  const float* tw = nullptr;       // [2L][ntheta n]: wx[i] at tw[i S + s], wy[i] at tw[(L + i) S + s]
  // the same taps transposed, one row per grid point g = row 2n + col (CSR):
  const int* rp = nullptr;         // [4 n^2 + 1]
  const int2* csw = nullptr;       // [nnz] (sample index s, weight wy wx as float bits)
  // the samples binned by grid tile for the forward gather: work item i covers
  // samples [it.z, it.z + it.w) of the sorted list, all in tile (it.x, it.y);
  // per sorted sample: (local tap origin oy << 16 | ox, sample index) and the
  // fractional parts (frac(2n x0), frac(2n y0)) that give its weights
  const int4* items = nullptr;
  const int2* so = nullptr;
  const float2* sf = nullptr;
  int ts = 0;                      // tile edge on the 2n grid
  int nitems = 0;
  float cxs = 0.f, cys = 0.f;      // weights exp(cxs (i - m - f)^2), cys exp(...)
  float2* W = nullptr;             // [nz, n, 2n]
  float2* G = nullptr;             // [nz, 2n, 2n]
  float2* Sh = nullptr;            // [ntheta n, nz] transformed sinogram (RT)
  // grid points whose CSR rows are longer than kHeavyRow, longest first; the
  // thread-per-point gather skips them and a wavefront per point sums them
  const int* heavy = nullptr;
  int nheavy = 0;
  // per-bin factor of the adjoint's line transform (NULL = none): the
  // detector filter of filtered back-projection, shifted by n/2 to the bins
  // of the c1d-modulated transform and scaled by sqrt(n / ntheta)
  const float* filt = nullptr;
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
  const float2* g = a.G + (int64_t)z * n2 * n2;
  const int L = 2 * a.m + 1;
  const int64_t S = (int64_t)a.ntheta * N;
  for (int e = 0; e < E; ++e) {
    const int tt = t + e * T;
    const int64_t si = (int64_t)th * N + tt;
    const int2 b = a.tb[si];
    const int bx = b.x, by = b.y;
    float wx[16];
    for (int i0 = 0; i0 < L; ++i0) wx[i0] = a.tw[i0 * S + si];
    float2 acc = make_float2(0.f, 0.f);
    for (int i1 = 0; i1 < L; ++i1) {
      const float wy = a.tw[(L + i1) * S + si];
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

// ---- forward pass C, tiled: gather through LDS tiles of the grid ------------
// The polar samples are binned by the grid tile holding the first tap of
// their 9x9 window (plan time).  A workgroup takes one work item (up to 256
// samples of one tile) for kTZ slices: it stages the tile plus the window
// margin, wrapped on the 2n grid, into LDS for each slice, and each thread
// gathers its sample's taps from LDS.  The weights come from the sample's
// fractional grid offset, exp(-pi^2/mu ((i - m - f)/(2n))^2) evaluated in
// FP32 (the FP64-built offset keeps the argument accurate).  The sums run in
// the order of the untiled gather (rows outer, columns inner).  Output: the
// line spectra Sh[(theta nz + z) n + t], scaled by c1d / n.
constexpr int kTS = 32, kTL = 9, kTP = kTS + kTL - 1, kTZ = 4;
__global__ __launch_bounds__(256) void tomo_tile_gather_k(TomoArgs a) {
  static_assert(kTZ == 4, "a staged point's slices are two float4");
  // slice-minor: point (r, c) holds its kTZ slices at tile[(r kTP + c) kTZ + q],
  // so a tap's four slices are two 16-byte LDS reads
  __shared__ float4 tile4[kTP * kTP * kTZ / 2];
  float2* tile = reinterpret_cast<float2*>(tile4);
  const int4 it = a.items[blockIdx.x];
  const int z0 = blockIdx.y * kTZ;
  const int n2 = 2 * a.n, P = a.ts + kTL - 1;
  const int64_t plane = (int64_t)n2 * n2;
  const int gy0 = it.y * a.ts, gx0 = it.x * a.ts;
  // staged element (q, r, c) by the compile-time pitch kTP (the divisions are
  // multiply-shifts), slice fastest so consecutive threads fill consecutive
  // LDS slots; the wrap onto the 2n grid is one conditional subtraction: the
  // tile origin is a multiple of ts below 2n and r, c < ts + 8, which keeps
  // origin + r below 4n for every even n >= 8
  for (int idx = threadIdx.x; idx < kTZ * kTP * kTP; idx += 256) {
    const int pt = idx / kTZ, q = idx - pt * kTZ;
    const int r = pt / kTP, c = pt - r * kTP;
    if (r >= P || c >= P) continue;  // tiles narrower than kTS (2n < kTS)
    const int z = z0 + q;
    int gr = gy0 + r, gc = gx0 + c;
    gr -= gr >= n2 ? n2 : 0;
    gc -= gc >= n2 ? n2 : 0;
    tile[idx] = z < a.nz ? a.G[z * plane + (int64_t)gr * n2 + gc] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  if ((int)threadIdx.x >= it.w) return;
  const int k = it.z + threadIdx.x;
  const int2 so = a.so[k];
  const float2 f = a.sf[k];
  const int ox = so.x & 0xffff, oy = so.x >> 16, si = so.y;
  float wx[kTL];
#pragma unroll
  for (int i = 0; i < kTL; ++i) {
    const float dx = (float)(i - a.m) - f.x;
    wx[i] = expf(a.cxs * dx * dx);
  }
  const int th = si / a.n, t = si % a.n;
  const float sc = -sgnf(t) / (float)a.n;
  const float4* tq = tile4 + (oy * kTP + ox) * 2;
  float2 acc[kTZ];
#pragma unroll
  for (int q = 0; q < kTZ; ++q) acc[q] = make_float2(0.f, 0.f);
#pragma unroll 1
  for (int i1 = 0; i1 < kTL; ++i1) {  // one tap row at a time (registers)
    const float dy = (float)(i1 - a.m) - f.y;
    const float wy = a.cys * expf(a.cxs * dy * dy);
    float2 racc[kTZ];
#pragma unroll
    for (int q = 0; q < kTZ; ++q) racc[q] = make_float2(0.f, 0.f);
#pragma unroll
    for (int i0 = 0; i0 < kTL; ++i0) {
      const float4 v01 = tq[(i1 * kTP + i0) * 2], v23 = tq[(i1 * kTP + i0) * 2 + 1];
      racc[0] = vfma(wx[i0], make_float2(v01.x, v01.y), racc[0]);
      racc[1] = vfma(wx[i0], make_float2(v01.z, v01.w), racc[1]);
      racc[2] = vfma(wx[i0], make_float2(v23.x, v23.y), racc[2]);
      racc[3] = vfma(wx[i0], make_float2(v23.z, v23.w), racc[3]);
    }
#pragma unroll
    for (int q = 0; q < kTZ; ++q) acc[q] = vfma(wy, racc[q], acc[q]);
  }
#pragma unroll
  for (int q = 0; q < kTZ; ++q)
    if (z0 + q < a.nz) a.Sh[((int64_t)th * a.nz + z0 + q) * a.n + t] = cscale(acc[q], sc);
}

// ---- forward pass D: line IFFT of the gathered spectra ----------------------
template <int N, int E, typename TO>
__global__ __launch_bounds__(256) void tomo_line_ifft_k(TomoArgs a, TO* __restrict__ out) {
  constexpr int T = N / E, ROWS = 256 / T;
  __shared__ float2 lds[ROWS * fft_buf<N>()];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int64_t line = (int64_t)blockIdx.x * ROWS + r;  // theta nz + z
  const bool valid = line < (int64_t)a.ntheta * a.nz;
  float2* s = lds + r * fft_buf<N>();
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, a.tw1);
  const float2* src = a.Sh + (valid ? line : 0) * N;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int tt = t + e * T;
    s[fpad(tt)] = src[tt];
  }
  __syncthreads();
  fft_lds<N, E, true>(s, t, twr);
  if (valid) {
    TO* dst = out + line * N;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int tt = t + e * T;
      st_c<TO>(dst, tt, cscale(s[fpad(tt)], -sgnf(tt) / (4.0f * N)));
    }
  }
}

// ---- adjoint pass C': line FFT into the z-contiguous sinogram ---------------
template <int N, int E, typename TI>
__global__ __launch_bounds__(256) void tomo_linefft_k(TomoArgs a, const TI* __restrict__ data) {
  constexpr int T = N / E, ROWS = 256 / T;
  __shared__ float2 lds[ROWS * fft_buf<N>()];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int64_t line = (int64_t)blockIdx.x * ROWS + r;  // (theta, z), z fastest
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
  if (valid) {
    for (int e = 0; e < E; ++e) {
      const int tt = t + e * T;
      const float sc = a.filt ? -sgnf(tt) * a.filt[tt] : -sgnf(tt);
      a.Sh[((int64_t)th * N + tt) * a.nz + z] = cscale(s[fpad(tt)], sc);
    }
  }
}

// ---- detector filter alone (Tomo._filter_sino): IFFT(h FFT(line)) -----------
// h already holds the 1/N of the inverse transform.  in may equal out.
template <int N, int E, typename TIO>
__global__ __launch_bounds__(256) void tomo_filter_k(const TIO* in, TIO* out, int64_t lines,
                                                     const float2* __restrict__ tw, const float* __restrict__ h) {
  constexpr int T = N / E, ROWS = 256 / T;
  __shared__ float2 lds[ROWS * fft_buf<N>()];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int64_t line = (int64_t)blockIdx.x * ROWS + r;
  const bool valid = line < lines;
  float2* s = lds + r * fft_buf<N>();
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, tw);
  const TIO* src = in + (valid ? line : 0) * N;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + e * T;
    s[fpad(k)] = valid ? ld_c<TIO>(src, k) : make_float2(0.f, 0.f);
  }
  __syncthreads();
  fft_lds<N, E, false>(s, t, twr);
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int k = t + e * T;
    s[fpad(k)] = cscale(s[fpad(k)], h[k]);
  }
  __syncthreads();
  fft_lds<N, E, true>(s, t, twr);
  if (valid) {
    TIO* dst = out + line * N;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = t + e * T;
      st_c<TIO>(dst, k, s[fpad(k)]);
    }
  }
}

// ---- adjoint pass C'': gather onto the grid (transpose of the forward taps) ---
// Grid point g sums w_k Sh[s_k, z] over its CSR row (packed 8-byte (sample,
// weight) pairs) in entry order: deterministic, no atomics.  The lanes of a
// wavefront run across ZB consecutive slices of P = 64 / ZB points, so one
// entry of a point is one (wave-uniform for ZB = 64) index load and one
// contiguous ZB x 8-byte read of the sample's slice block -- every byte of a
// fetched line is used by the instruction that fetched it.  Entries are taken
// kGatherBatch at a time so that many independent reads are in flight.
//
// A workgroup (4 wavefronts) owns PT consecutive grid points and ZB slices;
// its results go through an LDS tile so the grid is written as rows of PT
// points per slice (PT = 32, more where a wave step takes 64 / ZB > 8 points;
// 64 points, LDS-bound at 4 workgroups per CU, measured 9% slower).  Rows through the centre of the grid collect a
// sample of every angle (at 256^2 with 384 angles the longest has 2005
// entries against a mean of 30): rows longer than kHeavyRow entries are left
// to tomo_grid_gather_heavy_k, launched after this kernel on the same stream,
// which overwrites those points.
constexpr int kHeavyRow = 64;
constexpr int kGatherBatch = 8;
constexpr int kGatherPoints = 32;  // grid points per workgroup of the light rows (at least 4 wave steps)
constexpr int kHeavyWaves = 4;     // wavefronts per heavy row

template <int ZB>
constexpr int gather_points() {
  return kGatherPoints > 4 * (64 / ZB) ? kGatherPoints : 4 * (64 / ZB);
}

// acc += sum over entries [k0, k1) of w Sh[s, z] in entry order, for the
// slice block starting at src = Sh + z (zin: this lane's slice exists)
__device__ __forceinline__ float2 gather_entries(const int2* __restrict__ csw, const float2* __restrict__ src,
                                                 int64_t nz, int k0, int k1, bool zin, float2 acc) {
  int k = k0;
  for (; k + kGatherBatch <= k1; k += kGatherBatch) {
    int2 cw[kGatherBatch];
    float2 v[kGatherBatch];
#pragma unroll
    for (int u = 0; u < kGatherBatch; ++u) cw[u] = csw[k + u];
#pragma unroll
    for (int u = 0; u < kGatherBatch; ++u) v[u] = zin ? src[(int64_t)cw[u].x * nz] : make_float2(0.f, 0.f);
#pragma unroll
    for (int u = 0; u < kGatherBatch; ++u) acc = vfma(__int_as_float(cw[u].y), v[u], acc);
  }
  for (; k < k1; ++k) {
    const int2 cw = csw[k];
    const float2 v = zin ? src[(int64_t)cw.x * nz] : make_float2(0.f, 0.f);
    acc = vfma(__int_as_float(cw.y), v, acc);
  }
  return acc;
}

template <int ZB, bool FULL>
__global__ __launch_bounds__(256) void tomo_grid_gather_k(TomoArgs a) {
  constexpr int P = 64 / ZB, PT = gather_points<ZB>(), TS = PT + 1;  // points per wave step, per workgroup; tile stride
  static_assert(PT % (4 * P) == 0, "whole wave steps per workgroup");
  __shared__ float2 tile[ZB * TS];
  const int64_t n4 = 4LL * a.n * a.n;
  const int64_t g0 = (int64_t)blockIdx.x * PT;
  const int z0 = blockIdx.y * ZB;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / 64), lane = threadIdx.x % 64;
  const int ps = lane / ZB, q = lane % ZB;
  const bool zin = FULL || z0 + q < a.nz;
  const float2* src = a.Sh + z0 + q;
#pragma unroll 1
  for (int it = 0; it < PT / 4 / P; ++it) {
    const int pl = wv * (PT / 4) + it * P + ps;  // point of the tile
    const int64_t g = g0 + pl;
    float2 acc = make_float2(0.f, 0.f);
    if (g < n4) {
      const int k0 = a.rp[g], k1 = a.rp[g + 1];
      if (k1 - k0 <= kHeavyRow) acc = gather_entries(a.csw, src, a.nz, k0, k1, zin, acc);
    }
    tile[q * TS + pl] = acc;
  }
  __syncthreads();
  // slice r of the tile is PT consecutive grid points, written contiguously
#pragma unroll 1
  for (int i = threadIdx.x; i < ZB * PT; i += 256) {
    const int r = i / PT, pl = i % PT;
    const int64_t g = g0 + pl;
    if (g < n4 && (FULL || z0 + r < a.nz)) a.G[(int64_t)(z0 + r) * n4 + g] = tile[r * TS + pl];
  }
}

// One workgroup per heavy grid point and ZB slices: its NC = kHeavyWaves P
// lane groups each sum one contiguous chunk of the point's entries (chunk
// c = wave P + group, entries [k0 + c L / NC, k0 + (c + 1) L / NC)), in entry order, and lane q
// then adds the chunks' sums for slice q in chunk order -- a fixed order, so
// the result does not depend on scheduling.
template <int ZB, bool FULL>
__global__ __launch_bounds__(64 * kHeavyWaves) void tomo_grid_gather_heavy_k(TomoArgs a) {
  constexpr int P = 64 / ZB, NC = kHeavyWaves * P;
  __shared__ float2 part[NC * ZB];
  const int64_t n4 = 4LL * a.n * a.n;
  const int g = a.heavy[blockIdx.x];
  const int z0 = blockIdx.y * ZB;
  const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int ps = lane / ZB, q = lane % ZB;
  const bool zin = FULL || z0 + q < a.nz;
  const int k0 = a.rp[g], L = a.rp[g + 1] - k0;
  const int c = wv * P + ps;
  const int kb = k0 + (int)((int64_t)L * c / NC), ke = k0 + (int)((int64_t)L * (c + 1) / NC);
  part[c * ZB + q] = gather_entries(a.csw, a.Sh + z0 + q, a.nz, kb, ke, zin, make_float2(0.f, 0.f));
  __syncthreads();
  if (threadIdx.x < ZB && (FULL || z0 + (int)threadIdx.x < a.nz)) {
    float2 t = make_float2(0.f, 0.f);
    for (int cc = 0; cc < NC; ++cc) t = cadd(t, part[cc * ZB + threadIdx.x]);
    a.G[(int64_t)(z0 + threadIdx.x) * n4 + g] = t;
  }
}

// The grid gather of the adjoint: the rows up to kHeavyRow entries, then the
// heavy rows.  ZB = 64 slices per workgroup where they divide the slab, else
// 32, 16, or 8 with the last block masked.
template <int ZB, bool FULL>
static void launch_grid_gather_zb(const TomoArgs& a, hipStream_t st) {
  const int64_t n4 = 4LL * a.n * a.n;
  const unsigned zb = (unsigned)((a.nz + ZB - 1) / ZB);
  constexpr int PT = gather_points<ZB>();
  hipLaunchKernelGGL((tomo_grid_gather_k<ZB, FULL>), dim3((unsigned)((n4 + PT - 1) / PT), zb), dim3(256), 0, st, a);
  if (a.nheavy)
    hipLaunchKernelGGL((tomo_grid_gather_heavy_k<ZB, FULL>), dim3((unsigned)a.nheavy, zb), dim3(64 * kHeavyWaves), 0,
                       st, a);
}

static hipError_t launch_grid_gather(const TomoArgs& a, hipStream_t st) {
  if (a.nz % 64 == 0)
    launch_grid_gather_zb<64, true>(a, st);
  else if (a.nz % 32 == 0)
    launch_grid_gather_zb<32, true>(a, st);
  else if (a.nz % 16 == 0)
    launch_grid_gather_zb<16, true>(a, st);
  else if (a.nz % 8 == 0)
    launch_grid_gather_zb<8, true>(a, st);
  else
    launch_grid_gather_zb<8, false>(a, st);
  return hipGetLastError();
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

// ---- any even n: the same passes on the run-time mixed-radix FFT --------------
// (fft_any.hpp; rows of the embedded slice and lines are one unit each,
// columns of the 2n grid C per unit).  Loads and stores apply the same sign
// modulations, embedding, cropping and deapodisation as the kernels above.
template <typename TI>
struct TomoRowsFwdAny {  // line = z n + i' of the n non-zero rows, length 2n
  static constexpr bool kInv = false, kTwo = false;
  TomoArgs a;
  const TI* obj;
  __device__ float2 load(int64_t line, int, int p) const {
    const int n = a.n, h = n / 2;
    if (p < h || p >= h + n) return make_float2(0.f, 0.f);
    const int ip = (int)(line % n), jp = p - h;
    return cscale(ld_c<TI>(obj + line * (int64_t)n, jp), a.phi[ip * n + jp] * sgnf(ip + h + p));
  }
  __device__ void mid(float2*, int64_t, int, int, int) const {}
  __device__ void store(int64_t line, int, int k, float2 v) const {
    a.W[line * (int64_t)(2 * a.n) + k] = cscale(v, sgnf(k));
  }
};

struct TomoColsFwdAny {  // unit = z groups + column group
  static constexpr bool kInv = false, kTwo = false;
  TomoArgs a;
  int C, groups;
  __device__ float2 load(int64_t unit, int c, int i) const {
    const int n = a.n, h = n / 2, N2 = 2 * n;
    const int col = (int)(unit % groups) * C + c;
    if (col >= N2 || i < h || i >= h + n) return make_float2(0.f, 0.f);
    return a.W[(unit / groups) * (int64_t)n * N2 + (int64_t)(i - h) * N2 + col];
  }
  __device__ void mid(float2*, int64_t, int, int, int) const {}
  __device__ void store(int64_t unit, int c, int k, float2 v) const {
    const int N2 = 2 * a.n;
    const int col = (int)(unit % groups) * C + c;
    if (col < N2) a.G[(unit / groups) * (int64_t)N2 * N2 + (int64_t)k * N2 + col] = cscale(v, sgnf(k));
  }
};

template <typename TO>
struct TomoLineIfftAny {  // line = theta nz + z, length n
  static constexpr bool kInv = true, kTwo = false;
  TomoArgs a;
  TO* out;
  __device__ float2 load(int64_t line, int, int t) const { return a.Sh[line * a.n + t]; }
  __device__ void mid(float2*, int64_t, int, int, int) const {}
  __device__ void store(int64_t line, int, int t, float2 v) const {
    st_c<TO>(out + line * a.n, t, cscale(v, -sgnf(t) / (4.0f * a.n)));
  }
};

template <typename TI>
struct TomoLineFftAny {  // line = theta nz + z, length n; z-contiguous output
  static constexpr bool kInv = false, kTwo = false;
  TomoArgs a;
  const TI* data;
  __device__ float2 load(int64_t line, int, int t) const {
    return cscale(ld_c<TI>(data + line * a.n, t), -sgnf(t));
  }
  __device__ void mid(float2*, int64_t, int, int, int) const {}
  __device__ void store(int64_t line, int, int t, float2 v) const {
    const int64_t th = line / a.nz, z = line % a.nz;
    const float sc = a.filt ? -sgnf(t) * a.filt[t] : -sgnf(t);
    a.Sh[(th * a.n + t) * a.nz + z] = cscale(v, sc);
  }
};

struct TomoColsAdjAny {  // unit = z groups + column group; keeps the n central rows
  static constexpr bool kInv = true, kTwo = false;
  TomoArgs a;
  int C, groups;
  __device__ float2 load(int64_t unit, int c, int k) const {
    const int N2 = 2 * a.n;
    const int col = (int)(unit % groups) * C + c;
    if (col >= N2) return make_float2(0.f, 0.f);
    return cscale(a.G[(unit / groups) * (int64_t)N2 * N2 + (int64_t)k * N2 + col], sgnf(k + col));
  }
  __device__ void mid(float2*, int64_t, int, int, int) const {}
  __device__ void store(int64_t unit, int c, int i, float2 v) const {
    const int n = a.n, h = n / 2, N2 = 2 * n;
    const int col = (int)(unit % groups) * C + c;
    if (col < N2 && i >= h && i < h + n)
      a.W[(unit / groups) * (int64_t)n * N2 + (int64_t)(i - h) * N2 + col] = cscale(v, sgnf(i) / (float)N2);
  }
};

template <typename TO>
struct TomoRowsAdjAny {  // line = z n + i', length 2n; crop and deapodise
  static constexpr bool kInv = true, kTwo = false;
  TomoArgs a;
  TO* out;
  __device__ float2 load(int64_t line, int, int k) const { return a.W[line * (int64_t)(2 * a.n) + k]; }
  __device__ void mid(float2*, int64_t, int, int, int) const {}
  __device__ void store(int64_t line, int, int p, float2 v) const {
    const int n = a.n, h = n / 2;
    if (p < h || p >= h + n) return;
    const int ip = (int)(line % n), jp = p - h;
    st_c<TO>(out + line * (int64_t)n, jp, cscale(v, sgnf(p) / (float)(2 * n) * a.phi[ip * n + jp]));
  }
};

template <typename TIO>
struct TomoFilterAny {  // IFFT(h FFT(line)), h with the 1/n of the inverse
  static constexpr bool kInv = false, kTwo = true;
  const TIO* in;
  TIO* out;
  int n;
  const float* h;
  __device__ float2 load(int64_t line, int, int k) const { return ld_c<TIO>(in + line * n, k); }
  __device__ void mid(float2* s, int64_t, int, int tid, int nthr) const {
    for (int k = tid; k < n; k += nthr) s[k] = cscale(s[k], h[k]);
  }
  __device__ void store(int64_t line, int, int k, float2 v) const { st_c<TIO>(out + line * n, k, v); }
};

static bool tomo_any_ffts(const TomoArgs& a, AnyFft* f2, AnyFft* f1) {
  if (!any_fft_factor(2 * (int64_t)a.n, f2) || !any_fft_factor(a.n, f1)) return false;
  f2->tw = a.tw2;
  f1->tw = a.tw1;
  return true;
}

template <typename TIO>
static hipError_t tomo_fwd_any(const TomoArgs& a, const TIO* obj, TIO* out, hipStream_t st) {
  AnyFft f2, f1;
  if (!tomo_any_ffts(a, &f2, &f1) || !a.items) return hipErrorInvalidValue;
  const int N2 = 2 * a.n;
  hipError_t e = any_launch_rows(TomoRowsFwdAny<TIO>{a, obj}, f2, (int64_t)a.nz * a.n, st);
  if (e != hipSuccess) return e;
  const int C = any_cols_c(N2, any_max_lds()), groups = (N2 + C - 1) / C;
  if ((e = any_launch_cols(TomoColsFwdAny{a, C, groups}, f2, C, (int64_t)a.nz * groups, st)) != hipSuccess) return e;
  hipLaunchKernelGGL(tomo_tile_gather_k, dim3((unsigned)a.nitems, (unsigned)((a.nz + kTZ - 1) / kTZ)), dim3(256), 0,
                     st, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return any_launch_rows(TomoLineIfftAny<TIO>{a, out}, f1, (int64_t)a.ntheta * a.nz, st);
}

template <typename TIO>
static hipError_t tomo_adj_any(const TomoArgs& a, const TIO* data, TIO* out, hipStream_t st) {
  AnyFft f2, f1;
  if (!tomo_any_ffts(a, &f2, &f1)) return hipErrorInvalidValue;
  const int N2 = 2 * a.n;
  hipError_t e = any_launch_rows(TomoLineFftAny<TIO>{a, data}, f1, (int64_t)a.ntheta * a.nz, st);
  if (e != hipSuccess) return e;
  if ((e = launch_grid_gather(a, st)) != hipSuccess) return e;
  const int C = any_cols_c(N2, any_max_lds()), groups = (N2 + C - 1) / C;
  if ((e = any_launch_cols(TomoColsAdjAny{a, C, groups}, f2, C, (int64_t)a.nz * groups, st)) != hipSuccess) return e;
  return any_launch_rows(TomoRowsAdjAny<TIO>{a, out}, f2, (int64_t)a.nz * a.n, st);
}

static bool tomo_pow2(int64_t n) { return n >= 8 && n <= 1024 && (n & (n - 1)) == 0; }

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
  if (a.items) {
    hipLaunchKernelGGL(tomo_tile_gather_k, dim3((unsigned)a.nitems, (unsigned)((a.nz + kTZ - 1) / kTZ)), dim3(256),
                       0, st, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL((tomo_line_ifft_k<N, EL, TIO>), dim3((unsigned)((lines + RL - 1) / RL)), dim3(256), 0, st, a,
                       out);
  } else {
    hipLaunchKernelGGL((tomo_gather_k<N, EL, TIO>), dim3((unsigned)((lines + RL - 1) / RL)), dim3(256), 0, st, a,
                       out);
  }
  return hipGetLastError();
}

template <int N2, typename TIO>
static hipError_t tomo_adj_n(const TomoArgs& a, const TIO* data, TIO* out, hipStream_t st) {
  constexpr int E = 8, N = N2 / 2;
  constexpr int C = tomo_cols_c<N2>();
  constexpr int EL = N >= 8 ? 8 : N;
  constexpr int RL = 256 / (N / EL);
  const int64_t lines = (int64_t)a.ntheta * a.nz;
  hipLaunchKernelGGL((tomo_linefft_k<N, EL, TIO>), dim3((unsigned)((lines + RL - 1) / RL)), dim3(256), 0, st, a,
                     data);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if ((e = launch_grid_gather(a, st)) != hipSuccess) return e;
  hipLaunchKernelGGL((tomo_cols_adj_k<N2, E, C>), dim3((unsigned)(a.nz * (N2 / C))), dim3(C * (N2 / E)), 0, st, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const int64_t rows = (int64_t)a.nz * N;
  hipLaunchKernelGGL((tomo_rows_adj_k<N2, E, TIO>), dim3((unsigned)((rows + 256 / (N2 / E) - 1) / (256 / (N2 / E)))),
                     dim3(256), 0, st, a, out);
  return hipGetLastError();
}

template <int N, typename TIO>
static hipError_t tomo_filter_n(const TIO* in, TIO* out, int64_t lines, const float2* tw, const float* h,
                                hipStream_t st) {
  constexpr int E = N >= 8 ? 8 : N;
  constexpr int RL = 256 / (N / E);
  if (lines == 0) return hipSuccess;
  hipLaunchKernelGGL((tomo_filter_k<N, E, TIO>), dim3((unsigned)((lines + RL - 1) / RL)), dim3(256), 0, st, in, out,
                     lines, tw, h);
  return hipGetLastError();
}

template <typename TIO>
static hipError_t tomo_filter_dispatch(int n, const void* in, void* out, int64_t lines, const float2* tw,
                                       const float* h, hipStream_t st) {
  if (!tomo_pow2(n)) {
    AnyFft f;
    if (!any_fft_factor(n, &f)) return hipErrorInvalidValue;
    f.tw = tw;
    return any_launch_rows(TomoFilterAny<TIO>{(const TIO*)in, (TIO*)out, n, h}, f, lines, st);
  }
#define HTAMD_TFILT_CASE(NN) \
  case NN: return tomo_filter_n<NN, TIO>((const TIO*)in, (TIO*)out, lines, tw, h, st);
  switch (n) {
    HTAMD_TFILT_CASE(8) HTAMD_TFILT_CASE(16) HTAMD_TFILT_CASE(32) HTAMD_TFILT_CASE(64) HTAMD_TFILT_CASE(128)
    HTAMD_TFILT_CASE(256) HTAMD_TFILT_CASE(512) HTAMD_TFILT_CASE(1024)
    default: return hipErrorInvalidValue;
  }
#undef HTAMD_TFILT_CASE
}

template <typename TIO>
static hipError_t tomo_dispatch(bool adjoint, const TomoArgs& a, const void* in, void* out, hipStream_t st) {
  if (!tomo_pow2(a.n))
    return adjoint ? tomo_adj_any<TIO>(a, (const TIO*)in, (TIO*)out, st)
                   : tomo_fwd_any<TIO>(a, (const TIO*)in, (TIO*)out, st);
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
  float* phi = nullptr;
  float2* tw2 = nullptr;
  float2* tw1 = nullptr;
  int2* tb = nullptr;
  float* tw = nullptr;
  int* rp = nullptr;
  int2* csw = nullptr;
  int4* items = nullptr;
  int2* so = nullptr;
  float2* sf = nullptr;
  int* heavy = nullptr;
  int nitems = 0, ts = 0, nheavy = 0;
  // detector filters (ramp, Shepp-Logan, Parzen): [3][n] with the 1/n of the
  // inverse transform for htamd_tomo_filter, and [3][n] shifted by n/2 and
  // scaled by sqrt(n / ntheta) for htamd_tomo_fbp
  float* filt = nullptr;
  float* filt_fbp = nullptr;
};

namespace htamd {

constexpr int kTomoFilters = 3;
// largest slice width: the 2n-point lines of the oversampled grid (two LDS
// buffers of 2n complex values) fit the 160 KiB of LDS up to n = 4096
constexpr int64_t kTomoAnyMax = 4096;

// h[k] on fftfreq(n) (src/holotomocupy/tomo.py:131-155), FP64:
// ramp 4n|f|, Shepp-Logan 4n|f| sinc(f), Parzen 4n|f| w(u), u = 2 (4n|f|),
// w = 1 - 6u^2 + 6u^3 (u <= 1/2) else 2 (1 - u)^3 -- the reference's u.
static double tomo_filter_h(int kind, int64_t k, int64_t n) {
  const double f = (double)(k < (n + 1) / 2 ? k : k - n) / (double)n;
  const double af = std::fabs(f) * 4.0 * (double)n;
  if (kind == 0) return af;
  if (kind == 1) return f == 0.0 ? af : af * std::sin(M_PI * f) / (M_PI * f);
  const double u = 2.0 * af;
  return af * (u <= 0.5 ? 1.0 - 6.0 * u * u + 6.0 * u * u * u : 2.0 * (1.0 - u) * (1.0 - u) * (1.0 - u));
}

}  // namespace htamd

extern "C" {

int htamd_tomo_plan_create(int64_t n, int64_t ntheta, const double* theta, const double* mask,
                           htamd_tomo_plan** plan) {
  if (!plan || !theta) return fail(HTAMD_ERR_ARG, "tomo_plan_create: NULL argument");
  *plan = nullptr;
  {
    AnyFft f2, f1;
    const bool any_ok = n % 2 == 0 && n >= 8 && n <= kTomoAnyMax && any_fft_factor(2 * n, &f2) &&
                        any_fft_factor(n, &f1);
    if (!tomo_pow2(n) && !any_ok)
      return fail(HTAMD_ERR_UNSUPPORTED, "tomo_plan_create: n must be even in [8, " + std::to_string(kTomoAnyMax) +
                                             "] (powers of two up to 1024 on the fastest kernels), got " +
                                             std::to_string(n));
  }
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
  const int L = 2 * p->m + 1;
  const int64_t S = ntheta * n, n2 = 2 * n;
  if (S * L * L > (int64_t)INT32_MAX) {
    delete p;
    return fail(HTAMD_ERR_UNSUPPORTED, "tomo_plan_create: too many gridding taps (ntheta n (2m+1)^2 >= 2^31)");
  }
  hipError_t e = hipGetDevice(&p->device);
  // Gridding taps of every polar sample s = k n + t (cuda_kernels.py:17-51) in
  // FP64: angles rounded to float32 as the reference stores them,
  // x0 = (t - n/2) cos / n, y0 = -(t - n/2) sin / n, first tap floor(2n x0) - m,
  // weights exp(-pi^2/mu d^2) (x) and pi/mu exp(-pi^2/mu d^2) (y).
  std::vector<int2> tb((size_t)S);
  std::vector<float> tw((size_t)(2 * L * S));
  const double cy = M_PI / p->mu, cx = -M_PI * M_PI / p->mu, inv2n = 1.0 / (double)n2;
  for (int64_t k = 0; k < ntheta; ++k) {
    const double th = (double)(float)theta[k];
    const double c = std::cos(th), sn = std::sin(th);
    for (int64_t t = 0; t < n; ++t) {
      const int64_t si = k * n + t;
      const double x0 = (double)(t - n / 2) / (double)n * c;
      const double y0 = -(double)(t - n / 2) / (double)n * sn;
      const int bx = (int)std::floor((double)n2 * x0) - p->m;
      const int by = (int)std::floor((double)n2 * y0) - p->m;
      tb[(size_t)si] = make_int2(bx, by);
      for (int i = 0; i < L; ++i) {
        const double dx = (bx + i) * inv2n - x0, dy = (by + i) * inv2n - y0;
        tw[(size_t)(i * S + si)] = (float)std::exp(cx * dx * dx);
        tw[(size_t)((L + i) * S + si)] = (float)(cy * std::exp(cx * dy * dy));
      }
    }
  }
  // The same taps grouped by grid point (CSR, counting sort; within a row the
  // samples appear in increasing order), for the adjoint's gather.
  const int64_t ng = n2 * n2;
  std::vector<int> rp((size_t)(ng + 1), 0);
  auto wrap = [&](int64_t v) { return (int)((n + v + n2) % n2); };
  for (int64_t si = 0; si < S; ++si)
    for (int i1 = 0; i1 < L; ++i1)
      for (int i0 = 0; i0 < L; ++i0)
        ++rp[(size_t)(wrap(tb[(size_t)si].y + i1) * n2 + wrap(tb[(size_t)si].x + i0) + 1)];
  for (int64_t g = 0; g < ng; ++g) rp[(size_t)(g + 1)] += rp[(size_t)g];
  // grid points with more than kHeavyRow taps, longest first (ties by index)
  std::vector<int> heavy;
  for (int64_t g = 0; g < ng; ++g)
    if (rp[(size_t)(g + 1)] - rp[(size_t)g] > kHeavyRow) heavy.push_back((int)g);
  std::stable_sort(heavy.begin(), heavy.end(), [&](int x, int y) {
    return rp[(size_t)x + 1] - rp[(size_t)x] > rp[(size_t)y + 1] - rp[(size_t)y];
  });
  std::vector<int> fill(rp.begin(), rp.end() - 1);
  std::vector<int2> csw((size_t)(S * L * L));
  for (int64_t si = 0; si < S; ++si)
    for (int i1 = 0; i1 < L; ++i1) {
      const float wy = tw[(size_t)((L + i1) * S + si)];
      for (int i0 = 0; i0 < L; ++i0) {
        const int64_t g = wrap(tb[(size_t)si].y + i1) * n2 + wrap(tb[(size_t)si].x + i0);
        const int k = fill[(size_t)g]++;
        const float w = wy * tw[(size_t)(i0 * S + si)];
        int wb;
        std::memcpy(&wb, &w, sizeof wb);
        csw[(size_t)k] = make_int2((int)si, wb);
      }
    }
  // Tile binning of the samples for the forward gather: tile of the window's
  // first tap (wrapped), sorted by tile (counting sort, samples in increasing
  // order within a tile), cut into work items of at most 256 samples.
  const int ts = (int)std::min<int64_t>(32, n2);
  const int ntx = (int)((n2 + ts - 1) / ts);  // the last tile is partial when ts does not divide 2n
  std::vector<int> tcount((size_t)(ntx * ntx + 1), 0);
  for (int64_t si = 0; si < S; ++si)
    ++tcount[(size_t)((wrap(tb[(size_t)si].y) / ts) * ntx + wrap(tb[(size_t)si].x) / ts + 1)];
  for (int i = 0; i < ntx * ntx; ++i) tcount[(size_t)(i + 1)] += tcount[(size_t)i];
  std::vector<int> tfill(tcount.begin(), tcount.end() - 1);
  std::vector<int2> so((size_t)S);
  std::vector<float2> sf((size_t)S);
  for (int64_t si = 0; si < S; ++si) {
    const int wx0 = wrap(tb[(size_t)si].x), wy0 = wrap(tb[(size_t)si].y);
    const int tile = (wy0 / ts) * ntx + wx0 / ts;
    const int k = tfill[(size_t)tile]++;
    so[(size_t)k] = make_int2(((wy0 % ts) << 16) | (wx0 % ts), (int)si);
    const int64_t kk = si / n, t = si % n;
    const double th = (double)(float)theta[kk];
    const double x2 = (double)n2 * ((double)(t - n / 2) / (double)n * std::cos(th));
    const double y2 = (double)n2 * (-(double)(t - n / 2) / (double)n * std::sin(th));
    sf[(size_t)k] = make_float2((float)(x2 - std::floor(x2)), (float)(y2 - std::floor(y2)));
  }
  std::vector<int4> items;
  for (int tile = 0; tile < ntx * ntx; ++tile)
    for (int b = tcount[(size_t)tile]; b < tcount[(size_t)(tile + 1)]; b += 256)
      items.push_back(make_int4(tile % ntx, tile / ntx, b, std::min(256, tcount[(size_t)(tile + 1)] - b)));
  p->nitems = (int)items.size();
  p->nheavy = (int)heavy.size();
  p->ts = ts;
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
  std::vector<float> filt((size_t)(kTomoFilters * n)), filt_fbp((size_t)(kTomoFilters * n));
  const double norm_const = std::sqrt((double)n / (double)ntheta);
  for (int kind = 0; kind < kTomoFilters; ++kind)
    for (int64_t k = 0; k < n; ++k) {
      filt[(size_t)(kind * n + k)] = (float)(tomo_filter_h(kind, k, n) / (double)n);
      // FFT(c1d d)[k] picks up h at bin (k - n/2) mod n
      filt_fbp[(size_t)(kind * n + k)] = (float)(tomo_filter_h(kind, (k + n / 2) % n, n) * norm_const);
    }
  auto twid = [](int64_t N) {
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
  if (e != hipSuccess || (e = up(&p->phi, phi)) != hipSuccess || (e = up(&p->tw2, twid(2 * n))) != hipSuccess ||
      (e = up(&p->tw1, twid(n))) != hipSuccess || (e = up(&p->tb, tb)) != hipSuccess ||
      (e = up(&p->tw, tw)) != hipSuccess || (e = up(&p->rp, rp)) != hipSuccess || (e = up(&p->csw, csw)) != hipSuccess ||
      (e = up(&p->items, items)) != hipSuccess ||
      (!heavy.empty() && (e = up(&p->heavy, heavy)) != hipSuccess) ||
      (e = up(&p->so, so)) != hipSuccess || (e = up(&p->sf, sf)) != hipSuccess ||
      (e = up(&p->filt, filt)) != hipSuccess || (e = up(&p->filt_fbp, filt_fbp)) != hipSuccess) {
    htamd_tomo_plan_destroy(p);
    return hip_fail(e, "tomo_plan_create");
  }
  *plan = p;
  return HTAMD_OK;
}

int htamd_tomo_plan_destroy(htamd_tomo_plan* p) {
  if (!p) return HTAMD_OK;
  (void)hipFree(p->phi);
  (void)hipFree(p->tw2);
  (void)hipFree(p->tw1);
  (void)hipFree(p->tb);
  (void)hipFree(p->tw);
  (void)hipFree(p->rp);
  (void)hipFree(p->csw);
  (void)hipFree(p->items);
  (void)hipFree(p->so);
  (void)hipFree(p->sf);
  (void)hipFree(p->heavy);
  (void)hipFree(p->filt);
  (void)hipFree(p->filt_fbp);
  delete p;
  return HTAMD_OK;
}

int htamd_tomo_workspace_size(const htamd_tomo_plan* p, int64_t nz, int64_t* bytes) {
  if (!p || !bytes || nz < 0) return fail(HTAMD_ERR_ARG, "tomo_workspace_size: bad arguments");
  *bytes = nz * (p->n * 2 * p->n + 4 * p->n * p->n) * 8 + p->ntheta * p->n * nz * 8;
  return HTAMD_OK;
}

static int tomo_run(const htamd_tomo_plan* p, bool adjoint, const void* in, int is_complex, void* out,
                    int64_t nz, void* ws, int64_t ws_bytes, void* stream, const char* what, int filter = -1) {
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
  a.phi = p->phi;
  a.tb = p->tb;
  a.tw = p->tw;
  a.rp = p->rp;
  a.csw = p->csw;
  a.tw2 = p->tw2;
  a.tw1 = p->tw1;
  a.items = p->m == (kTL - 1) / 2 ? p->items : nullptr;  // the tiled gather is built for 9x9 windows
  a.so = p->so;
  a.sf = p->sf;
  a.ts = p->ts;
  a.nitems = p->nitems;
  a.heavy = p->heavy;
  a.nheavy = p->nheavy;
  a.cxs = (float)(-M_PI * M_PI / p->mu / (4.0 * (double)p->n * (double)p->n));
  a.cys = (float)(M_PI / p->mu);
  a.W = (float2*)ws;
  a.G = (float2*)ws + nz * p->n * 2 * p->n;
  a.Sh = a.G + nz * 4 * p->n * p->n;
  if (filter >= 0) a.filt = p->filt_fbp + (int64_t)filter * p->n;
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

int htamd_tomo_fbp(const htamd_tomo_plan* p, const void* data, int is_complex, int filter, void* out, int64_t nz,
                   void* workspace, int64_t workspace_bytes, void* stream) {
  if (filter < 0 || filter >= kTomoFilters)
    return fail(HTAMD_ERR_ARG, "tomo_fbp: filter must be 0 (ramp), 1 (shepp) or 2 (parzen)");
  return tomo_run(p, true, data, is_complex, out, nz, workspace, workspace_bytes, stream, "tomo_fbp", filter);
}

int htamd_tomo_filter(const htamd_tomo_plan* p, const void* data, int is_complex, int filter, void* out, int64_t nz,
                      void* stream) {
  if (!p) return fail(HTAMD_ERR_ARG, "tomo_filter: plan is NULL");
  if (filter < 0 || filter >= kTomoFilters)
    return fail(HTAMD_ERR_ARG, "tomo_filter: filter must be 0 (ramp), 1 (shepp) or 2 (parzen)");
  if (nz < 0) return fail(HTAMD_ERR_ARG, "tomo_filter: negative slice count");
  if (nz == 0) return HTAMD_OK;
  if (!data || !out) return fail(HTAMD_ERR_ARG, "tomo_filter: NULL buffer");
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != p->device) return fail(HTAMD_ERR_ARG, "tomo_filter: plan belongs to another device");
  const float* h = p->filt + (int64_t)filter * p->n;
  const int64_t lines = p->ntheta * nz;
  hipError_t e = is_complex ? tomo_filter_dispatch<float2>((int)p->n, data, out, lines, p->tw1, h, (hipStream_t)stream)
                            : tomo_filter_dispatch<float>((int)p->n, data, out, lines, p->tw1, h, (hipStream_t)stream);
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "tomo_filter");
}

}  // extern "C"
