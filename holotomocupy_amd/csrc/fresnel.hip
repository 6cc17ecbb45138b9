// Separable Fresnel propagator D / DT for gfx950.
//
// The reference builds D as crop(IFFT2(FFT2(sympad(psi)) * fker)) on a
// 2nz x 2n grid (src/holotomocupy/propagation.py:70-112).  The transfer
// function and the symmetric pad / crop factor per axis, and a symmetric
// pad followed by a length-2n FFT is a DCT-II, so each axis reduces to
//     D_x = C^-1 diag(lambda_k) C,   lambda_k = exp(-i pi lambda z (k/(2 n vs))^2)
// with C the length-n DCT-II.  C is evaluated through a length-n complex FFT
// (Makhoul reordering); the post-twiddle of C, the chirp and the pre-twiddle
// of C^-1 collapse into one 2x2 map on the bin pair (k, n-k):
//     U_k   = A_k V_k + B_k V_{n-k}
//     U_n-k = C_k V_k + A_k V_{n-k}
// with A_k = (l_k + l_{n-k})/2n, B_k = (l_k - l_{n-k}) e^{+i pi k/n}/2n,
// C_k = (l_k - l_{n-k}) e^{-i pi k/n}/2n, tables built in FP64 on the host.
// DT uses the tables of conj(lambda).  Each axis is one HBM pass: read the
// field, FFT in LDS, pair map, inverse FFT in LDS, write the field.
#include <mutex>

#include "fft_any.hpp"
#include "fresnel_core.hpp"
#include "htamd_internal.hpp"

namespace htamd {


// Rows of length N, ROWS = 256/(N/E) rows per workgroup.
template <int N, int E>
__global__ __launch_bounds__(kRowThreads) void fresnel_rows_k(const float2* in, int64_t in_stride,
                                                              float2* out, int64_t out_stride,
                                                              int nz, int64_t nrows,
                                                              const float2* __restrict__ tw,
                                                              const float2* __restrict__ tab) {
  constexpr int T = N / E;
  constexpr int ROWS = kRowThreads / T;
  __shared__ float2 lds[ROWS * fft_buf<N>()];
  const int r = threadIdx.x / T, t = threadIdx.x % T;
  const int64_t row = (int64_t)blockIdx.x * ROWS + r;
  const bool valid = row < nrows;
  const int64_t fb = valid ? row / nz : 0, fy = valid ? row % nz : 0;
  float2* s = lds + r * fft_buf<N>();
  const float2* src = in + fb * in_stride + fy * N;
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, tw);
  // FFT slots t + e T of the row (natural elements dct_nat of them) straight
  // into registers for the first transform stage
  float2 h[E];
#pragma unroll
  for (int e = 0; e < E; ++e) h[e] = valid ? src[dct_nat<N>(t + e * T)] : make_float2(0.f, 0.f);
  // the last stage writes the row to out in natural order from registers
  fresnel_1d_from<N, E, kMaxRadix, false, LineStore<N>>(
      h, s, t, twr, tab, LineStore<N>{valid ? out + fb * out_stride + fy * N : nullptr, t, T});
}

// Columns of length N (= nz) of a [batch, N, n] field, C adjacent columns per
// workgroup, in place.  Workgroups are mapped XCD-contiguously so neighbouring
// column tiles (which share cache lines) run under the same L2.
template <int N, int E, int C>
__global__ __launch_bounds__(C*(N / E)) void fresnel_cols_k(float2* data, int64_t stride, int n,
                                                            int tiles_per_field,
                                                            const float2* __restrict__ tw,
                                                            const float2* __restrict__ tab) {
  constexpr int T = N / E;
  constexpr int NT = C * T;
  constexpr int LS = col_stride<N, C>();
  __shared__ float2 lds[C * LS];
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t b = tile / tiles_per_field;
  const int x0 = (int)(tile % tiles_per_field) * C;
  float2* base = data + b * stride + x0;
  const int cc = threadIdx.x / T, t = threadIdx.x % T;
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, tw);
  col_tile_xfer<N, E, C, true>(lds, base, n, threadIdx.x);  // FFT slot q holds row dct_nat(q)
  __syncthreads();
  fresnel_1d<N, E>(lds + cc * LS, t, twr, tab);
  col_tile_xfer<N, E, C, false>(lds, base, n, threadIdx.x);
}

bool fft_size_supported(int64_t n) {
  return n >= kFftMin && n <= kFftMax && (n & (n - 1)) == 0;
}

bool fresnel_any_supported(int64_t n) {
  AnyFft f;
  return n >= kFftMin && n <= kAnyFresnelMax && n % 2 == 0 && any_fft_factor(n, &f) &&
         (int64_t)2 * n * (int64_t)sizeof(float2) <= any_max_lds();
}

int64_t any_max_lds() {
  static std::mutex mu;
  static int64_t cache[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 65536;
  std::lock_guard<std::mutex> lock(mu);
  if (!cache[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) v = 65536;
    cache[dev] = v;
  }
  return cache[dev];
}

// ---- any even length (fft_any.hpp) ---------------------------------------------
// Row pass: line = (field b, row y), slot p of the Makhoul-ordered line holds
// element dct_nat(p); FFT, pair map, inverse FFT, written back in the same
// order.  in may equal out (a line is read whole before it is written).
struct FresnelRowsAny {
  static constexpr bool kInv = false, kTwo = true;
  const float2* in;
  float2* out;
  int64_t in_stride, out_stride;
  int nz, n;
  const float2* tab;
  __device__ float2 load(int64_t line, int, int p) const {
    return in[(line / nz) * in_stride + (line % nz) * n + dct_nat_rt(p, n)];
  }
  __device__ void mid(float2* s, int64_t, int, int tid, int nthr) const { dct_pair_map_rt(s, n, tab, tid, nthr); }
  __device__ void store(int64_t line, int, int p, float2 v) const {
    out[(line / nz) * out_stride + (line % nz) * n + dct_nat_rt(p, n)] = v;
  }
};

// Column pass in place: unit = (field b, group of C columns from x0).
struct FresnelColsAny {
  static constexpr bool kInv = false, kTwo = true;
  float2* data;
  int64_t stride;
  int nz, n, C, groups;
  const float2* tab;
  __device__ int64_t off(int64_t unit, int c, int p) const {
    return (unit / groups) * stride + (int64_t)dct_nat_rt(p, nz) * n + (unit % groups) * C + c;
  }
  __device__ bool ok(int64_t unit, int c) const { return (unit % groups) * C + c < n; }
  __device__ float2 load(int64_t unit, int c, int p) const {
    return ok(unit, c) ? data[off(unit, c, p)] : make_float2(0.f, 0.f);
  }
  __device__ void mid(float2* s, int64_t, int, int tid, int nthr) const { dct_pair_map_rt(s, nz, tab, tid, nthr); }
  __device__ void store(int64_t unit, int c, int p, float2 v) const {
    if (ok(unit, c)) data[off(unit, c, p)] = v;
  }
};

static hipError_t fresnel_rows_any(int64_t n, int64_t nz, const float2* in, int64_t in_stride, float2* out,
                                   int64_t out_stride, int64_t batch, const float2* tw, const float2* tab,
                                   hipStream_t st) {
  AnyFft f;
  if (!any_fft_factor(n, &f)) return hipErrorInvalidValue;
  f.tw = tw;
  const FresnelRowsAny op{in, out, in_stride, out_stride, (int)nz, (int)n, tab};
  return any_launch_rows(op, f, batch * nz, st);
}

static hipError_t fresnel_cols_any(int64_t nz, int64_t n, float2* data, int64_t stride, int64_t batch,
                                   const float2* tw, const float2* tab, hipStream_t st) {
  AnyFft f;
  if (!any_fft_factor(nz, &f)) return hipErrorInvalidValue;
  f.tw = tw;
  const int C = any_cols_c((int)nz, any_max_lds());
  const int groups = (int)((n + C - 1) / C);
  const FresnelColsAny op{data, stride, (int)nz, (int)n, C, groups, tab};
  return any_launch_cols(op, f, C, batch * groups, st);
}

template <int N>
static hipError_t rows_n(const float2* in, int64_t in_stride, float2* out, int64_t out_stride,
                         int64_t nz, int64_t batch, const float2* tw, const float2* tab,
                         hipStream_t st) {
  constexpr int E = 8;
  constexpr int ROWS = kRowThreads / (N / E);
  const int64_t nrows = batch * nz;
  const int64_t grid = (nrows + ROWS - 1) / ROWS;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((fresnel_rows_k<N, E>), dim3((unsigned)grid), dim3(kRowThreads), 0, st, in,
                     in_stride, out, out_stride, (int)nz, nrows, tw, tab);
  return hipGetLastError();
}

hipError_t launch_fresnel_rows(int64_t n, int64_t nz, const float2* in, int64_t in_stride,
                               float2* out, int64_t out_stride, int64_t batch, const float2* tw,
                               const float2* tab, hipStream_t st) {
  if (!fft_size_supported(n)) return fresnel_rows_any(n, nz, in, in_stride, out, out_stride, batch, tw, tab, st);
#define HTAMD_ROWS_CASE(NN) \
  case NN: return rows_n<NN>(in, in_stride, out, out_stride, nz, batch, tw, tab, st);
  switch (n) {
    HTAMD_ROWS_CASE(8) HTAMD_ROWS_CASE(16) HTAMD_ROWS_CASE(32) HTAMD_ROWS_CASE(64)
    HTAMD_ROWS_CASE(128) HTAMD_ROWS_CASE(256) HTAMD_ROWS_CASE(512) HTAMD_ROWS_CASE(1024)
    HTAMD_ROWS_CASE(2048)
    default: return hipErrorInvalidValue;
  }
#undef HTAMD_ROWS_CASE
}

template <int N, int C>
static hipError_t cols_nc(int64_t n, float2* data, int64_t stride, int64_t batch, const float2* tw,
                          const float2* tab, hipStream_t st) {
  constexpr int E = col_e<N>();
  const int64_t tiles_per_field = n / C;
  const int64_t grid = batch * tiles_per_field;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((fresnel_cols_k<N, E, C>), dim3((unsigned)grid), dim3(C * (N / E)), 0, st,
                     data, stride, (int)n, (int)tiles_per_field, tw, tab);
  return hipGetLastError();
}

// Columns per workgroup: up to 16, bounded by 512 threads and the 160 KiB of
// LDS of a CU, and dividing the row length n.  Column reads and writes are
// C * 8 contiguous bytes per row, so the widest tile that fits is the fastest
// (2048-point columns: 40 us per field at C = 4, 26 us at C = 8).
template <int N, int C>
constexpr bool fcols_fit() {
  return C * (N / col_e<N>()) <= 512 && (int64_t)C * col_stride<N, C>() * 8 <= 160 * 1024;
}

template <int N>
static hipError_t cols_n(int64_t n, float2* data, int64_t stride, int64_t batch, const float2* tw,
                         const float2* tab, hipStream_t st) {
  constexpr int CMAX = fcols_fit<N, 16>() ? 16 : fcols_fit<N, 8>() ? 8 : fcols_fit<N, 4>() ? 4
                       : fcols_fit<N, 2>() ? 2 : 1;
  if (CMAX >= 16 && n % 16 == 0) return cols_nc<N, (CMAX >= 16 ? 16 : 1)>(n, data, stride, batch, tw, tab, st);
  if (CMAX >= 8 && n % 8 == 0) return cols_nc<N, (CMAX >= 8 ? 8 : 1)>(n, data, stride, batch, tw, tab, st);
  if (CMAX >= 4 && n % 4 == 0) return cols_nc<N, (CMAX >= 4 ? 4 : 1)>(n, data, stride, batch, tw, tab, st);
  if (CMAX >= 2 && n % 2 == 0) return cols_nc<N, (CMAX >= 2 ? 2 : 1)>(n, data, stride, batch, tw, tab, st);
  return cols_nc<N, 1>(n, data, stride, batch, tw, tab, st);
}

hipError_t launch_fresnel_cols(int64_t nz, int64_t n, float2* data, int64_t stride, int64_t batch,
                               const float2* tw, const float2* tab, hipStream_t st) {
  if (!fft_size_supported(nz)) return fresnel_cols_any(nz, n, data, stride, batch, tw, tab, st);
#define HTAMD_COLS_CASE(NN) \
  case NN: return cols_n<NN>(n, data, stride, batch, tw, tab, st);
  switch (nz) {
    HTAMD_COLS_CASE(8) HTAMD_COLS_CASE(16) HTAMD_COLS_CASE(32) HTAMD_COLS_CASE(64)
    HTAMD_COLS_CASE(128) HTAMD_COLS_CASE(256) HTAMD_COLS_CASE(512) HTAMD_COLS_CASE(1024)
    HTAMD_COLS_CASE(2048)
    default: return hipErrorInvalidValue;
  }
#undef HTAMD_COLS_CASE
}

}  // namespace htamd
