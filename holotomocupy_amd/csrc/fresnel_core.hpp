// Device core of the separable Fresnel propagator shared by the operator
// kernels (fresnel.hip) and the fused BH gradient (bh.hip).
//
// Each axis of D = crop(IFFT2(FFT2(sympad psi) fker)) (src/holotomocupy/propagation.py:70-112)
// is C^-1 diag(lambda_k) C with C the length-n DCT-II, evaluated through a
// length-n complex FFT in LDS (Makhoul reordering) and a 2x2 map on each bin
// pair (k, n-k); see fresnel.hip for the derivation and capi.hip for the FP64
// tables.
#pragma once
#include "fft_lds.hpp"

namespace htamd {

// Pair map on the FFT output held in LDS (see file comment).  Table layout:
// tab[k] = A_k, tab[H+k] = B_k, tab[2H+k] = C_k (H = N/2), except that the two
// self-paired bins use tab[0] = l_0/N and tab[H] = l_{N/2}/N.
template <int N, int E, bool WAVE = false>
__device__ __forceinline__ void dct_pair_map(float2* s, int t, const float2* __restrict__ tab) {
  constexpr int T = N / E;
  constexpr int H = N / 2;
#pragma unroll
  for (int q = 0; q < E / 2; ++q) {
    const int k = t + q * T;
    if (k == 0) {
      s[0] = cmul(tab[0], s[0]);
      s[fpad(H)] = cmul(tab[H], s[fpad(H)]);
    } else {
      const int pk = fpad(k), pm = fpad(N - k);
      const float2 v1 = s[pk], v2 = s[pm];
      const float2 a = tab[k], b = tab[H + k], c = tab[2 * H + k];
      s[pk] = cadd(cmul(a, v1), cmul(b, v2));
      s[pm] = cadd(cmul(c, v1), cmul(a, v2));
    }
    // keep the compiler from hoisting all 3 E/2 table loads at once (registers)
    if (q % 4 == 3) asm volatile("" ::: "memory");
  }
  lds_sync<WAVE>();
}

// epi: output hook of the inverse's last stage (see NoEpi).
template <int N, int E, int RM = kMaxRadix, bool WAVE = false, typename EPI = NoEpi>
__device__ __forceinline__ void fresnel_1d(float2* s, int t, const FftTw<N, E, RM>& tw,
                                           const float2* __restrict__ tab, const EPI& epi = EPI()) {
  fft_lds<N, E, false, RM, WAVE>(s, t, tw);
  dct_pair_map<N, E, WAVE>(s, t, tab);
  fft_lds<N, E, true, RM, WAVE, EPI>(s, t, tw, epi);
}

// fresnel_1d of a line whose FFT inputs the thread holds: h[e] is FFT slot
// t + e N/E (the forward transform's first stage runs on them in registers,
// so the line is never written to s in slot order first).
template <int N, int E, int RM = kMaxRadix, bool WAVE = false, typename EPI = NoEpi>
__device__ __forceinline__ void fresnel_1d_from(const float2 (&h)[E], float2* s, int t, const FftTw<N, E, RM>& tw,
                                                const float2* __restrict__ tab, const EPI& epi = EPI()) {
  constexpr int R0 = stage_radix<N, E, 1, RM>();
  stockham_first_from<N, E, false, RM, WAVE>(h, s, t);
  stockham<N, E, R0, 0, false, RM, WAVE>(s, t, tw.w);
  dct_pair_map<N, E, WAVE>(s, t, tab);
  fft_lds<N, E, true, RM, WAVE, EPI>(s, t, tw, epi);
}

// Output "hook" that only drops the last stage's closing ordering point, for
// a caller that reads back nothing but its own outputs (element t + e N/E at
// s[fpad(t + e N/E)]): a thread's LDS accesses complete in order.
struct OwnOutputs {
  static constexpr bool active = false, store = true, sync_after = false;
  __device__ __forceinline__ float2 operator()(int, float2 v) const { return v; }
};

// Output hook that writes a line's result to global memory in natural order:
// element t + e N/E of the transform is natural element dct_nat of it.
template <int N>
struct LineStore {
  static constexpr bool active = true, store = false, sync_after = false;
  float2* dst;  // the line, natural order; nullptr: not stored
  int t, T;
  __device__ __forceinline__ float2 operator()(int e, float2 v) const {
    if (dst) dst[dct_nat<N>(t + e * T)] = v;
    return v;
  }
};

// The pair map from src to dst (each thread reads and writes its own pairs).
template <int N, int E, bool WAVE = false>
__device__ __forceinline__ void dct_pair_map_pp(const float2* src, float2* dst, int t,
                                                const float2* __restrict__ tab) {
  constexpr int T = N / E;
  constexpr int H = N / 2;
#pragma unroll
  for (int q = 0; q < E / 2; ++q) {
    const int k = t + q * T;
    if (k == 0) {
      dst[0] = cmul(tab[0], src[0]);
      dst[fpad(H)] = cmul(tab[H], src[fpad(H)]);
    } else {
      const int pk = fpad(k), pm = fpad(N - k);
      const float2 v1 = src[pk], v2 = src[pm];
      const float2 a = tab[k], b = tab[H + k], c = tab[2 * H + k];
      dst[pk] = cadd(cmul(a, v1), cmul(b, v2));
      dst[pm] = cadd(cmul(c, v1), cmul(a, v2));
    }
    if (q % 4 == 3) asm volatile("" ::: "memory");
  }
  lds_sync<WAVE>();
}

// fresnel_1d of a with b as the second buffer (one ordering point per FFT
// stage): the transform, the pair map and the inverse move the data 2 S + 1
// times (S stages each), an odd count, so the result is always in b.  epi:
// output hook of the inverse's last stage (a pointwise step on the result,
// see NoEpi).
template <int N, int E, int RM = kMaxRadix, bool WAVE = false, typename EPI = NoEpi>
__device__ __forceinline__ void fresnel_1d_pp(float2* a, float2* b, int t, const FftTw<N, E, RM>& tw,
                                              const float2* __restrict__ tab, const EPI& epi = EPI()) {
  constexpr bool odd = fft_stages<N, E, RM>() % 2 == 1;
  fft_lds_pp<N, E, false, RM, WAVE>(a, b, t, tw);  // -> odd ? b : a
  if constexpr (odd) {
    dct_pair_map_pp<N, E, WAVE>(b, a, t, tab);
    fft_lds_pp<N, E, true, RM, WAVE, EPI>(a, b, t, tw, epi);  // a -> b
  } else {
    dct_pair_map_pp<N, E, WAVE>(a, b, t, tab);
    fft_lds_pp<N, E, true, RM, WAVE, EPI>(b, a, t, tw, epi);  // b -> b (even count)
  }
}

// Order in which a row-pass thread visits its E slots t + e N/E when it
// writes their natural elements dct_nat(t + e N/E) to HBM: 0, E-1, 1, E-2, ...
// Natural elements 2m and 2m + 1 share a cache line and belong to slots e and
// E-1-e of threads t and N/E-1-t, so in this order the two halves of a line
// are written in consecutive iterations instead of E-1 apart, and lines are
// not written back half-filled (K3: 2722 -> 2178 MiB written per 16-angle
// launch at 2048^2, the algorithmic figure).
template <int E>
__device__ __forceinline__ int pair_order(int k) {
  return (k & 1) ? E - 1 - (k >> 1) : (k >> 1);
}

// Column-pass elements per thread: 32 at N = 2048 (64 threads per column, so a
// 4-column tile is one 256-thread workgroup), else 8.
template <int N>
constexpr int col_e() {
  return N >= 2048 ? N / 64 : 8;
}

constexpr int kRowThreads = 256;

// Global rows of a column transform held in Makhoul order.  FFT slot
// q = q0 + e*T (T = N/E, 0 <= q0 < T, e = 0..E-1) holds natural row dct_nat(q),
// which is 2q for e < E/2 and 2N-1-2q for e >= E/2: two affine runs in e, so
// the element offsets row*n + c need one multiply-add each instead of the
// index arithmetic of dct_nat.
// Offsets are 32-bit: one field (N rows of n elements) stays below 2^31.
template <int N, int E>
struct MakhoulRows {
  int a, b, step;
  __device__ __forceinline__ MakhoulRows(int q0, int c, int n)
      : a(2 * q0 * n + c), b((2 * N - 1 - 2 * q0) * n + c), step(2 * (N / E) * n) {}
  __device__ __forceinline__ int off(int e) const { return e < E / 2 ? a + e * step : b - e * step; }
};

// LDS slot fpad(q0 + e*T) of the same element: affine in e when T is a
// multiple of 16 (fpad adds one slot per 16).
template <int N, int E>
__device__ __forceinline__ int slot_run(int q0, int e) {
  constexpr int T = N / E;
  if constexpr (T % 16 == 0)
    return fpad(q0) + e * (T / 16 * 17);
  else
    return fpad(q0 + e * T);
}

// LDS slot fpad(dct_slot(t + e*T)) of row element t + e*T (T = N/E even): the
// element's parity is t's, so its FFT slot runs by +-T/2 per e, and by
// +-17T/32 after padding when T is a multiple of 32.
template <int N, int E>
__device__ __forceinline__ int row_slot(int t, int e) {
  constexpr int T = N / E;
  if constexpr (T % 32 == 0) {
    constexpr int st = T / 32 * 17;
    return fpad(dct_slot<N>(t)) + ((t & 1) ? -e * st : e * st);
  } else {
    return fpad(dct_slot<N>(t + e * T));
  }
}

// Move a C-column tile between global memory (rows of stride n, Makhoul
// order down the columns) and the LDS tile (column c at c * LS): element e of
// thread tid is column tid % C, slot tid / C + e*T.
template <int N, int E, int C, bool LOAD>
__device__ __forceinline__ void col_tile_xfer(float2* lds, float2* base, int n, int tid) {
  constexpr int LS = col_stride<N, C>();
  const int c = tid % C, q0 = tid / C;
  const MakhoulRows<N, E> rows(q0, c, n);
  float2* l = lds + c * LS;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if constexpr (LOAD)
      l[slot_run<N, E>(q0, e)] = base[rows.off(e)];
    else
      base[rows.off(e)] = l[slot_run<N, E>(q0, e)];
  }
}

}  // namespace htamd
