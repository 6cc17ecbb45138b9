// LDS-resident Stockham FFT for power-of-two lengths on CDNA4 (wave64).
//
// One transform of length N is carried by T = N/E threads; each thread owns E
// complex values per stage and performs E/R radix-R butterflies.  The data
// lives in LDS between stages: in place with two workgroup barriers per stage,
// or (fft_lds_pp) alternating between two buffers with one.
//
// LDS layout: element i of a transform buffer lives in slot fpad(i) = i +
// i/16 (one unused slot after every 16 elements), so the strided writes of the
// early Stockham stages spread over all LDS banks instead of hitting two;
// fft_buf<N>() is the buffer size in float2 slots.  Every access to a buffer
// handed to fft_lds goes through fpad.
//
// Twiddles: each thread needs one base twiddle w = exp(-2 pi i k/(NS R)) per
// butterfly and stage (NS > 1); they are read once per kernel from a table
// tw[m] = exp(-2 pi i m/N) generated in FP64 on the host, held in registers
// (FftTw), and the powers w^2..w^(R-1) are formed with at most three complex
// products.  The inverse transform uses the conjugates.  Output is in natural
// order (Stockham autosort), unnormalised.
#pragma once
#include "htamd_common.hpp"

namespace htamd {

__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }
template <int N>
constexpr int fft_buf() {
  return N + N / 16;
}

// Stride (float2 slots) between the transform buffers of a C-column LDS tile
// that is filled with consecutive lanes running down the columns, 16/C
// elements per column per 16 lanes: c * stride = c * 16/C (mod 16) puts the C
// columns of one 16-lane access on disjoint bank ranges.
template <int N, int C>
constexpr int col_stride() {
  constexpr int B = fft_buf<N>();
  constexpr int want = C >= 16 ? 1 : 16 / C;
  return B + ((want - B % 16) % 16 + 16) % 16;
}

// In-register DFTs.  Forward: X[k] = sum_r a[r] exp(-2 pi i r k / R).
template <bool INV>
__host__ __device__ __forceinline__ void dft2(float2& a0, float2& a1) {
  float2 t = a0;
  a0 = cadd(t, a1);
  a1 = csub(t, a1);
}

template <bool INV>
__host__ __device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
  float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
  float2 t2 = cadd(a1, a3), t3 = mul_mi<INV>(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}

template <bool INV>
__host__ __device__ __forceinline__ void dft8(float2* a) {
  constexpr float h = 0.70710678118654752440f;
  float2 e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
  float2 o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
  dft4<INV>(e0, e1, e2, e3);
  dft4<INV>(o0, o1, o2, o3);
  // o_k *= exp(-+ 2 pi i k / 8)
  o1 = INV ? make_float2(h * (o1.x - o1.y), h * (o1.x + o1.y))
           : make_float2(h * (o1.x + o1.y), h * (o1.y - o1.x));
  o2 = mul_mi<INV>(o2);
  o3 = INV ? make_float2(-h * (o3.x + o3.y), h * (o3.x - o3.y))
           : make_float2(h * (o3.y - o3.x), -h * (o3.x + o3.y));
  a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
  a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
  a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
  a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
}

// multiply by W16^m = exp(-+ 2 pi i m / 16), m a compile-time constant
template <int m, bool INV>
__device__ __forceinline__ float2 mul_w16(float2 v) {
  constexpr int q = m & 15;
  if constexpr (q == 0) {
    return v;
  } else if constexpr (q == 4) {
    return mul_mi<INV>(v);
  } else if constexpr (q == 8) {
    return make_float2(-v.x, -v.y);
  } else {
    // cos(2 pi q / 16), q = 0..15
    constexpr double kC[16] = {1.0, 0.92387953251128674, 0.70710678118654752, 0.38268343236508977,
                               0.0, -0.38268343236508977, -0.70710678118654752, -0.92387953251128674,
                               -1.0, -0.92387953251128674, -0.70710678118654752, -0.38268343236508977,
                               0.0, 0.38268343236508977, 0.70710678118654752, 0.92387953251128674};
    constexpr float c = (float)kC[q];
    constexpr float sn = (float)kC[(q + 12) & 15];  // sin(x) = cos(x - pi/2)
    constexpr float s = INV ? sn : -sn;
    return make_float2(c * v.x - s * v.y, c * v.y + s * v.x);
  }
}

// radix 16 as 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2
template <bool INV>
__device__ __forceinline__ void dft16(float2* a) {
  float2 y[16];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    float2 b0 = a[n2], b1 = a[4 + n2], b2 = a[8 + n2], b3 = a[12 + n2];
    dft4<INV>(b0, b1, b2, b3);  // over n1 -> k1
    y[n2 * 4 + 0] = b0;
    y[n2 * 4 + 1] = b1;
    y[n2 * 4 + 2] = b2;
    y[n2 * 4 + 3] = b3;
  }
  // W16^(n2 k1)
  y[5] = mul_w16<1, INV>(y[5]);
  y[6] = mul_w16<2, INV>(y[6]);
  y[7] = mul_w16<3, INV>(y[7]);
  y[9] = mul_w16<2, INV>(y[9]);
  y[10] = mul_w16<4, INV>(y[10]);
  y[11] = mul_w16<6, INV>(y[11]);
  y[13] = mul_w16<3, INV>(y[13]);
  y[14] = mul_w16<6, INV>(y[14]);
  y[15] = mul_w16<9, INV>(y[15]);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 b0 = y[k1], b1 = y[4 + k1], b2 = y[8 + k1], b3 = y[12 + k1];
    dft4<INV>(b0, b1, b2, b3);  // over n2 -> k2
    a[k1] = b0;
    a[k1 + 4] = b1;
    a[k1 + 8] = b2;
    a[k1 + 12] = b3;
  }
}

template <int R, bool INV>
__device__ __forceinline__ void dft(float2* a) {
  if constexpr (R == 16) {
    dft16<INV>(a);
  } else if constexpr (R == 2) {
    dft2<INV>(a[0], a[1]);
  } else if constexpr (R == 4) {
    dft4<INV>(a[0], a[1], a[2], a[3]);
  } else if constexpr (R == 8) {
    dft8<INV>(a);
  } else {
    static_assert(R == 2 || R == 4 || R == 8 || R == 16, "unsupported radix");
  }
}

// ---- compile-time stage schedule -------------------------------------------
// Stage radix: at most RM (in-register DFTs up to radix 16; RM = 8 keeps the
// butterfly temporaries in fewer registers); a thread owning E > RM values
// performs E/R butterflies per stage.
constexpr int kMaxRadix = 16;
template <int N, int E, int NS, int RM = kMaxRadix>
constexpr int stage_radix() {
  constexpr int R0 = E < RM ? E : RM;
  return (N / NS) >= R0 ? R0 : (N / NS);
}
// number of base twiddles a thread holds for stages starting at span NS
template <int N, int E, int NS = 1, int RM = kMaxRadix>
constexpr int tw_count() {
  if constexpr (NS >= N) {
    return 0;
  } else {
    constexpr int R = stage_radix<N, E, NS, RM>();
    return (NS > 1 ? E / R : 0) + tw_count<N, E, NS * R, RM>();
  }
}

template <int N, int E, int RM = kMaxRadix>
struct FftTw {
  static constexpr int K = tw_count<N, E, 1, RM>() > 0 ? tw_count<N, E, 1, RM>() : 1;
  float2 w[K];
};

template <int N, int E, int NS, int OFF, int RM>
__device__ __forceinline__ void load_tw_rec(float2* w, int t, const float2* __restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int R = stage_radix<N, E, NS, RM>();
    constexpr int T = N / E;
    if constexpr (NS > 1) {
#pragma unroll
      for (int b = 0; b < E / R; ++b) {
        const int k = (t + b * T) & (NS - 1);
        w[OFF + b] = tw[k * (N / (NS * R))];
      }
    }
    load_tw_rec<N, E, NS * R, OFF + (NS > 1 ? E / R : 0), RM>(w, t, tw);
  }
}

// Load this thread's base twiddles (t in [0, N/E)).  Issue early: the loads
// overlap with whatever the kernel does before the first transform.
template <int N, int E, int RM = kMaxRadix>
__device__ __forceinline__ void fft_load_twiddles(FftTw<N, E, RM>& r, int t, const float2* __restrict__ tw) {
  load_tw_rec<N, E, 1, 0, RM>(r.w, t, tw);
}

// Ordering of LDS accesses between the phases of a transform.  WAVE: every
// thread that touches the buffer belongs to one wavefront, whose LDS
// instructions complete in issue order, so keeping the compiler from moving
// accesses across the point is enough; otherwise a workgroup barrier.
template <bool WAVE>
__device__ __forceinline__ void lds_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// powers p[r] = w^r, r = 1..R-1 (p[0] unused), products of depth <= 4
template <int R, bool INV>
__device__ __forceinline__ void tw_powers(float2 w, float2* p) {
  if (INV) w.y = -w.y;
  p[1] = w;
  if constexpr (R > 2) {
    p[2] = cmul(w, w);
    p[3] = cmul(p[2], w);
  }
  if constexpr (R > 4) {
    p[4] = cmul(p[2], p[2]);
    p[5] = cmul(p[4], w);
    p[6] = cmul(p[3], p[3]);
    p[7] = cmul(p[4], p[3]);
  }
  if constexpr (R > 8) {
    p[8] = cmul(p[4], p[4]);
#pragma unroll
    for (int r = 1; r < 8; ++r) p[8 + r] = cmul(p[8], p[r]);
  }
}

// Output hook of a transform's last stage.  There a thread's outputs are the
// elements t + e N/E, e = b + r E/R for butterfly b and output r, so a
// pointwise step on the transform's result can run in registers as the stage
// writes it: epi(e, v) returns the value stored for element t + e N/E.  A
// hook whose `store` is false consumes the results itself and the stage
// writes nothing.  `sync_after` false drops the stage's closing ordering
// point, for callers that then read back only their own outputs or leave the
// buffer alone until an ordering point of their own.
struct NoEpi {
  static constexpr bool active = false, store = true, sync_after = true;
  __device__ __forceinline__ float2 operator()(int, float2 v) const { return v; }
};

// One Stockham stage: span NS (product of radices done so far), radix R,
// reading src and writing dst.  In place (src == dst) the reads of every
// thread must finish before any write (MID: a second ordering point); with two
// buffers the one ordering point after the writes is enough.  EPI: the output
// hook, last stage only.
template <int N, int E, int R, int NS, bool INV, bool WAVE = false, bool MID = true, typename EPI = NoEpi>
__device__ __forceinline__ void stockham_stage(const float2* src, float2* dst, int t, const float2* w,
                                               const EPI& epi = EPI()) {
  constexpr int T = N / E;
  constexpr int NB = E / R;
  float2 a[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    if constexpr ((N / R) % 16 == 0) {
      const int pj = fpad(j);  // fpad(j + 16 m) = fpad(j) + 17 m
#pragma unroll
      for (int r = 0; r < R; ++r) a[b][r] = src[pj + r * (N / R / 16) * 17];
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) a[b][r] = src[fpad(j + r * (N / R))];
    }
  }
  if constexpr (MID && EPI::store) lds_sync<WAVE>();  // a stage that writes nothing reads only
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int j = t + b * T;
    const int k = j & (NS - 1);
    if constexpr (NS > 1) {
      // butterflies b of one thread share k (and the twiddles) when T % NS == 0
      constexpr bool kShared = (T % NS) == 0;
      float2 p[R];
      tw_powers<R, INV>(w[kShared ? 0 : b], p);
#pragma unroll
      for (int r = 1; r < R; ++r) a[b][r] = cmul(a[b][r], p[r]);
    }
    dft<R, INV>(a[b]);
    if constexpr (EPI::active) {
      static_assert(NS * R == N, "the output hook runs on the last stage");
#pragma unroll
      for (int r = 0; r < R; ++r) a[b][r] = epi(b + r * NB, a[b][r]);
    }
    if constexpr (EPI::store) {
      const int base = (j - k) * R + k;
      if constexpr (NS % 16 == 0) {
        const int pb = fpad(base);
#pragma unroll
        for (int r = 0; r < R; ++r) dst[pb + r * (NS / 16) * 17] = a[b][r];
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) dst[fpad(base + r * NS)] = a[b][r];
      }
    }
  }
  if constexpr (EPI::sync_after) lds_sync<WAVE>();
}

// In-place stages from span NS; the last stage applies the output hook epi.
template <int N, int E, int NS, int OFF, bool INV, int RM, bool WAVE, typename EPI = NoEpi>
__device__ __forceinline__ void stockham(float2* s, int t, const float2* w, const EPI& epi = EPI()) {
  if constexpr (NS < N) {
    constexpr int R = stage_radix<N, E, NS, RM>();
    if constexpr (NS * R >= N) {
      stockham_stage<N, E, R, NS, INV, WAVE, true, EPI>(s, s, t, w + OFF, epi);
    } else {
      stockham_stage<N, E, R, NS, INV, WAVE>(s, s, t, w + OFF);
      stockham<N, E, NS * R, OFF + (NS > 1 ? E / R : 0), INV, RM, WAVE, EPI>(s, t, w, epi);
    }
  }
}

// Ping-pong stages: stage i reads buffer i % 2 of (a, b) and writes the other;
// the last stage applies the output hook epi.
template <int N, int E, int NS, int OFF, bool INV, int RM, bool WAVE, bool FLIP, typename EPI = NoEpi>
__device__ __forceinline__ void stockham_pp(float2* a, float2* b, int t, const float2* w, const EPI& epi = EPI()) {
  if constexpr (NS < N) {
    constexpr int R = stage_radix<N, E, NS, RM>();
    if constexpr (NS * R >= N) {
      stockham_stage<N, E, R, NS, INV, WAVE, false, EPI>(FLIP ? b : a, FLIP ? a : b, t, w + OFF, epi);
    } else {
      stockham_stage<N, E, R, NS, INV, WAVE, false>(FLIP ? b : a, FLIP ? a : b, t, w + OFF);
      stockham_pp<N, E, NS * R, OFF + (NS > 1 ? E / R : 0), INV, RM, WAVE, !FLIP, EPI>(a, b, t, w, epi);
    }
  }
}

// First stage (span 1) of a transform whose inputs the thread holds in
// registers: h[e] is element t + e N/E, the elements the stage would read
// (butterfly b, input r: e = b + r E/R).  Writes dst, then one ordering point;
// nothing is read from dst, so in place needs no second one.
template <int N, int E, bool INV, int RM, bool WAVE>
__device__ __forceinline__ void stockham_first_from(const float2 (&h)[E], float2* dst, int t) {
  constexpr int R = stage_radix<N, E, 1, RM>();
  constexpr int T = N / E, NB = E / R;
  float2 a[NB][R];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int r = 0; r < R; ++r) a[b][r] = h[b + r * NB];
    dft<R, INV>(a[b]);
    const int j = t + b * T;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[fpad(j * R + r)] = a[b][r];
  }
  lds_sync<WAVE>();
}

// Number of Stockham stages of a length-N transform.
template <int N, int E, int RM = kMaxRadix, int NS = 1>
constexpr int fft_stages() {
  if constexpr (NS >= N)
    return 0;
  else
    return 1 + fft_stages<N, E, RM, NS * stage_radix<N, E, NS, RM>()>();
}

// Transform the buffer s (fft_buf<N>() slots, element i at s[fpad(i)]) in
// place; every thread of the workgroup must call this (it contains barriers;
// with WAVE, every thread of the wavefront that owns the buffer, which must
// be all N/E threads of the transform).  t in [0, N/E) is the thread's index
// within this FFT and tw its registers from fft_load_twiddles with the same t.
template <int N, int E, bool INV, int RM = kMaxRadix, bool WAVE = false, typename EPI = NoEpi>
__device__ __forceinline__ void fft_lds(float2* s, int t, const FftTw<N, E, RM>& tw, const EPI& epi = EPI()) {
  static_assert((N & (N - 1)) == 0 && N >= 2, "N must be a power of two");
  static_assert(E <= N && (E == 2 || E == 4 || E == 8 || E == 16 || E == 32), "E must be 2..32");
  static_assert(!WAVE || N / E == 64, "a wavefront-private transform takes exactly 64 threads");
  stockham<N, E, 1, 0, INV, RM, WAVE, EPI>(s, t, tw.w, epi);
}

// The same transform of a with b as the second buffer: one ordering point per
// stage instead of two.  The result is in b when fft_stages is odd, else in a.
// epi: output hook of the last stage (see NoEpi).
template <int N, int E, bool INV, int RM = kMaxRadix, bool WAVE = false, typename EPI = NoEpi>
__device__ __forceinline__ void fft_lds_pp(float2* a, float2* b, int t, const FftTw<N, E, RM>& tw,
                                           const EPI& epi = EPI()) {
  static_assert((N & (N - 1)) == 0 && N >= 2, "N must be a power of two");
  static_assert(E <= N && (E == 2 || E == 4 || E == 8 || E == 16 || E == 32), "E must be 2..32");
  static_assert(!WAVE || N / E == 64, "a wavefront-private transform takes exactly 64 threads");
  stockham_pp<N, E, 1, 0, INV, RM, WAVE, false, EPI>(a, b, t, tw.w, epi);
}

// Makhoul reordering that turns a length-N DCT-II into a length-N FFT:
// natural index i -> FFT input slot (even i -> i/2, odd i -> N-1-i/2).
template <int N>
__device__ __forceinline__ int dct_slot(int i) {
  return (i & 1) ? (N - 1 - (i >> 1)) : (i >> 1);
}

// Natural index of FFT slot p under that reordering (inverse of dct_slot).
template <int N>
__device__ __forceinline__ int dct_nat(int p) {
  return p < N / 2 ? 2 * p : 2 * (N - 1 - p) + 1;
}

}  // namespace htamd
