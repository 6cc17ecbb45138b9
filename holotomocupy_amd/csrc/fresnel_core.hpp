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
template <int N, int E>
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
  __syncthreads();
}

template <int N, int E>
__device__ __forceinline__ void fresnel_1d(float2* s, int t, const FftTw<N, E>& tw,
                                           const float2* __restrict__ tab) {
  fft_lds<N, E, false>(s, t, tw);
  dct_pair_map<N, E>(s, t, tab);
  fft_lds<N, E, true>(s, t, tw);
}

// Column-pass elements per thread: 32 at N = 2048 (64 threads per column, so a
// 4-column tile is one 256-thread workgroup), else 8.
template <int N>
constexpr int col_e() {
  return N >= 2048 ? N / 64 : 8;
}

constexpr int kRowThreads = 256;

}  // namespace htamd
