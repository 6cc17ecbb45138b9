// LDS-resident mixed-radix Stockham FFT of any length N (run-time length).
//
// The power-of-two engine (fft_lds.hpp) fixes N, the radices and the
// per-thread work at compile time.  Detector and object sizes of real scans
// are not powers of two (1536 = 2^9 3, 2500 = 2^2 5^4, 1608 = 2^3 3 67, the
// oversampled Radon grids 2n), so this engine takes the factorisation at run
// time: stages of radix 8, 4, 2, 3 and 5 use in-register butterflies, and any
// other prime factor p runs as a direct length-p DFT stage (p complex
// products per output, p^2 / p per point: cheap for the small primes of
// real sizes, and correct for any).
//
// Stage s (radix R, span NS = product of the radices before it) maps
//   out[(j - k) R + k + q NS] = sum_r in[j + r N/R] exp(-2 pi i r (k + q NS) / (NS R)),
//   j < N/R, k = j mod NS, q < R,
// i.e. the Stockham autosort form of the power-of-two engine; the result is in
// natural order.  Two LDS buffers of N slots alternate between stages (read
// one, write the other, one workgroup barrier per stage), so a thread may own
// any number of butterflies: a line of length N is carried by any number of
// threads.  Twiddles come from a table tw[m] = exp(-2 pi i m / N) built in
// FP64 on the host; the inverse uses the conjugates.  Unnormalised.
//
// The stage bodies are __host__ __device__ so that the index arithmetic can
// be run on the CPU against a direct DFT.
#pragma once
#include "fft_lds.hpp"

namespace htamd {

constexpr int kAnyMaxStages = 24;

struct AnyFft {
  int N = 0;
  int nst = 0;
  unsigned char radix[kAnyMaxStages] = {};
  const float2* tw = nullptr;  // [N] exp(-2 pi i m / N)
};

// Radix order of a length: 8s, then one 4 or 2, then 3s, 5s, then the
// remaining primes in increasing order.  Returns false when N < 1 or the
// length needs more than kAnyMaxStages stages.
inline bool any_fft_factor(int64_t N, AnyFft* f) {
  if (N < 1 || N > (1 << 24)) return false;
  f->N = (int)N;
  f->nst = 0;
  int64_t m = N;
  auto push = [&](int r) {
    if (f->nst >= kAnyMaxStages) return false;
    f->radix[f->nst++] = (unsigned char)r;
    m /= r;
    return true;
  };
  while (m % 8 == 0)
    if (!push(8)) return false;
  if (m % 4 == 0 && !push(4)) return false;
  if (m % 2 == 0 && !push(2)) return false;
  while (m % 3 == 0)
    if (!push(3)) return false;
  while (m % 5 == 0)
    if (!push(5)) return false;
  for (int64_t p = 7; m > 1; p += 2) {
    if (p * p > m) p = m;  // what remains is prime
    while (m % p == 0) {
      if (p > 255) return false;  // a radix is stored in one byte
      if (!push((int)p)) return false;
    }
  }
  return true;
}

template <bool INV>
__host__ __device__ __forceinline__ float2 tw_dir(float2 w) {
  return INV ? make_float2(w.x, -w.y) : w;
}

template <bool INV>
__host__ __device__ __forceinline__ void dft3(float2* a) {
  constexpr float s3 = 0.86602540378443864676f;  // sin(2 pi / 3)
  const float2 t1 = cadd(a[1], a[2]);
  const float2 t2 = make_float2(a[0].x - 0.5f * t1.x, a[0].y - 0.5f * t1.y);
  // forward: X1 = t2 - i s3 (a1 - a2), X2 = t2 + i s3 (a1 - a2)
  const float2 t3 = mul_mi<INV>(cscale(csub(a[1], a[2]), s3));
  a[0] = cadd(a[0], t1);
  a[1] = cadd(t2, t3);
  a[2] = csub(t2, t3);
}

template <bool INV>
__host__ __device__ __forceinline__ void dft5(float2* a) {
  constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
  constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
  constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
  constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
  const float2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
  const float2 t3 = csub(a[1], a[4]), t4 = csub(a[2], a[3]);
  const float2 b1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
  const float2 b2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
  // forward: X1 = b1 - i d1, X4 = b1 + i d1, X2 = b2 - i d2, X3 = b2 + i d2
  const float2 d1 = mul_mi<INV>(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
  const float2 d2 = mul_mi<INV>(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
  a[0] = cadd(a[0], cadd(t1, t2));
  a[1] = cadd(b1, d1);
  a[4] = csub(b1, d1);
  a[2] = cadd(b2, d2);
  a[3] = csub(b2, d2);
}

template <int R, bool INV>
__host__ __device__ __forceinline__ void any_dft(float2* a) {
  if constexpr (R == 2) {
    dft2<INV>(a[0], a[1]);
  } else if constexpr (R == 3) {
    dft3<INV>(a);
  } else if constexpr (R == 4) {
    dft4<INV>(a[0], a[1], a[2], a[3]);
  } else if constexpr (R == 5) {
    dft5<INV>(a);
  } else {
    static_assert(R == 8, "in-register radices are 2, 3, 4, 5 and 8");
    dft8<INV>(a);
  }
}

// One stage with an in-register radix-R butterfly; this thread's butterflies
// are j = tid, tid + nthr, ...
template <int R, bool INV>
__host__ __device__ __forceinline__ void any_stage_r(const float2* src, float2* dst, int N, int NS,
                                                     const float2* __restrict__ tw, int tid, int nthr) {
  const int M = N / R, ts = N / (NS * R);
  for (int j = tid; j < M; j += nthr) {
    const int k = j % NS;
    float2 a[R];
#pragma unroll
    for (int r = 0; r < R; ++r) a[r] = src[j + r * M];
    if (NS > 1) {
      const int step = k * ts;  // r k ts < N
#pragma unroll
      for (int r = 1; r < R; ++r) a[r] = cmul(a[r], tw_dir<INV>(tw[r * step]));
    }
    any_dft<R, INV>(a);
    const int base = (j - k) * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[base + r * NS] = a[r];
  }
}

// One stage of any prime radix R as a direct DFT: one output per work item.
template <bool INV>
__host__ __device__ __forceinline__ void any_stage_dft(const float2* src, float2* dst, int N, int R, int NS,
                                                       const float2* __restrict__ tw, int tid, int nthr) {
  const int M = N / R, ts = N / (NS * R);
  for (int o = tid; o < N; o += nthr) {
    const int j = o % M, q = o / M;
    const int k = j % NS;
    const int step = (k + q * NS) * ts;  // < N
    float2 acc = src[j];
    int m = 0;
    for (int r = 1; r < R; ++r) {
      m += step;
      if (m >= N) m -= N;
      const float2 w = tw_dir<INV>(tw[m]);
      const float2 v = src[j + r * M];
      acc.x += v.x * w.x - v.y * w.y;
      acc.y += v.x * w.y + v.y * w.x;
    }
    dst[(j - k) * R + k + q * NS] = acc;
  }
}

template <bool INV>
__host__ __device__ __forceinline__ void any_stage(const float2* src, float2* dst, int N, int R, int NS,
                                                   const float2* __restrict__ tw, int tid, int nthr) {
  switch (R) {
    case 2: any_stage_r<2, INV>(src, dst, N, NS, tw, tid, nthr); break;
    case 3: any_stage_r<3, INV>(src, dst, N, NS, tw, tid, nthr); break;
    case 4: any_stage_r<4, INV>(src, dst, N, NS, tw, tid, nthr); break;
    case 5: any_stage_r<5, INV>(src, dst, N, NS, tw, tid, nthr); break;
    case 8: any_stage_r<8, INV>(src, dst, N, NS, tw, tid, nthr); break;
    default: any_stage_dft<INV>(src, dst, N, R, NS, tw, tid, nthr); break;
  }
}

// Transform the line held in a[0..N) using b[0..N) as the second buffer;
// returns the buffer that holds the result (a or b).  Every thread of the
// workgroup calls this with the same stage count (it holds one barrier per
// stage); tid in [0, nthr) is the thread's index among the line's nthr
// threads.  The caller's writes to a must be behind a barrier.
template <bool INV>
__device__ __forceinline__ float2* fft_any(float2* a, float2* b, const AnyFft& f, int tid, int nthr) {
  int NS = 1;
  for (int s = 0; s < f.nst; ++s) {
    const int R = f.radix[s];
    any_stage<INV>(a, b, f.N, R, NS, f.tw, tid, nthr);
    __syncthreads();
    NS *= R;
    float2* t = a;
    a = b;
    b = t;
  }
  return a;
}

// ---- generic line kernels ------------------------------------------------------
// An operator Op supplies what happens around the transform(s):
//   static constexpr bool kInv   direction of the first transform
//   static constexpr bool kTwo   a second transform (direction !kInv) follows
//   float2 load(int64_t unit, int c, int p)    input for slot p of line c of the unit
//   void mid(float2* s, int64_t unit, int c, int tid, int nthr)   between the two
//                                               transforms (kTwo), on line c's result
//   void store(int64_t unit, int c, int p, float2 v)               output of slot p
// Rows: one unit per line, lines [blockIdx.x * lpb, ...) with tpl threads each
// (c = 0).  Columns: unit blockIdx.x holds C lines (c < C) that the loads and
// stores visit with consecutive threads on consecutive c, so that C adjacent
// columns of a row are one contiguous access.  LDS: 2N slots per line.

template <class Op>
__device__ __forceinline__ float2* any_transforms(const Op& op, float2* a, const AnyFft& f, int64_t unit, int c,
                                                  int tid, int nthr) {
  float2* res = fft_any<Op::kInv>(a, a + f.N, f, tid, nthr);
  if constexpr (Op::kTwo) {
    op.mid(res, unit, c, tid, nthr);
    __syncthreads();
    res = fft_any<!Op::kInv>(res, res == a ? a + f.N : a, f, tid, nthr);
  }
  return res;
}

template <class Op>
__global__ __launch_bounds__(256) void any_rows_k(Op op, AnyFft f, int64_t nlines, int tpl) {
  extern __shared__ float2 any_lds[];
  const int lpb = (int)blockDim.x / tpl;
  const int r = (int)threadIdx.x / tpl, t = (int)threadIdx.x % tpl;
  const int64_t line = (int64_t)blockIdx.x * lpb + r;
  const bool valid = line < nlines;
  float2* a = any_lds + (size_t)r * 2 * f.N;
  for (int p = t; p < f.N; p += tpl) a[p] = valid ? op.load(line, 0, p) : make_float2(0.f, 0.f);
  __syncthreads();
  float2* res = any_transforms(op, a, f, valid ? line : 0, 0, t, tpl);
  if (valid)
    for (int p = t; p < f.N; p += tpl) op.store(line, 0, p, res[p]);
}

template <class Op>
__global__ __launch_bounds__(256) void any_cols_k(Op op, AnyFft f, int C) {
  extern __shared__ float2 any_lds[];
  const int64_t unit = blockIdx.x;
  const int tpc = (int)blockDim.x / C;
  const int c = (int)threadIdx.x / tpc, t = (int)threadIdx.x % tpc;
  const int tot = C * f.N;
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int cc = idx % C, p = idx / C;
    any_lds[(size_t)cc * 2 * f.N + p] = op.load(unit, cc, p);
  }
  __syncthreads();
  float2* a = any_lds + (size_t)c * 2 * f.N;
  const int off = (int)(any_transforms(op, a, f, unit, c, t, tpc) - a);  // the same for every line
  for (int idx = threadIdx.x; idx < tot; idx += blockDim.x) {
    const int cc = idx % C, p = idx / C;
    op.store(unit, cc, p, any_lds[(size_t)cc * 2 * f.N + off + p]);
  }
}

// Largest dynamic LDS a workgroup may use on the current device (bytes).
int64_t any_max_lds();

// Threads per line of the row kernel and lines per 256-thread workgroup.
inline int any_row_tpl(int N) { return N <= 512 ? 64 : N <= 1024 ? 128 : 256; }

template <class Op>
hipError_t any_launch_rows(const Op& op, const AnyFft& f, int64_t nlines, hipStream_t st) {
  if (nlines <= 0) return hipSuccess;
  int tpl = any_row_tpl(f.N);
  const int64_t lds_line = (int64_t)2 * f.N * sizeof(float2);
  const int64_t cap = any_max_lds();
  if (lds_line > cap) return hipErrorInvalidValue;
  while (tpl < 256 && (256 / tpl) * lds_line > cap) tpl *= 2;
  const int lpb = 256 / tpl;
  const size_t lds = (size_t)(lpb * lds_line);
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)any_rows_k<Op>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(any_rows_k<Op>, dim3((unsigned)((nlines + lpb - 1) / lpb)), dim3(256), lds, st, op, f, nlines,
                     tpl);
  return hipGetLastError();
}

// Columns per workgroup: the largest of 16, 8, 4, 2, 1 whose lines fit the LDS.
inline int any_cols_c(int N, int64_t cap) {
  int C = 16;
  while (C > 1 && (int64_t)C * 2 * N * (int64_t)sizeof(float2) > cap) C /= 2;
  return C;
}

template <class Op>
hipError_t any_launch_cols(const Op& op, const AnyFft& f, int C, int64_t units, hipStream_t st) {
  if (units <= 0) return hipSuccess;
  const size_t lds = (size_t)C * 2 * f.N * sizeof(float2);
  if ((int64_t)lds > any_max_lds()) return hipErrorInvalidValue;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute((const void*)any_cols_k<Op>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)lds);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(any_cols_k<Op>, dim3((unsigned)units), dim3(256), lds, st, op, f, C);
  return hipGetLastError();
}

// Natural index of slot p of a Makhoul-ordered line of even length N (the
// run-time form of dct_nat).
__device__ __forceinline__ int dct_nat_rt(int p, int N) { return p < N / 2 ? 2 * p : 2 * (N - 1 - p) + 1; }

// Pair map of the 1-D Fresnel operator on a line of even length N (the
// run-time form of dct_pair_map; same table layout).
__device__ __forceinline__ void dct_pair_map_rt(float2* s, int N, const float2* __restrict__ tab, int tid,
                                                int nthr) {
  const int H = N / 2;
  for (int k = tid; k < H; k += nthr) {
    if (k == 0) {
      s[0] = cmul(tab[0], s[0]);
      s[H] = cmul(tab[H], s[H]);
    } else {
      const float2 v1 = s[k], v2 = s[N - k];
      const float2 a = tab[k], b = tab[H + k], c = tab[2 * H + k];
      s[k] = cadd(cmul(a, v1), cmul(b, v2));
      s[N - k] = cadd(cmul(c, v1), cmul(a, v2));
    }
  }
}

}  // namespace htamd
