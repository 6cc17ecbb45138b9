// Common device helpers for the htamd gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace htamd {

constexpr int kNumXcd = 8;  // MI355X: 256 CUs in 8 XCDs, workgroups dealt round-robin

__host__ __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__host__ __device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__host__ __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__host__ __device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__host__ __device__ __forceinline__ float2 mul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// exp(i x) for real or complex x
__device__ __forceinline__ float2 expi(float2 x) {
  float s, c;
  sincosf(x.x, &s, &c);
  const float a = expf(-x.y);
  return make_float2(a * c, a * s);
}

// exp(i x) with the hardware transcendental unit (v_sin/v_cos/v_exp); for the
// object phases of the cascade (|x| of order 1-100 rad) the absolute error is
// ~1e-6, inside the 1e-5 operator tolerance.
__device__ __forceinline__ float2 expi_fast(float2 x) {
  float s, c;
  __sincosf(x.x, &s, &c);
  const float a = __expf(-x.y);
  return make_float2(a * c, a * s);
}

// Workgroup -> logical tile mapping that gives each XCD a contiguous range of
// tiles, so tiles that share cache lines are resident in the same L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t bid, int64_t nblocks) {
  const int64_t xcd = bid % kNumXcd;
  const int64_t local = bid / kNumXcd;
  const int64_t per = nblocks / kNumXcd;
  const int64_t rem = nblocks % kNumXcd;
  return xcd * per + (xcd < rem ? xcd : rem) + local;
}

}  // namespace htamd
