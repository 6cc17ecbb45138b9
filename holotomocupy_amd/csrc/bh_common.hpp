// Pieces shared by the fused BH passes (gradient and bilinear form): value
// conversions between the object type and complex, the staged-row LDS span,
// the plan that binds a propagator and a shift plan, and the stage-timing hook.
#pragma once
#include "../../include/htamd.h"
#include "fresnel_core.hpp"
#include "htamd_bspline.hpp"
#include "htamd_internal.hpp"

namespace htamd {

template <typename TO> __device__ __forceinline__ float2 to_c(TO v);
template <> __device__ __forceinline__ float2 to_c<float>(float v) { return make_float2(v, 0.f); }
template <> __device__ __forceinline__ float2 to_c<float2>(float2 v) { return v; }
template <typename TO> __device__ __forceinline__ TO from_c(float2 v);
template <> __device__ __forceinline__ float from_c<float>(float2 v) { return v.x; }
template <> __device__ __forceinline__ float2 from_c<float2>(float2 v) { return v; }

// Row-staged gathers.  All pixels of a detector row sample the same four
// object rows (iy .. iy+3), so those rows are combined along y once per
// (angle, row) into LDS with coalesced reads of the coefficients, and each
// pixel then reads 4 LDS values instead of 16 scattered global ones.
__device__ __forceinline__ unsigned char* dyn_lds() {
  extern __shared__ __align__(16) unsigned char htamd_dyn_lds[];
  return htamd_dyn_lds;
}

template <int N>
__device__ __forceinline__ void row_span(int npsi, float mb, float rx, int& lo, int& hi) {
  // the sampling position is monotone in the pixel index, so the row's taps
  // lie within [floor(x(0)) - 1, floor(x(N-1)) + 2] (the floors of bs_pos, which
  // every gather uses)
  const int i0 = bs_pos(0, N, npsi, mb, rx).i, i1 = bs_pos(N - 1, N, npsi, mb, rx).i;
  lo = max(min(i0, i1) - 1, 0);
  hi = min(max(i0, i1) + 3, npsi);
}

// Optional timing hook: events[2k] / events[2k+1] are recorded before / after
// stage k (0 prefilter, 1 rows forward, 2 columns, 3 rows adjoint, 4 Sadj,
// 5 prefilter of the gradient, 6 reductions) on the launch stream.
struct StageEvents {
  hipEvent_t* ev = nullptr;
  int count = 0;
  void mark(int idx, hipStream_t st) const {
    if (ev && idx < count && ev[idx]) (void)hipEventRecord(ev[idx], st);
  }
};

inline int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

}  // namespace htamd

struct htamd_bh_plan {
  // binds the plans of one geometry; the fused passes take their tables from them
  const htamd_prop_plan* prop = nullptr;
  const htamd_shift_plan* shift = nullptr;
  htamd::StageEvents events;
};

