// Internal host-side interfaces between the C ABI (capi.hip) and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "htamd_bspline.hpp"

namespace htamd {

void set_error(const std::string& msg);
// set the message and return `code` / HTAMD_ERR_HIP
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* where);
// tw[m] = exp(-2 pi i m / N), m < N, generated in FP64, uploaded to *dst
hipError_t upload_twiddles(float2** dst, int64_t N);

// ---- Fresnel propagator (fresnel.hip) -------------------------------------
// Smallest / largest power-of-two transform length handled by the LDS FFT.
constexpr int64_t kFftMin = 8;
constexpr int64_t kFftMax = 2048;
bool fft_size_supported(int64_t n);
// Other even lengths in [kFftMin, kAnyFresnelMax] whose line fits the LDS run
// on the run-time mixed-radix FFT (fft_any.hpp), slower than the power-of-two
// kernels.
constexpr int64_t kAnyFresnelMax = 8192;
bool fresnel_any_supported(int64_t n);

// Row pass over `batch` fields of nz rows of length n; field b starts at
// in + b*in_stride / out + b*out_stride (elements); in may equal out.
// tab: pair table (3*(n/2) float2) of the 1-D operator, tw: twiddles (n float2).
hipError_t launch_fresnel_rows(int64_t n, int64_t nz, const float2* in, int64_t in_stride,
                               float2* out, int64_t out_stride, int64_t batch, const float2* tw,
                               const float2* tab, hipStream_t st);
// Column pass, in place: `batch` fields [nz, n] at data + b*stride.
hipError_t launch_fresnel_cols(int64_t nz, int64_t n, float2* data, int64_t stride, int64_t batch,
                               const float2* tw, const float2* tab, hipStream_t st);

}  // namespace htamd

// Plan objects behind the opaque handles of include/htamd.h (device tables of
// a propagator plan live on plan->device).
struct htamd_prop_plan {
  int64_t n = 0, nz = 0, ndist = 0;
  int device = -1;
  float2* tw_x = nullptr;
  float2* tw_y = nullptr;
  float2* tab_x = nullptr;  // [ndist][2][3*(n/2)]
  float2* tab_y = nullptr;  // [ndist][2][3*(nz/2)]
};

struct htamd_shift_plan {
  int64_t n = 0, npsi = 0, nz = 0, nzpsi = 0;
  int is_complex = 1;
  htamd::PrefilterTaps taps{};
  // FFT twiddles of the object grid (coeffback), allocated on the current
  // device when npsi and nzpsi are transform sizes of the LDS FFT, else NULL
  float2* tw_x = nullptr;
  float2* tw_y = nullptr;
};
