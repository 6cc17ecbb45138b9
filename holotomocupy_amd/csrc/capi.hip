// C ABI of libhtamd.so: plans own FP64-generated constant tables in device
// memory; callers own every data buffer and pass a HIP stream.  No exception
// crosses the ABI; every entry point returns 0 on success and a non-zero code
// otherwise, with a message readable through htamd_last_error().
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/htamd.h"
#include "htamd_internal.hpp"

namespace htamd {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int hip_fail(hipError_t e, const char* where) {
  return fail(HTAMD_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

static std::vector<float2> make_twiddles(int64_t N) {
  std::vector<float2> tw((size_t)N);
  for (int64_t m = 0; m < N; ++m) {
    const double a = -2.0 * M_PI * (double)m / (double)N;
    tw[(size_t)m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  return tw;
}

hipError_t upload_twiddles(float2** dst, int64_t N) {
  const std::vector<float2> tw = make_twiddles(N);
  hipError_t e = hipMalloc((void**)dst, tw.size() * sizeof(float2));
  if (e == hipSuccess) e = hipMemcpy(*dst, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice);
  return e;
}

// Pair table of the 1-D operator C^-1 diag(l) C of length N (see fresnel.hip),
// l_k = exp(sign * i * coef * k^2), coef = pi*lambda*z/(4 N^2 vs^2), FP64.
static void fill_pair_table(float2* tab, int64_t N, double coef, double sign) {
  using cd = std::complex<double>;
  const int64_t H = N / 2;
  auto lam = [&](int64_t k) {
    const double ph = sign * coef * (double)k * (double)k;
    return cd(std::cos(ph), std::sin(ph));
  };
  const double inv = 1.0 / (double)N;
  auto put = [&](int64_t idx, cd v) { tab[idx] = make_float2((float)v.real(), (float)v.imag()); };
  put(0, lam(0) * inv);
  put(H, lam(H) * inv);
  put(2 * H, cd(0.0, 0.0));
  for (int64_t k = 1; k < H; ++k) {
    const cd lk = lam(k), lnk = lam(N - k);
    const cd a = 0.5 * (lk + lnk) * inv;
    const cd b = 0.5 * (lk - lnk) * inv;
    const double th = M_PI * (double)k / (double)N;
    put(k, a);
    put(H + k, b * cd(std::cos(th), std::sin(th)));
    put(2 * H + k, b * cd(std::cos(th), -std::sin(th)));
  }
}

template <typename T>
static hipError_t upload(T** dst, const std::vector<T>& src) {
  hipError_t e = hipMalloc((void**)dst, src.size() * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

}  // namespace htamd

using namespace htamd;



extern "C" {

const char* htamd_last_error(void) { return g_last_error.c_str(); }

int htamd_abi_version(void) { return HTAMD_ABI_VERSION; }

int htamd_prop_plan_create(int64_t n, int64_t nz, int64_t ndist, double wavelength,
                           double voxelsize, const double* distance, htamd_prop_plan** plan) {
  if (!plan) return fail(HTAMD_ERR_ARG, "prop_plan_create: plan is NULL");
  *plan = nullptr;
  if (!(fft_size_supported(n) || fresnel_any_supported(n)) || !(fft_size_supported(nz) || fresnel_any_supported(nz)))
    return fail(HTAMD_ERR_UNSUPPORTED,
                "prop_plan_create: n and nz must be even sizes in [8, 8192] (powers of two up to 2048 on the "
                "fastest kernels), got n=" + std::to_string(n) + " nz=" + std::to_string(nz));
  if (ndist < 1 || !distance) return fail(HTAMD_ERR_ARG, "prop_plan_create: need ndist >= 1 distances");
  if (!(voxelsize > 0.0) || !(wavelength > 0.0))
    return fail(HTAMD_ERR_ARG, "prop_plan_create: wavelength and voxelsize must be > 0");

  auto* p = new htamd_prop_plan();
  p->n = n;
  p->nz = nz;
  p->ndist = ndist;
  hipError_t e = hipGetDevice(&p->device);
  if (e != hipSuccess) {
    delete p;
    return hip_fail(e, "prop_plan_create: hipGetDevice");
  }
  const int64_t hx = 3 * (n / 2), hy = 3 * (nz / 2);
  std::vector<float2> tx((size_t)(ndist * 2 * hx)), ty((size_t)(ndist * 2 * hy));
  for (int64_t j = 0; j < ndist; ++j) {
    const double lz = wavelength * distance[j];
    const double cx = M_PI * lz / (4.0 * (double)n * (double)n * voxelsize * voxelsize);
    const double cy = M_PI * lz / (4.0 * (double)nz * (double)nz * voxelsize * voxelsize);
    for (int adj = 0; adj < 2; ++adj) {
      const double sign = adj ? 1.0 : -1.0;
      fill_pair_table(tx.data() + (j * 2 + adj) * hx, n, cx, sign);
      fill_pair_table(ty.data() + (j * 2 + adj) * hy, nz, cy, sign);
    }
  }
  if ((e = upload(&p->tw_x, make_twiddles(n))) != hipSuccess ||
      (e = upload(&p->tw_y, make_twiddles(nz))) != hipSuccess ||
      (e = upload(&p->tab_x, tx)) != hipSuccess || (e = upload(&p->tab_y, ty)) != hipSuccess) {
    htamd_prop_plan_destroy(p);
    return hip_fail(e, "prop_plan_create: table upload");
  }
  *plan = p;
  return HTAMD_OK;
}

int htamd_prop_plan_destroy(htamd_prop_plan* p) {
  if (!p) return HTAMD_OK;
  hipFree(p->tw_x);
  hipFree(p->tw_y);
  hipFree(p->tab_x);
  hipFree(p->tab_y);
  delete p;
  return HTAMD_OK;
}

int htamd_prop_apply_strided(const htamd_prop_plan* p, const void* in, int64_t in_stride, void* out,
                             int64_t out_stride, int64_t batch, int64_t j, int adjoint, void* stream) {
  if (!p) return fail(HTAMD_ERR_ARG, "prop_apply: plan is NULL");
  if (batch < 0) return fail(HTAMD_ERR_ARG, "prop_apply: negative batch");
  if (j < 0 || j >= p->ndist)
    return fail(HTAMD_ERR_ARG, "prop_apply: distance index " + std::to_string(j) + " out of range");
  if (batch == 0) return HTAMD_OK;
  if (!in || !out) return fail(HTAMD_ERR_ARG, "prop_apply: NULL buffer");
  const int64_t field = p->n * p->nz;
  if (in_stride < field || out_stride < field)
    return fail(HTAMD_ERR_ARG, "prop_apply: field strides must be >= nz*n");
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != p->device)
    return fail(HTAMD_ERR_ARG, "prop_apply: plan was created on device " + std::to_string(p->device) +
                                   " but device " + std::to_string(dev) + " is current");
  const int a = adjoint ? 1 : 0;
  const float2* tabx = p->tab_x + (j * 2 + a) * 3 * (p->n / 2);
  const float2* taby = p->tab_y + (j * 2 + a) * 3 * (p->nz / 2);
  hipStream_t st = (hipStream_t)stream;
  hipError_t e = launch_fresnel_rows(p->n, p->nz, (const float2*)in, in_stride, (float2*)out,
                                     out_stride, batch, p->tw_x, tabx, st);
  if (e != hipSuccess) return hip_fail(e, "prop_apply: row pass");
  e = launch_fresnel_cols(p->nz, p->n, (float2*)out, out_stride, batch, p->tw_y, taby, st);
  if (e != hipSuccess) return hip_fail(e, "prop_apply: column pass");
  return HTAMD_OK;
}

int htamd_prop_pass(const htamd_prop_plan* p, const void* in, int64_t in_stride, void* out,
                    int64_t out_stride, int64_t batch, int64_t j, int adjoint, int axis, void* stream) {
  if (!p) return fail(HTAMD_ERR_ARG, "prop_pass: plan is NULL");
  if (batch < 0 || j < 0 || j >= p->ndist || (axis != 0 && axis != 1))
    return fail(HTAMD_ERR_ARG, "prop_pass: bad batch, distance or axis");
  if (batch == 0) return HTAMD_OK;
  if (!in || !out) return fail(HTAMD_ERR_ARG, "prop_pass: NULL buffer");
  const int64_t field = p->n * p->nz;
  if (in_stride < field || out_stride < field)
    return fail(HTAMD_ERR_ARG, "prop_pass: field strides must be >= nz*n");
  if (axis == 1 && in != out) return fail(HTAMD_ERR_ARG, "prop_pass: the column pass is in place");
  const int a = adjoint ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (axis == 0)
    e = launch_fresnel_rows(p->n, p->nz, (const float2*)in, in_stride, (float2*)out, out_stride, batch,
                            p->tw_x, p->tab_x + (j * 2 + a) * 3 * (p->n / 2), st);
  else
    e = launch_fresnel_cols(p->nz, p->n, (float2*)out, out_stride, batch, p->tw_y,
                            p->tab_y + (j * 2 + a) * 3 * (p->nz / 2), st);
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "prop_pass");
}

int htamd_prop_apply(const htamd_prop_plan* p, const void* in, void* out, int64_t batch, int64_t j,
                     int adjoint, void* stream) {
  const int64_t field = p ? p->n * p->nz : 0;
  return htamd_prop_apply_strided(p, in, field, out, field, batch, j, adjoint, stream);
}

}  // extern "C"
