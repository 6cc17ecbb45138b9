// C ABI of the cubic B-spline shift family (see include/htamd.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/htamd.h"
#include "htamd_bspline.hpp"
#include "htamd_internal.hpp"

using namespace htamd;

namespace htamd {

// Periodic inverse of the [1 4 1] circulant of length N, i.e. the impulse
// response of the reference's ifB3 factor 1/(4 + 2 cos(2 pi k/N))
// (shift.py:25-30), sampled at offsets |d| <= kPfH; FP64.  Offsets that alias
// (N < 2 kPfH + 1) are kept once, nearest to zero first.
static void prefilter_taps(int64_t N, float* w) {
  const int H = kPfH;
  std::vector<char> used((size_t)N, 0);
  for (int i = 0; i <= 2 * H; ++i) w[i] = 0.0f;
  for (int a = 0; a <= H; ++a) {
    for (int sgn = 1; sgn >= -1; sgn -= 2) {
      if (a == 0 && sgn < 0) continue;
      const int d = sgn * a;
      const int64_t rho = ((d % N) + N) % N;
      if (used[(size_t)rho]) continue;
      used[(size_t)rho] = 1;
      double s = 0.0;
      for (int64_t k = 0; k < N; ++k) {
        const double th = 2.0 * M_PI * (double)k / (double)N;
        s += std::cos(th * (double)rho) / (4.0 + 2.0 * std::cos(th));
      }
      w[d + H] = (float)(s / (double)N);
    }
  }
}

template <typename T>
static BsArgs<T> base_args(const htamd_shift_plan* p, int64_t batch, const float* r, int64_t rs,
                           const float* m, int64_t ms) {
  BsArgs<T> a;
  a.n = (int)p->n;
  a.nz = (int)p->nz;
  a.npsi = (int)p->npsi;
  a.nzpsi = (int)p->nzpsi;
  a.batch = batch;
  a.r = r;
  a.rs = rs;
  a.m = m;
  a.ms = ms;
  return a;
}

static int check_common(const htamd_shift_plan* p, int64_t batch, const float* r, const float* m,
                        const char* what) {
  if (!p) return fail(HTAMD_ERR_ARG, std::string(what) + ": plan is NULL");
  if (batch < 0) return fail(HTAMD_ERR_ARG, std::string(what) + ": negative batch");
  if (batch > 0 && (!r || !m)) return fail(HTAMD_ERR_ARG, std::string(what) + ": NULL shift or magnification");
  return HTAMD_OK;
}

}  // namespace htamd

extern "C" {

int htamd_shift_plan_create(int64_t n, int64_t npsi, int64_t nz, int64_t nzpsi, int is_complex,
                            htamd_shift_plan** plan) {
  if (!plan) return fail(HTAMD_ERR_ARG, "shift_plan_create: plan is NULL");
  *plan = nullptr;
  if (n < 1 || npsi < 1 || nz < 1 || nzpsi < 1)
    return fail(HTAMD_ERR_ARG, "shift_plan_create: sizes must be positive");
  if (n > (1 << 20) || npsi > (1 << 20) || nz > (1 << 20) || nzpsi > (1 << 20))
    return fail(HTAMD_ERR_UNSUPPORTED, "shift_plan_create: sizes above 2^20 are not supported");
  auto* p = new htamd_shift_plan();
  p->n = n;
  p->npsi = npsi;
  p->nz = nz;
  p->nzpsi = nzpsi;
  p->is_complex = is_complex ? 1 : 0;
  prefilter_taps(npsi, p->taps.wx);
  prefilter_taps(nzpsi, p->taps.wy);
  if (fft_size_supported(npsi) && fft_size_supported(nzpsi)) {
    hipError_t e = upload_twiddles(&p->tw_x, npsi);
    if (e == hipSuccess) e = upload_twiddles(&p->tw_y, nzpsi);
    if (e != hipSuccess) {
      htamd_shift_plan_destroy(p);
      return hip_fail(e, "shift_plan_create");
    }
  }
  *plan = p;
  return HTAMD_OK;
}

int htamd_shift_plan_destroy(htamd_shift_plan* p) {
  if (!p) return HTAMD_OK;
  (void)hipFree(p->tw_x);
  (void)hipFree(p->tw_y);
  delete p;
  return HTAMD_OK;
}

int htamd_bspline_prefilter(const htamd_shift_plan* p, const void* in, void* out, int64_t batch,
                            void* stream) {
  if (!p) return fail(HTAMD_ERR_ARG, "bspline_prefilter: plan is NULL");
  if (batch < 0) return fail(HTAMD_ERR_ARG, "bspline_prefilter: negative batch");
  if ((p->npsi % 2) || (p->nzpsi % 2))
    return fail(HTAMD_ERR_UNSUPPORTED, "bspline_prefilter: npsi and nzpsi must be even");
  if (batch == 0) return HTAMD_OK;
  if (!in || !out || in == out)
    return fail(HTAMD_ERR_ARG, "bspline_prefilter: needs distinct non-NULL input and output");
  hipError_t e = p->is_complex
      ? bs_launch_prefilter<float2>((const float2*)in, (float2*)out, batch, (int)p->nzpsi,
                                    (int)p->npsi, p->taps, (hipStream_t)stream)
      : bs_launch_prefilter<float>((const float*)in, (float*)out, batch, (int)p->nzpsi, (int)p->npsi,
                                   p->taps, (hipStream_t)stream);
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "bspline_prefilter");
}

int htamd_bspline_interp(const htamd_shift_plan* p, const void* c, const float* r, int64_t r_stride,
                         const float* m, int64_t m_stride, void* out, int64_t batch, void* stream) {
  int s = check_common(p, batch, r, m, "bspline_interp");
  if (s || batch == 0) return s;
  if (!c || !out) return fail(HTAMD_ERR_ARG, "bspline_interp: NULL buffer");
  hipError_t e;
  if (p->is_complex) {
    auto a = base_args<float2>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float2*)c;
    a.out = (float2*)out;
    e = bs_launch_forward<float2>(a, 0, (hipStream_t)stream);
  } else {
    auto a = base_args<float>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float*)c;
    a.out = (float*)out;
    e = bs_launch_forward<float>(a, 0, (hipStream_t)stream);
  }
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "bspline_interp");
}

int htamd_bspline_interp_adj(const htamd_shift_plan* p, const void* g, const float* r,
                             int64_t r_stride, const float* m, int64_t m_stride, void* c_out,
                             int64_t batch, void* stream) {
  int s = check_common(p, batch, r, m, "bspline_interp_adj");
  if (s || batch == 0) return s;
  if (!g || !c_out) return fail(HTAMD_ERR_ARG, "bspline_interp_adj: NULL buffer");
  hipError_t e;
  if (p->is_complex) {
    auto a = base_args<float2>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float2*)g;
    a.out = (float2*)c_out;
    e = bs_launch_adjoint<float2>(a, (hipStream_t)stream);
  } else {
    auto a = base_args<float>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float*)g;
    a.out = (float*)c_out;
    e = bs_launch_adjoint<float>(a, (hipStream_t)stream);
  }
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "bspline_interp_adj");
}

int htamd_bspline_d1(const htamd_shift_plan* p, const void* c, const void* c1, const float* r,
                     int64_t r_stride, const float* m, int64_t m_stride, const float* dr,
                     int64_t dr_stride, void* out, int64_t batch, void* stream) {
  int s = check_common(p, batch, r, m, "bspline_d1");
  if (s || batch == 0) return s;
  if (!c || !c1 || !dr || !out) return fail(HTAMD_ERR_ARG, "bspline_d1: NULL buffer");
  hipError_t e;
  if (p->is_complex) {
    auto a = base_args<float2>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float2*)c;
    a.c1 = (const float2*)c1;
    a.dr1 = dr;
    a.d1s = dr_stride;
    a.out = (float2*)out;
    e = bs_launch_forward<float2>(a, 1, (hipStream_t)stream);
  } else {
    auto a = base_args<float>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float*)c;
    a.c1 = (const float*)c1;
    a.dr1 = dr;
    a.d1s = dr_stride;
    a.out = (float*)out;
    e = bs_launch_forward<float>(a, 1, (hipStream_t)stream);
  }
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "bspline_d1");
}

int htamd_bspline_d2(const htamd_shift_plan* p, const void* c, const void* c1, const void* c2,
                     const float* r, int64_t r_stride, const float* m, int64_t m_stride,
                     const float* dr1, int64_t dr1_stride, const float* dr2, int64_t dr2_stride,
                     void* out, int64_t batch, void* stream) {
  int s = check_common(p, batch, r, m, "bspline_d2");
  if (s || batch == 0) return s;
  if (!c || !c1 || !c2 || !dr1 || !dr2 || !out) return fail(HTAMD_ERR_ARG, "bspline_d2: NULL buffer");
  hipError_t e;
  if (p->is_complex) {
    auto a = base_args<float2>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float2*)c;
    a.c1 = (const float2*)c1;
    a.c2 = (const float2*)c2;
    a.dr1 = dr1;
    a.d1s = dr1_stride;
    a.dr2 = dr2;
    a.d2s = dr2_stride;
    a.out = (float2*)out;
    e = bs_launch_forward<float2>(a, 2, (hipStream_t)stream);
  } else {
    auto a = base_args<float>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float*)c;
    a.c1 = (const float*)c1;
    a.c2 = (const float*)c2;
    a.dr1 = dr1;
    a.d1s = dr1_stride;
    a.dr2 = dr2;
    a.d2s = dr2_stride;
    a.out = (float*)out;
    e = bs_launch_forward<float>(a, 2, (hipStream_t)stream);
  }
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "bspline_d2");
}

int htamd_bspline_d1_adj_workspace_size(const htamd_shift_plan* p, int64_t batch, int64_t* bytes) {
  if (!p || !bytes) return fail(HTAMD_ERR_ARG, "bspline_d1_adj_workspace_size: NULL argument");
  BsArgs<float> a;
  a.n = (int)p->n;
  a.nz = (int)p->nz;
  a.batch = batch < 0 ? 0 : batch;
  *bytes = bs_dadj_partials<float>(a) * (int64_t)sizeof(float2);
  return HTAMD_OK;
}

int htamd_bspline_d1_adj(const htamd_shift_plan* p, const void* c, const void* g, const float* r,
                         int64_t r_stride, const float* m, int64_t m_stride, void* out1, float* out2,
                         void* workspace, int64_t workspace_bytes, int64_t batch, void* stream) {
  int s = check_common(p, batch, r, m, "bspline_d1_adj");
  if (s || batch == 0) return s;
  if (!c || !g || !out1 || !out2) return fail(HTAMD_ERR_ARG, "bspline_d1_adj: NULL buffer");
  int64_t need = 0;
  htamd_bspline_d1_adj_workspace_size(p, batch, &need);
  if (!workspace || workspace_bytes < need)
    return fail(HTAMD_ERR_ARG, "bspline_d1_adj: workspace of " + std::to_string(need) +
                                   " bytes required, got " + std::to_string(workspace_bytes));
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (p->is_complex) {
    auto a = base_args<float2>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float2*)g;
    a.out = (float2*)out1;
    e = bs_launch_adjoint<float2>(a, st);
    if (e == hipSuccess) {
      auto b = base_args<float2>(p, batch, r, r_stride, m, m_stride);
      b.c = (const float2*)c;
      e = bs_launch_dadj_reduce<float2>(b, (const float2*)g, (float2*)workspace, out2, st);
    }
  } else {
    auto a = base_args<float>(p, batch, r, r_stride, m, m_stride);
    a.c = (const float*)g;
    a.out = (float*)out1;
    e = bs_launch_adjoint<float>(a, st);
    if (e == hipSuccess) {
      auto b = base_args<float>(p, batch, r, r_stride, m, m_stride);
      b.c = (const float*)c;
      e = bs_launch_dadj_reduce<float>(b, (const float*)g, (float2*)workspace, out2, st);
    }
  }
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "bspline_d1_adj");
}

}  // extern "C"

// ---- allocation-free variant of htamd_bspline_d1_adj -------------------------
// The per-tile partial sums live in a file-scope device array instead of a
// caller workspace.  Host code never passes `scratch` itself to a kernel (on
// the host that name is only a stand-in); it passes the device address that
// hipGetSymbolAddress returns, looked up once per device and kept here.
namespace {

constexpr int64_t kD1AdjScratchBytes = int64_t(4) << 20;  // 4 MiB: 128 angles of 2048^2
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kD1AdjScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// *out = device address of the current device's copy of `scratch`; the
// lookup runs once per device
int scratch_on_current_device(const char* what, void** out) {
  *out = nullptr;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, what);
  if (dev < 0 || dev >= kMaxDevices)
    return fail(HTAMD_ERR_UNSUPPORTED, std::string(what) + ": device index " + std::to_string(dev) + " out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, what);
    if (!p) return fail(HTAMD_ERR_HIP, std::string(what) + ": hipGetSymbolAddress returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return HTAMD_OK;
}

}  // namespace

extern "C" {

int htamd_bspline_d1_adj_noalloc_max_bytes(int64_t* bytes) {
  if (!bytes) return fail(HTAMD_ERR_ARG, "bspline_d1_adj_noalloc_max_bytes: NULL");
  *bytes = kD1AdjScratchBytes;
  return HTAMD_OK;
}

int htamd_bspline_d1_adj_noalloc(const htamd_shift_plan* p, const void* c, const void* g, const float* r,
                                 int64_t r_stride, const float* m, int64_t m_stride, void* out1, float* out2,
                                 int64_t batch, void* stream) {
  int s = check_common(p, batch, r, m, "bspline_d1_adj_noalloc");
  if (s || batch == 0) return s;
  if (!c || !g || !out1 || !out2) return fail(HTAMD_ERR_ARG, "bspline_d1_adj_noalloc: NULL buffer");
  int64_t need = 0;
  htamd_bspline_d1_adj_workspace_size(p, batch, &need);
  if (need > kD1AdjScratchBytes)
    return fail(HTAMD_ERR_UNSUPPORTED, "bspline_d1_adj_noalloc: " + std::to_string(batch) + " angles need " +
                                           std::to_string(need) + " bytes of scratch, the fixed maximum is " +
                                           std::to_string(kD1AdjScratchBytes) +
                                           " (split the batch or use htamd_bspline_d1_adj)");
  void* ws = nullptr;
  s = scratch_on_current_device("bspline_d1_adj_noalloc", &ws);
  if (s) return s;
  // every partial is written by the first reduction kernel before the second
  // reads it, so the scratch needs no clearing
  return htamd_bspline_d1_adj(p, c, g, r, r_stride, m, m_stride, out1, out2, ws, kD1AdjScratchBytes, batch,
                              stream);
}

}  // extern "C"
