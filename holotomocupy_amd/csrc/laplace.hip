// 3-D Laplacian of the BH regulariser (src/holotomocupy/rec_mpi.py:400-494):
// the 7-point stencil `_lap` (rec_mpi.py:430-436), periodic in y and x (the
// reference's cp.roll), with the z neighbours read from the volume and zero
// outside it (the ghost slices of rec_mpi.py:402-416 at the global boundary).
//
//   out[k] = add[k] + scale * lap(in)[z0 + k],   k = 0 .. nz_out-1
//
// Slice z0 + k may lie outside [0, nz_in) (the reference evaluates the
// Laplacian on ghost slices for its biharmonic).  A volume is processed as
// float streams: a complex64 pixel is two floats whose stencils are
// independent.  The summation order follows `_lap`: a + c + roll(b,-1,y) +
// roll(b,1,y) + roll(b,-1,x) + roll(b,1,x) - 6 b.
#include <string>

#include "../../include/htamd.h"
#include "htamd_common.hpp"
#include "htamd_internal.hpp"

namespace htamd {

constexpr int kLapThreads = 256;

__global__ __launch_bounds__(kLapThreads) void lap3d_k(const float* __restrict__ in, int64_t nz_in, int ny,
                                                        int row, int cw, int64_t z0, const float* add,
                                                        float scale, float* out) {
  const int f = blockIdx.x * kLapThreads + threadIdx.x;  // float index within the row
  if (f >= row) return;
  const int y = blockIdx.y;
  const int64_t k = blockIdx.z;
  const int64_t s = z0 + k;
  const int64_t plane = (int64_t)ny * row;
  const int64_t o = k * plane + (int64_t)y * row + f;
  auto at = [&](int64_t zz, int yy, int ff) -> float {
    return (zz >= 0 && zz < nz_in) ? in[zz * plane + (int64_t)yy * row + ff] : 0.0f;
  };
  const int yp = y + 1 == ny ? 0 : y + 1;
  const int ym = y == 0 ? ny - 1 : y - 1;
  const int fp = f + cw >= row ? f + cw - row : f + cw;
  const int fm = f - cw < 0 ? f - cw + row : f - cw;
  const float b = at(s, y, f);
  float v = at(s - 1, y, f) + at(s + 1, y, f);
  v += at(s, yp, f);
  v += at(s, ym, f);
  v += at(s, y, fp);
  v += at(s, y, fm);
  v -= 6.0f * b;
  out[o] = add ? add[o] + scale * v : scale * v;
}

}  // namespace htamd

using namespace htamd;

extern "C" {

int htamd_laplacian3d(const void* in, int64_t nz_in, int64_t ny, int64_t nx, int is_complex, int64_t z0,
                      int64_t nz_out, const void* add, float scale, void* out, void* stream) {
  if (nz_in < 0 || ny <= 0 || nx <= 0 || nz_out < 0 || (nz_in > 0 && !in) || (nz_out > 0 && !out))
    return fail(HTAMD_ERR_ARG, "laplacian3d: bad arguments");
  const int64_t row = nx * (is_complex ? 2 : 1);
  if (row > (int64_t)1 << 30 || ny > 65535 || nz_out > 65535)
    return fail(HTAMD_ERR_ARG, "laplacian3d: volume too large for one launch (ny, slices <= 65535)");
  if (out == in) return fail(HTAMD_ERR_ARG, "laplacian3d: out must not alias in");
  if (nz_out == 0) return HTAMD_OK;
  const dim3 grid((unsigned)((row + kLapThreads - 1) / kLapThreads), (unsigned)ny, (unsigned)nz_out);
  hipLaunchKernelGGL(lap3d_k, grid, dim3(kLapThreads), 0, (hipStream_t)stream, (const float*)in, nz_in, (int)ny,
                     (int)row, is_complex ? 2 : 1, z0, (const float*)add, scale, (float*)out);
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "laplacian3d");
}

}  // extern "C"
