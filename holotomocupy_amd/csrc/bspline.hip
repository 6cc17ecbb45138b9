// Cubic B-spline shift family for gfx950: prefilter (coeff), S, Sadj, dS,
// d2S, dSadj.  Semantics follow src/holotomocupy/shift.py:18-221 and the
// kernels of src/holotomocupy/cuda_kernels.py:131-843; the implementation is
// gather-only (no atomics) so every result is deterministic.
//
// Sampling point of output pixel (tx, ty) of angle b (cuda_kernels.py:172-181):
//   x = (m_b (tx - n/2) - r_b[1] + (m_b - 1)/2) + npsi/2      (same for y, r_b[0])
// taps floor(x) + j, j = -1..2, weights phi(dx - j), phi = 6 * cubic B-spline.
#include <climits>

#include "htamd_bspline.hpp"
#include "htamd_internal.hpp"

namespace htamd {

// ---------------------------------------------------------------------------
// S: out[b, ty, tx] = sum_{jy,jx} wy wx c[b, iy+jy, ix+jx]   (taps in range)
// dS:  -(dwx wy drx + wx dwy dry) c + wx wy c1
// d2S: (d2wx wy d1x d2x + dwx dwy (d1x d2y + d1y d2x) + wx d2wy d1y d2y) c
//      - (dwx wy d1x + wx dwy d1y) c1 - (dwx wy d2x + wx dwy d2y) c2
// (d2S pairs Deltar1 with c1, as cuda_kernels.py:437-450 does.)
enum { BS_S = 0, BS_D1 = 1, BS_D2 = 2 };

constexpr int kTX = 64, kTY = 4;

template <typename T, int MODE>
__global__ __launch_bounds__(kTX* kTY) void bs_forward_k(BsArgs<T> a) {
  const int64_t bid = blockIdx.x;
  const int txt = (int)(bid % a.tiles_x);
  const int64_t rest = bid / a.tiles_x;
  const int tyt = (int)(rest % a.tiles_y);
  const int64_t b = rest / a.tiles_y;
  const int tx = txt * kTX + (int)(threadIdx.x % kTX);
  const int ty = tyt * kTY + (int)(threadIdx.x / kTX);
  if (tx >= a.n || ty >= a.nz) return;
  const float mb = a.m[b * a.ms];
  const BsPos px = bs_pos(tx, a.n, a.npsi, mb, a.r[b * a.rs + 1]);
  const BsPos py = bs_pos(ty, a.nz, a.nzpsi, mb, a.r[b * a.rs + 0]);
  const int ix = px.i - 1, iy = py.i - 1;
  const float xd = px.d, yd = py.d;
  float wx[4], wy[4];
  bs_w(xd, wx);
  bs_w(yd, wy);
  const int64_t plane = (int64_t)a.nzpsi * a.npsi;
  const T* c = a.c + b * plane;
  T acc = vzero<T>();
  if constexpr (MODE == BS_S) {
#pragma unroll
    for (int jy = 0; jy < 4; ++jy) {
      const int yy = iy + jy;
      if (yy < 0 || yy >= a.nzpsi) continue;
      T row = vzero<T>();
#pragma unroll
      for (int jx = 0; jx < 4; ++jx) {
        const int xx = ix + jx;
        if (xx < 0 || xx >= a.npsi) continue;
        row = vfma(wx[jx], c[(int64_t)yy * a.npsi + xx], row);
      }
      acc = vfma(wy[jy], row, acc);
    }
  } else {
    float dwx[4], dwy[4];
    bs_dw(xd, dwx);
    bs_dw(yd, dwy);
    const T* c1 = a.c1 + b * plane;
    if constexpr (MODE == BS_D1) {
      const float drx = a.dr1[b * a.d1s + 1], dry = a.dr1[b * a.d1s + 0];
#pragma unroll
      for (int jy = 0; jy < 4; ++jy) {
        const int yy = iy + jy;
        if (yy < 0 || yy >= a.nzpsi) continue;
#pragma unroll
        for (int jx = 0; jx < 4; ++jx) {
          const int xx = ix + jx;
          if (xx < 0 || xx >= a.npsi) continue;
          const int64_t k = (int64_t)yy * a.npsi + xx;
          const float w = dwx[jx] * wy[jy] * drx + wx[jx] * dwy[jy] * dry;
          acc = vfma(-w, c[k], acc);
          acc = vfma(wx[jx] * wy[jy], c1[k], acc);
        }
      }
    } else {
      float d2wx[4], d2wy[4];
      bs_d2w(xd, d2wx);
      bs_d2w(yd, d2wy);
      const T* c2 = a.c2 + b * plane;
      const float d1x = a.dr1[b * a.d1s + 1], d1y = a.dr1[b * a.d1s + 0];
      const float d2x = a.dr2[b * a.d2s + 1], d2y = a.dr2[b * a.d2s + 0];
      const float cross = d1x * d2y + d1y * d2x;
#pragma unroll
      for (int jy = 0; jy < 4; ++jy) {
        const int yy = iy + jy;
        if (yy < 0 || yy >= a.nzpsi) continue;
#pragma unroll
        for (int jx = 0; jx < 4; ++jx) {
          const int xx = ix + jx;
          if (xx < 0 || xx >= a.npsi) continue;
          const int64_t k = (int64_t)yy * a.npsi + xx;
          const float w = d2wx[jx] * wy[jy] * d1x * d2x + dwx[jx] * dwy[jy] * cross +
                          wx[jx] * d2wy[jy] * d1y * d2y;
          const float w1 = dwx[jx] * wy[jy] * d1x + wx[jx] * dwy[jy] * d1y;
          const float w2 = dwx[jx] * wy[jy] * d2x + wx[jx] * dwy[jy] * d2y;
          acc = vfma(w, c[k], acc);
          acc = vfma(-w1, c1[k], acc);
          acc = vfma(-w2, c2[k], acc);
        }
      }
    }
  }
  a.out[(b * a.nz + ty) * (int64_t)a.n + tx] = acc;
}

// S with a wavefront per 64 columns x kSRY consecutive output rows.  The rows'
// y windows overlap (three of four taps at magnification 1), so the wave walks
// the coefficient rows its outputs need in ascending order, forms each row's
// x-tap sum once and adds it to every output whose window holds that row --
// the same sums in the same order as bs_forward_k<.., BS_S>, with about
// (kSRY + 3) / kSRY row reads per output instead of four.
constexpr int kSRY = 8;

template <typename T>
__global__ __launch_bounds__(256) void bs_interp_rows_k(BsArgs<T> a) {
  constexpr int RY = kSRY;
  const int64_t bid = blockIdx.x;
  const int txt = (int)(bid % a.tiles_x);
  const int64_t rest = bid / a.tiles_x;
  const int tyt = (int)(rest % a.tiles_y);
  const int64_t b = rest / a.tiles_y;
  const int tx = txt * 64 + (int)(threadIdx.x % 64);
  const int ty0 = tyt * (4 * RY) + __builtin_amdgcn_readfirstlane((int)(threadIdx.x / 64)) * RY;
  if (ty0 >= a.nz) return;  // the whole wave
  const float mb = a.m[b * a.ms];
  const BsPos px = bs_pos(tx, a.n, a.npsi, mb, a.r[b * a.rs + 1]);
  const int ix = px.i - 1;
  float wx[4];
  bs_w(px.d, wx);
  const float ry = a.r[b * a.rs + 0];
  int iy[RY];
  float wy[RY][4];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const BsPos py = bs_pos(ty0 + r, a.nz, a.nzpsi, mb, ry);
    iy[r] = py.i - 1;
    bs_w(py.d, wy[r]);
  }
  const T* c = a.c + b * (int64_t)a.nzpsi * a.npsi;
  T acc[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) acc[r] = vzero<T>();
  // rows past nz are evaluated like the others and not stored; their windows
  // only extend the walk
  int ylo = iy[0], yhi = iy[0];
#pragma unroll
  for (int r = 1; r < RY; ++r) {
    ylo = min(ylo, iy[r]);
    yhi = max(yhi, iy[r]);
  }
  ylo = max(ylo, 0);
  yhi = min(yhi + 3, a.nzpsi - 1);
  for (int yy = ylo; yy <= yhi; ++yy) {
    T row = vzero<T>();
#pragma unroll
    for (int jx = 0; jx < 4; ++jx) {
      const int xx = ix + jx;
      if (xx < 0 || xx >= a.npsi) continue;
      row = vfma(wx[jx], c[(int64_t)yy * a.npsi + xx], row);
    }
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const int d = yy - iy[r];
      if (d >= 0 && d < 4) {
        const float w = d == 0 ? wy[r][0] : d == 1 ? wy[r][1] : d == 2 ? wy[r][2] : wy[r][3];
        acc[r] = vfma(w, row, acc[r]);
      }
    }
  }
  if (tx < a.n) {
#pragma unroll
    for (int r = 0; r < RY; ++r)
      if (ty0 + r < a.nz) a.out[(b * a.nz + ty0 + r) * (int64_t)a.n + tx] = acc[r];
  }
}

// ---------------------------------------------------------------------------
// Sadj as a gather over the coefficient grid: c[b, iy, ix] =
//   sum_{ty: iy in taps(y(ty))} wy(ty) sum_{tx: ix in taps(x(tx))} wx(tx) g[b, ty, tx].
// Candidate pixels are bracketed analytically with a one-pixel margin and each
// is re-evaluated with the forward's exact float expression, so the weights
// are bitwise those of S.  Tiles of 64 x kAdjTY coefficients: lane l of wave w
// computes column l of rows w*kAdjR .. w*kAdjR + kAdjR - 1, so its x weights
// serve kAdjR outputs.  With magnification >= 1 every candidate bracket holds
// fewer than 8 entries (an entry past 8 is the outer margin, whose weight is
// zero), and the tile's y weights are evaluated once into LDS; smaller
// magnifications walk the brackets per thread.  Each output sums detector rows
// in ascending order either way.
constexpr int kAdjR = 4, kAdjTY = 4 * kAdjR, kAdjK = 8;

template <typename T>
__global__ __launch_bounds__(256) void bs_adjoint_k(BsArgs<T> a) {
  // here a.c is the input g [B, nz, n] and a.out the coefficients [B, nzpsi, npsi]
  constexpr int R = kAdjR, TY = kAdjTY, K = kAdjK;
  __shared__ float s_wy[TY][K];
  __shared__ int s_ylo[TY];
  const int64_t bid = blockIdx.x;
  const int txt = (int)(bid % a.tiles_x);
  const int64_t rest = bid / a.tiles_x;
  const int tyt = (int)(rest % a.tiles_y);
  const int64_t b = rest / a.tiles_y;
  const int t = threadIdx.x;
  const int ix = txt * 64 + (t & 63);
  const int r0 = (t >> 6) * R;
  const bool xin = ix < a.npsi;
  const float mb = a.m[b * a.ms];
  const float rx = a.r[b * a.rs + 1], ry = a.r[b * a.rs + 0];
  const T* g = a.c + b * (int64_t)a.nz * a.n;
  T acc[R];
#pragma unroll
  for (int rr = 0; rr < R; ++rr) acc[rr] = vzero<T>();
  if (mb >= 1.0f) {  // uniform over the workgroup (one angle)
    if (t < TY * K) {
      const int row = t / K, k = t % K, iy = tyt * TY + row;
      float w = 0.0f;
      int ylo = 0, yhi = -1;
      if (iy < a.nzpsi) {
        bs_candidates(iy, a.nz, a.nzpsi, mb, ry, ylo, yhi);
        if (ylo + k <= yhi) w = bs_reach(bs_pos(ylo + k, a.nz, a.nzpsi, mb, ry), iy);
      }
      s_wy[row][k] = w;
      if (k == 0) s_ylo[row] = ylo;
    }
    __syncthreads();
    if (xin) {
      int xlo, xhi;
      bs_candidates(ix, a.n, a.npsi, mb, rx, xlo, xhi);
      float wxs[K];
#pragma unroll
      for (int k = 0; k < K; ++k)
        wxs[k] = xlo + k <= xhi ? bs_reach(bs_pos(xlo + k, a.n, a.npsi, mb, rx), ix) : 0.0f;
      // the R outputs' candidate detector rows overlap (three of four at
      // magnification 1): walk their union in ascending order, form each
      // row's x sum once and add it to every output that weights it -- per
      // output the same terms in the same order as a walk of its own rows
      // over the rows that exist (rows past the plane have no candidates)
      int ylo = INT_MAX, yhi = INT_MIN;
#pragma unroll
      for (int rr = 0; rr < R; ++rr) {
        if (tyt * TY + r0 + rr >= a.nzpsi) continue;
        ylo = min(ylo, s_ylo[r0 + rr]);
        yhi = max(yhi, s_ylo[r0 + rr]);
      }
      yhi += K - 1;
      for (int ty = ylo; ty <= yhi; ++ty) {
        float wv[R];
        bool any = false;
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
          const int d = ty - s_ylo[r0 + rr];
          wv[rr] = d >= 0 && d < K ? s_wy[r0 + rr][d] : 0.0f;
          any = any || wv[rr] != 0.0f;
        }
        if (!any) continue;  // wave-uniform: the rows are the wave's
        const T* gr = g + (int64_t)ty * a.n + xlo;
        T row = vzero<T>();
#pragma unroll
        for (int kk = 0; kk < K; ++kk)
          if (wxs[kk] != 0.0f) row = vfma(wxs[kk], gr[kk], row);
#pragma unroll
        for (int rr = 0; rr < R; ++rr)
          if (wv[rr] != 0.0f) acc[rr] = vfma(wv[rr], row, acc[rr]);
      }
    }
  } else if (xin) {
    int xlo, xhi;
    bs_candidates(ix, a.n, a.npsi, mb, rx, xlo, xhi);
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      const int iy = tyt * TY + r0 + rr;
      if (iy >= a.nzpsi) continue;
      int ylo, yhi;
      bs_candidates(iy, a.nz, a.nzpsi, mb, ry, ylo, yhi);
      for (int ty = ylo; ty <= yhi; ++ty) {
        const float wyv = bs_reach(bs_pos(ty, a.nz, a.nzpsi, mb, ry), iy);
        if (wyv == 0.0f) continue;
        T row = vzero<T>();
        for (int tx = xlo; tx <= xhi; ++tx) {
          const float wxv = bs_reach(bs_pos(tx, a.n, a.npsi, mb, rx), ix);
          if (wxv != 0.0f) row = vfma(wxv, g[(int64_t)ty * a.n + tx], row);
        }
        acc[rr] = vfma(wyv, row, acc[rr]);
      }
    }
  }
  if (!xin) return;
#pragma unroll
  for (int rr = 0; rr < R; ++rr) {
    const int iy = tyt * TY + r0 + rr;
    if (iy < a.nzpsi) a.out[(b * a.nzpsi + iy) * (int64_t)a.npsi + ix] = acc[rr];
  }
}

// ---------------------------------------------------------------------------
// dSadj reductions: per angle, out2[b,0] = Re<g, dt1>, out2[b,1] = Re<g, dt2>
// with dt1 = -sum wx dwy c, dt2 = -sum dwx wy c at every pixel.  Stage 1 writes
// one partial pair per workgroup; stage 2 sums them per angle in a fixed order.
template <typename T>
__global__ __launch_bounds__(kTX* kTY) void bs_dadj_partial_k(BsArgs<T> a, const T* __restrict__ g,
                                                              float2* __restrict__ partial) {
  const int64_t bid = blockIdx.x;
  const int txt = (int)(bid % a.tiles_x);
  const int64_t rest = bid / a.tiles_x;
  const int tyt = (int)(rest % a.tiles_y);
  const int64_t b = rest / a.tiles_y;
  const int tx = txt * kTX + (int)(threadIdx.x % kTX);
  const int ty = tyt * kTY + (int)(threadIdx.x / kTX);
  float s1 = 0.0f, s2 = 0.0f;
  if (tx < a.n && ty < a.nz) {
    const float mb = a.m[b * a.ms];
    const BsPos px = bs_pos(tx, a.n, a.npsi, mb, a.r[b * a.rs + 1]);
    const BsPos py = bs_pos(ty, a.nz, a.nzpsi, mb, a.r[b * a.rs + 0]);
    const int ix = px.i - 1, iy = py.i - 1;
    float wx[4], wy[4], dwx[4], dwy[4];
    bs_w(px.d, wx);
    bs_w(py.d, wy);
    bs_dw(px.d, dwx);
    bs_dw(py.d, dwy);
    const T* c = a.c + b * (int64_t)a.nzpsi * a.npsi;
    T dt1 = vzero<T>(), dt2 = vzero<T>();
#pragma unroll
    for (int jy = 0; jy < 4; ++jy) {
      const int yy = iy + jy;
      if (yy < 0 || yy >= a.nzpsi) continue;
#pragma unroll
      for (int jx = 0; jx < 4; ++jx) {
        const int xx = ix + jx;
        if (xx < 0 || xx >= a.npsi) continue;
        const T cv = c[(int64_t)yy * a.npsi + xx];
        dt1 = vfma(-wx[jx] * dwy[jy], cv, dt1);
        dt2 = vfma(-dwx[jx] * wy[jy], cv, dt2);
      }
    }
    const T gv = g[(b * a.nz + ty) * (int64_t)a.n + tx];
    s1 = vredot(gv, dt1);
    s2 = vredot(gv, dt2);
  }
  const float2 tot = block_sum2<kTX * kTY>(make_float2(s1, s2));
  if (threadIdx.x == 0) partial[bid] = tot;
}

__global__ __launch_bounds__(256) void bs_dadj_final_k(const float2* __restrict__ partial,
                                                       int64_t per_angle, float* __restrict__ out2) {
  const int64_t b = blockIdx.x;
  double s1 = 0.0, s2 = 0.0;
  for (int64_t i = threadIdx.x; i < per_angle; i += 256) {
    const float2 p = partial[b * per_angle + i];
    s1 += (double)p.x;
    s2 += (double)p.y;
  }
  __shared__ double r1[256], r2[256];
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out2[2 * b + 0] = (float)r1[0];
    out2[2 * b + 1] = (float)r2[0];
  }
}

// ---------------------------------------------------------------------------
// Prefilter (coeff): periodic separable FIR with the exact periodic inverse of
// the [1 4 1] circulant (shift.py:25-30, 74-82), taps |d| <= kPfH (host-built,
// FP64; the dropped tail is below 1e-8 relative).  One pass: a (TY+2H) x
// (TX+2H) tile with periodic halo is staged in LDS, filtered along x into a
// second LDS tile, then along y, and written once.
constexpr int kPfTX = 64, kPfTY = 32, kPfR = 8;

template <typename T>
__global__ __launch_bounds__(256) void bs_prefilter_k(const T* __restrict__ in, T* __restrict__ out,
                                                      int nzpsi, int npsi, int tiles_x, int tiles_y,
                                                      PrefilterTaps taps, PlaneTab<T> tab) {
  constexpr int H = kPfH;
  constexpr int LX = kPfTX + 2 * H, LY = kPfTY + 2 * H;
  __shared__ T s_in[LY][LX + 1];
  __shared__ T s_tmp[LY][kPfTX + 1];
  const int64_t bid = blockIdx.x;
  const int txt = (int)(bid % tiles_x);
  const int64_t rest = bid / tiles_x;
  const int tyt = (int)(rest % tiles_y);
  const int64_t b = rest / tiles_y;
  const int x0 = txt * kPfTX, y0 = tyt * kPfTY;
  const int64_t plane = (int64_t)nzpsi * npsi;
  const T* src = tab.in ? tab.in[b] : in + b * plane;
  for (int i = threadIdx.x; i < LY * LX; i += 256) {
    const int ly = i / LX, lx = i % LX;
    int gy = (y0 + ly - H) % nzpsi;
    gy = gy < 0 ? gy + nzpsi : gy;
    int gx = (x0 + lx - H) % npsi;
    gx = gx < 0 ? gx + npsi : gx;
    s_in[ly][lx] = src[(int64_t)gy * npsi + gx];
  }
  __syncthreads();
  // each thread filters kPfR consecutive outputs from a register window of
  // kPfR + 2H inputs (one LDS read per input instead of one per tap);
  // along x consecutive lanes take consecutive rows (distinct LDS banks)
  constexpr int R = kPfR, WIN = kPfR + 2 * H;
  for (int i = threadIdx.x; i < LY * (kPfTX / R); i += 256) {
    const int ly = i % LY, gx = (i / LY) * R;
    T win[WIN];
#pragma unroll
    for (int k = 0; k < WIN; ++k) win[k] = s_in[ly][gx + k];
#pragma unroll
    for (int o = 0; o < R; ++o) {
      T acc = vzero<T>();
#pragma unroll
      for (int d = 0; d <= 2 * H; ++d) acc = vfma(taps.wx[d], win[o + d], acc);
      s_tmp[ly][gx + o] = acc;
    }
  }
  __syncthreads();
  T* dst = tab.out ? tab.out[b] : out + b * plane;
  for (int i = threadIdx.x; i < kPfTX * (kPfTY / R); i += 256) {
    const int lx = i % kPfTX, gy = (i / kPfTX) * R;
    if (x0 + lx >= npsi) continue;
    T win[WIN];
#pragma unroll
    for (int k = 0; k < WIN; ++k) win[k] = s_tmp[gy + k][lx];
#pragma unroll
    for (int o = 0; o < R; ++o) {
      if (y0 + gy + o >= nzpsi) break;
      T acc = vzero<T>();
#pragma unroll
      for (int d = 0; d <= 2 * H; ++d) acc = vfma(taps.wy[d], win[o + d], acc);
      dst[(int64_t)(y0 + gy + o) * npsi + x0 + lx] = acc;
    }
  }
}

// Column-strip form for planes at least kPsTX wide.  A block of kPsNT threads
// covers kPsW = kPsTX + 2H columns (thread t owns periodic column
// x0 - H + t) and a strip of `sy` output rows.  Each thread slides a register
// window of R + 2H rows down its column (coalesced row loads, one periodic
// wrap test per row); per R-row chunk it writes R y-filtered values to LDS,
// and after one barrier each of kPsTX / 8 x R work items filters 8
// consecutive outputs of one row from a register window of 8 + 2H LDS values
// and stores them.  The LDS rows hold one pad slot per 8 values, so the 16
// lanes of one ds_read_b64 pass hit distinct banks.
constexpr int kPsTX = 256, kPsW = kPsTX + 2 * kPfH, kPsNT = 320, kPsR = 8;
constexpr int kPsRS = kPsW + kPsW / 8 + 1;  // padded LDS row, in elements
static_assert(kPsNT >= kPsW && kPsNT % 64 == 0, "one thread per column of the panel");
static_assert(kPsR * (kPsTX / 8) <= kPsNT, "one x work item per thread");

template <typename T>
__global__ __launch_bounds__(kPsNT) void bs_prefilter_strip_k(const T* __restrict__ in, T* __restrict__ out,
                                                              int nzpsi, int npsi, int tiles_x, int strips,
                                                              int sy, PrefilterTaps taps, PlaneTab<T> tab) {
  constexpr int H = kPfH, R = kPsR, WIN = kPsR + 2 * kPfH, G = kPsTX / 8;
  __shared__ T tmp[R * kPsRS];
  const int64_t bid = blockIdx.x;
  const int strip = (int)(bid % strips);
  const int64_t rest = bid / strips;
  const int txt = (int)(rest % tiles_x);
  const int64_t b = rest / tiles_x;
  const int x0 = txt * kPsTX, y0 = strip * sy;
  const int t = threadIdx.x;
  const int64_t plane = (int64_t)nzpsi * npsi;
  const bool loader = t < kPsW;
  int gx = (x0 - H + t) % npsi;
  gx = gx < 0 ? gx + npsi : gx;
  const T* src = (tab.in ? tab.in[b] : in + b * plane) + gx;
  int gy = (y0 - H) % nzpsi;
  gy = gy < 0 ? gy + nzpsi : gy;
  T win[WIN];
#pragma unroll
  for (int k = 0; k < WIN; ++k) {
    win[k] = loader ? src[(int64_t)gy * npsi] : vzero<T>();
    gy = gy + 1 == nzpsi ? 0 : gy + 1;
  }
  T* dst = tab.out ? tab.out[b] : out + b * plane;
  const int ystop = min(sy, nzpsi - y0);
  const int xr = t / G, xg = t % G;  // x work item: row xr, outputs 8 xg .. 8 xg + 7
  for (int c = 0; c < ystop; c += R) {
    if (loader) {
#pragma unroll
      for (int o = 0; o < R; ++o) {
        T acc = vzero<T>();
#pragma unroll
        for (int d = 0; d <= 2 * H; ++d) acc = vfma(taps.wy[d], win[o + d], acc);
        tmp[o * kPsRS + t + t / 8] = acc;
      }
    }
    if (c + R < ystop) {  // slide and fetch the next chunk's rows before the x pass
#pragma unroll
      for (int k = 0; k < 2 * H; ++k) win[k] = win[k + R];
#pragma unroll
      for (int k = 2 * H; k < WIN; ++k) {
        win[k] = loader ? src[(int64_t)gy * npsi] : vzero<T>();
        gy = gy + 1 == nzpsi ? 0 : gy + 1;
      }
    }
    __syncthreads();
    if (xr < R && c + xr < ystop) {
      T xw[WIN];
#pragma unroll
      for (int k = 0; k < WIN; ++k) {
        const int col = 8 * xg + k;
        xw[k] = tmp[xr * kPsRS + col + col / 8];
      }
      T* row = dst + (int64_t)(y0 + c + xr) * npsi;
#pragma unroll
      for (int o = 0; o < 8; ++o) {
        const int xx = x0 + 8 * xg + o;
        if (xx >= npsi) break;
        T acc = vzero<T>();
#pragma unroll
        for (int d = 0; d <= 2 * H; ++d) acc = vfma(taps.wx[d], xw[o + d], acc);
        row[xx] = acc;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// host launchers
template <typename T>
static void tiles_of(BsArgs<T>& a, bool coeff_grid) {
  const int w = coeff_grid ? a.npsi : a.n, h = coeff_grid ? a.nzpsi : a.nz;
  a.tiles_x = (w + kTX - 1) / kTX;
  a.tiles_y = (h + kTY - 1) / kTY;
}

template <typename T>
hipError_t bs_launch_forward(BsArgs<T> a, int mode, hipStream_t st) {
  if (mode == BS_S) {
    a.tiles_x = (a.n + 63) / 64;
    a.tiles_y = (a.nz + 4 * kSRY - 1) / (4 * kSRY);
    const int64_t grid = a.batch * (int64_t)a.tiles_x * a.tiles_y;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((bs_interp_rows_k<T>), dim3((unsigned)grid), dim3(256), 0, st, a);
    return hipGetLastError();
  }
  tiles_of(a, false);
  const int64_t grid = a.batch * (int64_t)a.tiles_x * a.tiles_y;
  if (grid == 0) return hipSuccess;
  if (mode == BS_D1)
    hipLaunchKernelGGL((bs_forward_k<T, BS_D1>), dim3((unsigned)grid), dim3(kTX * kTY), 0, st, a);
  else
    hipLaunchKernelGGL((bs_forward_k<T, BS_D2>), dim3((unsigned)grid), dim3(kTX * kTY), 0, st, a);
  return hipGetLastError();
}

template <typename T>
hipError_t bs_launch_adjoint(BsArgs<T> a, hipStream_t st) {
  a.tiles_x = (a.npsi + 63) / 64;
  a.tiles_y = (a.nzpsi + kAdjTY - 1) / kAdjTY;
  const int64_t grid = a.batch * (int64_t)a.tiles_x * a.tiles_y;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((bs_adjoint_k<T>), dim3((unsigned)grid), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <typename T>
int64_t bs_dadj_partials(const BsArgs<T>& a0) {
  BsArgs<T> a = a0;
  tiles_of(a, false);
  return a.batch * (int64_t)a.tiles_x * a.tiles_y;
}

template <typename T>
hipError_t bs_launch_dadj_reduce(BsArgs<T> a, const T* g, float2* partial, float* out2,
                                 hipStream_t st) {
  tiles_of(a, false);
  const int64_t per_angle = (int64_t)a.tiles_x * a.tiles_y;
  const int64_t grid = a.batch * per_angle;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((bs_dadj_partial_k<T>), dim3((unsigned)grid), dim3(kTX * kTY), 0, st, a, g,
                     partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(bs_dadj_final_k, dim3((unsigned)a.batch), dim3(256), 0, st, partial, per_angle,
                     out2);
  return hipGetLastError();
}

template <typename T>
hipError_t bs_launch_prefilter(const T* in, T* out, int64_t batch, int nzpsi, int npsi,
                               const PrefilterTaps& taps, hipStream_t st, PlaneTab<T> tab) {
  if (npsi >= kPsTX) {
    // strips of sy rows (a multiple of kPsR): 128 rows (halo overhead 24 / 128)
    // unless that leaves fewer than 256 blocks; measured on 16 and 72 planes of
    // 1024^2 and 8 of 2048^2, taller strips beat more blocks at every size
    const int tiles_x = (npsi + kPsTX - 1) / kPsTX;
    int sy = 128;
    while (sy > 2 * kPsR && batch * tiles_x * (int64_t)((nzpsi + sy - 1) / sy) < 256) sy /= 2;
    const int strips = (nzpsi + sy - 1) / sy;
    const int64_t grid = batch * (int64_t)tiles_x * strips;
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL((bs_prefilter_strip_k<T>), dim3((unsigned)grid), dim3(kPsNT), 0, st, in, out, nzpsi, npsi,
                       tiles_x, strips, sy, taps, tab);
    return hipGetLastError();
  }
  const int tx = (npsi + kPfTX - 1) / kPfTX, ty = (nzpsi + kPfTY - 1) / kPfTY;
  const int64_t grid = batch * (int64_t)tx * ty;
  if (grid == 0) return hipSuccess;
  hipLaunchKernelGGL((bs_prefilter_k<T>), dim3((unsigned)grid), dim3(256), 0, st, in, out, nzpsi,
                     npsi, tx, ty, taps, tab);
  return hipGetLastError();
}

#define HTAMD_BS_INST(T)                                                                         \
  template hipError_t bs_launch_forward<T>(BsArgs<T>, int, hipStream_t);                         \
  template hipError_t bs_launch_adjoint<T>(BsArgs<T>, hipStream_t);                              \
  template int64_t bs_dadj_partials<T>(const BsArgs<T>&);                                        \
  template hipError_t bs_launch_dadj_reduce<T>(BsArgs<T>, const T*, float2*, float*, hipStream_t); \
  template hipError_t bs_launch_prefilter<T>(const T*, T*, int64_t, int, int, const PrefilterTaps&, \
                                             hipStream_t, PlaneTab<T>);
HTAMD_BS_INST(float)
HTAMD_BS_INST(float2)

}  // namespace htamd
