// Fused BH gradient of the holotomography data term on gfx950.
//
// Computes what Rec.gradients_cascade computes for a chunk of angles
// (src/holotomocupy/rec_mpi.py:327-353, cascade 517-813):
//   x22 = S_j(coeff(proj)), x12 = exp(i x22), x0 = D_j(prb_j x12),
//   y0  = 2/N (x0 - d x0/|x0|), y12 = DT_j(y0),
//   gprb_j  = sum_angles y12 conj(x12),   y22 = -i y12 conj(prb_j) conj(x12),
//   (dpsi, dr) = dSadj(coeff(proj), y22),  gproj = coeff(sum_j dpsi),  gpos = dr
// (each scaled by its rho^2), with the F3 recompute of the reference done once
// per pixel and no [angles, ndist, nz, n] intermediate other than one work
// field.  Passes per (angle, distance), all HBM-streaming:
//   K1 rows:    S-gather + exp + probe product, D_x                 -> T
//   K2 columns: D_y, residual against sqrt(data), DT_y (objective)  -> T
//   K3 rows:    DT_x, probe-gradient accumulation over angles,
//               y22 and the shift-gradient partial sums            -> T
//   K4 object:  gather-form Sadj summed over distances              -> dpsi
// then the B-spline prefilter of dpsi and two small deterministic reductions.
#include <climits>

#include "bh_common.hpp"

namespace htamd {

struct BhArgs {
  int n = 0, nz = 0, npsi = 0, nzpsi = 0, ndist = 0;
  int64_t nang = 0;
  const void* c = nullptr;           // prefiltered projections [nang, nzpsi, npsi] (object type)
  const float* pos = nullptr;        // [nang, ndist, 2] (y, x)
  const float* mag = nullptr;        // [nang, ndist]
  const float2* prb = nullptr;       // [nprb, ndist, nz, n]
  const int* pidx = nullptr;         // [nang] probe of each angle (NULL: probe 0)
  const float* data = nullptr;       // [nang, ndist, nz, n] sqrt(intensity)
  const float* const* data_p = nullptr;  // or per angle: data_p[b] -> [ndist, nz, n]
  const float2* const* prb_p = nullptr;  // or per probe slot: prb_p[p] -> [ndist, nz, n]
  float2* const* gprb_p = nullptr;       // or per probe slot: gprb_p[p] -> [ndist, nz, n]
  const float4* scales = nullptr;    // [nang]: (2/data_size, rho_obj^2, rho_pos^2, unused)
  float2* T = nullptr;               // work field [nang, ndist, nz, n]
  float2* gprb = nullptr;            // [nprb, ndist, nz, n]
  const float* rho_prb = nullptr;    // [nprb] rho_prb^2
  int accumulate_prb = 0;
  float2* part_r = nullptr;          // [nang*ndist*nz] shift-gradient partials
  double* part_obj = nullptr;        // [column tiles] objective partials
  const float2* twx = nullptr;
  const float2* twy = nullptr;
  const float2* tabx = nullptr;      // [ndist][2][3*(n/2)]
  const float2* taby = nullptr;      // [ndist][2][3*(nz/2)]
};

// 16-tap B-spline gather at sampling position px of a row whose y-taps start at iy
// with weights wy (and dwy).  v = sum wy wx c; with DERIV also
// dt1 = -sum dwy wx c and dt2 = -sum wy dwx c (the dsadj weights,
// cuda_kernels.py:735-748).
template <typename TO, bool DERIV>
__device__ __forceinline__ void s_gather(const TO* __restrict__ c, int npsi, int nzpsi, BsPos px, int iy,
                                         const float* wy, const float* dwy, TO& v, TO& dt1, TO& dt2) {
  const int ix = px.i - 1;
  float wx[4], dwx[4];
  bs_w(px.d, wx);
  if (DERIV) bs_dw(px.d, dwx);
  v = vzero<TO>();
  dt1 = vzero<TO>();
  dt2 = vzero<TO>();
#pragma unroll
  for (int jy = 0; jy < 4; ++jy) {
    const int yy = iy + jy;
    if (yy < 0 || yy >= nzpsi) continue;
    TO rv = vzero<TO>(), rd = vzero<TO>();
#pragma unroll
    for (int jx = 0; jx < 4; ++jx) {
      const int xx = ix + jx;
      if (xx < 0 || xx >= npsi) continue;
      const TO cv = c[yy * npsi + xx];
      rv = vfma(wx[jx], cv, rv);
      if (DERIV) rd = vfma(dwx[jx], cv, rd);
    }
    v = vfma(wy[jy], rv, v);
    if (DERIV) {
      dt1 = vfma(-dwy[jy], rv, dt1);
      dt2 = vfma(-wy[jy], rd, dt2);
    }
  }
}

// Row-staged gather.  All pixels of a detector row sample the same four
// object rows (iy .. iy+3), so those rows are combined along y once per
// (angle, row) into LDS with coalesced reads of c, and each pixel then reads 4
// LDS values instead of 16 scattered global ones.  cy[x - lo] = sum_jy wy c,
// cdy[x - lo] = sum_jy dwy c over the columns [lo, hi) the row's taps reach
// (the same sums as s_gather, combined along y first).
template <typename TO, int T, bool DERIV>
__device__ __forceinline__ void stage_row(const TO* __restrict__ cb, int npsi, int nzpsi, int iy,
                                          const float* wy, const float* dwy, int lo, int hi, int t,
                                          TO* cy, TO* cdy) {
  for (int x = lo + t; x < hi; x += T) {
    TO acc = vzero<TO>(), dacc = vzero<TO>();
#pragma unroll
    for (int jy = 0; jy < 4; ++jy) {
      const int yy = iy + jy;
      if (yy < 0 || yy >= nzpsi) continue;
      const TO cv = cb[(int64_t)yy * npsi + x];
      acc = vfma(wy[jy], cv, acc);
      if (DERIV) dacc = vfma(dwy[jy], cv, dacc);
    }
    cy[x - lo] = acc;
    if (DERIV) cdy[x - lo] = dacc;
  }
}

template <typename TO, bool DERIV>
__device__ __forceinline__ void row_gather(const TO* cy, const TO* cdy, int lo, int hi, BsPos px, TO& v,
                                           TO& dt1, TO& dt2) {
  const int ix = px.i - 1;
  float wx[4], dwx[4];
  bs_w(px.d, wx);
  if (DERIV) bs_dw(px.d, dwx);
  v = vzero<TO>();
  dt1 = vzero<TO>();
  dt2 = vzero<TO>();
#pragma unroll
  for (int jx = 0; jx < 4; ++jx) {
    const int xx = ix + jx;
    if (xx < lo || xx >= hi) continue;  // within the row's span this is xx outside [0, npsi)
    const TO cv = cy[xx - lo];
    v = vfma(wx[jx], cv, v);
    if (DERIV) {
      dt2 = vfma(-dwx[jx], cv, dt2);
      dt1 = vfma(-wx[jx], cdy[xx - lo], dt1);
    }
  }
}

// Sum of (s1, s2) over the T threads of one row group; every thread calls it.
// The sum is returned to the row's thread t == 0 (and, for T <= 64, to all).
template <int T>
__device__ __forceinline__ float2 row_sum2(float2 v, float2* red) {
  constexpr int W = T < 64 ? T : 64;
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off, 64);
    v.y += __shfl_xor(v.y, off, 64);
  }
  if constexpr (T > 64) {
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x % T == 0) {
      const int w0 = threadIdx.x >> 6;
      v = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < T / 64; ++k) {
        v.x += red[w0 + k].x;
        v.y += red[w0 + k].y;
      }
    }
  }
  return v;
}

// ---- K1: rows, forward ------------------------------------------------------
// Workgroup = ROWS row units (j, ty); loops over the angles of the launch.
// Four waves per SIMD (at most 128 VGPRs): the staged gather waits on memory
// for about half of its cycles, and the allocator's free choice at N = 2048
// is 129 registers, i.e. three waves (1.74 against 1.95 ms per 16 angles x 4
// distances at 2048^2, no spills at any N).
template <int N, int E, typename TO, bool STAGED>
__global__ __launch_bounds__(kRowThreads) __attribute__((amdgpu_waves_per_eu(4))) void bh_fwd_rows_k(BhArgs a) {
  constexpr int T = N / E, ROWS = kRowThreads / T;
  __shared__ float2 lds[ROWS * fft_buf<N>()];
  const int r = threadIdx.x / T, t0 = threadIdx.x % T;
  const int64_t nunits = (int64_t)a.ndist * a.nz;
  const int64_t unit = xcd_tile(blockIdx.x, gridDim.x) * ROWS + r;  // XCD-contiguous row bands
  const bool valid = unit < nunits;
  const int j = valid ? (int)(unit / a.nz) : 0, ty = valid ? (int)(unit % a.nz) : 0;
  float2* s = lds + r * fft_buf<N>();
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t0, a.twx);
  const float2* tab = a.tabx + (int64_t)(j * 2 + 0) * 3 * (N / 2);
  const int64_t plane = (int64_t)a.nzpsi * a.npsi;
  for (int64_t b = 0; b < a.nang; ++b) {
    // recompute thread-dependent addresses every angle instead of keeping
    // them live across the loop (register pressure)
    int t = t0;
    asm volatile("" : "+v"(t));
    const int64_t bj = b * a.ndist + j;
    const float mb = a.mag[bj], ry = a.pos[2 * bj], rx = a.pos[2 * bj + 1];
    const BsPos py = bs_pos(ty, a.nz, a.nzpsi, mb, ry);
    const int iy = py.i - 1;
    float wy[4];
    bs_w(py.d, wy);
    const TO* cb = (const TO*)a.c + b * plane;
    const int p = a.pidx ? a.pidx[b] : 0;
    const float2* prow = (a.prb_p ? a.prb_p[p] : a.prb + (int64_t)p * a.ndist * a.nz * N) + ((int64_t)j * a.nz + ty) * N;
    if constexpr (STAGED) {
      TO* cy = (TO*)dyn_lds() + r * a.npsi;
      int lo, hi;
      row_span<N>(a.npsi, mb, rx, lo, hi);
      stage_row<TO, T, false>(cb, a.npsi, a.nzpsi, iy, wy, wy, lo, hi, t, cy, cy);
      __syncthreads();
      // the probe value of the next iteration is loaded one iteration ahead
      float2 pnext = prow[dct_nat<N>(t)];
#pragma unroll 2
      for (int e = 0; e < E; ++e) {
        const int i = dct_nat<N>(t + e * T);  // the pixel of the thread's FFT slot t + e T
        const float2 pv = pnext;
        if (e + 1 < E) pnext = prow[dct_nat<N>(t + (e + 1) * T)];
        TO v, d1, d2;
        row_gather<TO, false>(cy, cy, lo, hi, bs_pos(i, N, a.npsi, mb, rx), v, d1, d2);
        s[fpad(t + e * T)] = cmul(pv, expi_fast(to_c(v)));
      }
    } else {
#pragma unroll 1
      for (int e = 0; e < E; ++e) {
        const int i = dct_nat<N>(t + e * T);
        TO v, d1, d2;
        s_gather<TO, false>(cb, a.npsi, a.nzpsi, bs_pos(i, N, a.npsi, mb, rx), iy, wy, wy, v, d1, d2);
        s[fpad(t + e * T)] = cmul(prow[i], expi_fast(to_c(v)));
      }
    }
    // the transform's first stage reads only the slots t + e T this thread
    // wrote, so no ordering point is needed before it; the last stage writes
    // the row to T in natural order from registers
    fresnel_1d<N, E, kMaxRadix, false, LineStore<N>>(
        s, t, twr, tab, LineStore<N>{valid ? a.T + (bj * a.nz + ty) * (int64_t)N : nullptr, t, T});
    __syncthreads();  // the next angle's gather writes s
  }
}

// K2's residual 2/N (x - d x/|x|) and the objective term (|x| - d)^2 of one
// element, applied to D_y's result as the last inverse stage writes it (output
// e of a thread is element t + e N/E, whose sqrt(data) the thread holds in
// dv[e]).  |x| and d/|x| come from one reciprocal square root (v_rsq_f32,
// ~1 ulp) instead of a correctly rounded square root and division; a thread's
// E objective terms are summed in FP32 and added to the FP64 total once.
template <int E>
struct BhResidual {
  static constexpr bool active = true, store = true, sync_after = true;
  const float (&dv)[E];
  float sc;
  float& objf;
  __device__ __forceinline__ float2 operator()(int e, float2 v) const {
    const float a2 = v.x * v.x + v.y * v.y;
    const float ri = __builtin_amdgcn_rsqf(a2);
    const float q = a2 * ri - dv[e];
    objf = fmaf(q, q, objf);
    const float g = sc * (1.0f - dv[e] * ri);
    return make_float2(g * v.x, g * v.y);
  }
};

// ---- K2: columns: D_y, residual, DT_y ---------------------------------------
// The transforms alternate between two LDS tiles (one barrier per FFT stage
// instead of two, and about 50 fewer registers live at N = 2048): D_y's result
// lands in the second tile and DT_y's back in the first, from which the tile
// is stored.  The column's sqrt(data) loads are issued with the tile load, so
// they land during D_y, and the residual is formed in D_y's last stage
// (BhResidual) rather than in a pass over the tile of its own.
template <int N, int E, int C>
__global__ __launch_bounds__(C*(N / E)) void bh_cols_k(BhArgs a, int tiles_per_field) {
  constexpr int T = N / E, NT = C * T, LS = col_stride<N, C>();
  __shared__ float2 lds[2 * C * LS];
  __shared__ double red[NT / 64 > 0 ? NT / 64 : 1];
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t f = tile / tiles_per_field;  // field b*ndist + j
  const int x0 = (int)(tile % tiles_per_field) * C;
  const int64_t b = f / a.ndist;
  const int j = (int)(f % a.ndist);
  const int cc = threadIdx.x / T, t = threadIdx.x % T;
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, a.twy);
  float2* base = a.T + f * (int64_t)N * a.n + x0;
  const float* dbase = (a.data_p ? a.data_p[b] + (int64_t)j * N * a.n : a.data + f * (int64_t)N * a.n) + x0;
  col_tile_xfer<N, E, C, true>(lds, base, a.n, threadIdx.x);  // FFT slot q holds row dct_nat(q)
  float dv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) dv[e] = dbase[dct_nat<N>(t + e * T) * a.n + cc];
  __syncthreads();
  float2* s0 = lds + cc * LS;
  float2* s = lds + C * LS + cc * LS;
  const float sc = a.scales[b].x;
  float objf = 0.0f;
  const BhResidual<E> res{dv, sc, objf};
  fresnel_1d_pp<N, E, kMaxRadix, false, BhResidual<E>>(s0, s, t, twr, a.taby + (int64_t)(j * 2 + 0) * 3 * (N / 2),
                                                       res);
  double obj = (double)objf;
  fresnel_1d_pp<N, E>(s, s0, t, twr, a.taby + (int64_t)(j * 2 + 1) * 3 * (N / 2));
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  col_tile_xfer<N, E, C, false>(lds, base, a.n, tid);
  if constexpr (NT >= 64) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) obj += __shfl_xor(obj, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = obj;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = 0.0;
#pragma unroll
      for (int k = 0; k < NT / 64; ++k) tot += red[k];
      a.part_obj[tile] = tot;
    }
  } else {
    __shared__ double all[NT];
    all[threadIdx.x] = obj;
    __syncthreads();
    if (threadIdx.x == 0) {
      double tot = 0.0;
      for (int k = 0; k < NT; ++k) tot += all[k];
      a.part_obj[tile] = tot;
    }
  }
}

// ---- K3: rows, adjoint + epilogue ------------------------------------------
template <int N, int E, typename TO, bool STAGED>
__global__ __launch_bounds__(kRowThreads) void bh_adj_rows_k(BhArgs a) {
  constexpr int T = N / E, ROWS = kRowThreads / T;
  __shared__ float2 lds[ROWS * fft_buf<N>()];
  __shared__ float2 lacc[ROWS * N];  // probe-gradient accumulator of the row
  __shared__ float2 red[kRowThreads / 64];
  const int r = threadIdx.x / T, t0 = threadIdx.x % T;
  const int64_t nunits = (int64_t)a.ndist * a.nz;
  const int64_t unit = xcd_tile(blockIdx.x, gridDim.x) * ROWS + r;  // XCD-contiguous row bands
  const bool valid = unit < nunits;
  const int j = valid ? (int)(unit / a.nz) : 0, ty = valid ? (int)(unit % a.nz) : 0;
  float2* s = lds + r * fft_buf<N>();
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t0, a.twx);
  const float2* tab = a.tabx + (int64_t)(j * 2 + 1) * 3 * (N / 2);
  const int64_t plane = (int64_t)a.nzpsi * a.npsi;
  // element dct_nat(t + e T) belongs to thread t in the zeroing, the
  // epilogue and the flushes alike, so no barrier orders them
  float2* acc = lacc + r * N;
  for (int e = 0; e < E; ++e) acc[dct_nat<N>(t0 + e * T)] = make_float2(0.f, 0.f);
  int pcur = a.pidx ? a.pidx[0] : 0;
  auto flush = [&](int p) {
    if (!valid) return;
    float2* g = (a.gprb_p ? a.gprb_p[p] : a.gprb + (int64_t)p * a.ndist * a.nz * N) + ((int64_t)j * a.nz + ty) * N;
    const int t = t0;
    const float rho = a.rho_prb[p];
    for (int e = 0; e < E; ++e) {
      const int i = dct_nat<N>(t + e * T);
      float2 v = cscale(acc[i], rho);
      if (a.accumulate_prb) v = cadd(g[i], v);
      g[i] = v;
      acc[i] = make_float2(0.f, 0.f);
    }
  };
  for (int64_t b = 0; b < a.nang; ++b) {
    int t = t0;
    asm volatile("" : "+v"(t));
    const int64_t bj = b * a.ndist + j;
    float2* row = a.T + (bj * a.nz + ty) * (int64_t)N;
    // the row's FFT slots t + e T (natural elements dct_nat of them) straight
    // into registers for the first transform stage
    float2 h[E];
#pragma unroll
    for (int e = 0; e < E; ++e) h[e] = valid ? row[dct_nat<N>(t + e * T)] : make_float2(0.f, 0.f);
    const float mb = a.mag[bj], ry = a.pos[2 * bj], rx = a.pos[2 * bj + 1];
    const BsPos py = bs_pos(ty, a.nz, a.nzpsi, mb, ry);
    const int iy = py.i - 1;
    float wy[4], dwy[4];
    bs_w(py.d, wy);
    bs_dw(py.d, dwy);
    const TO* cb = (const TO*)a.c + b * plane;
    TO* cy = (TO*)dyn_lds() + 2 * r * a.npsi;
    TO* cdy = cy + a.npsi;
    int lo = 0, hi = 0;
    if constexpr (STAGED) {
      row_span<N>(a.npsi, mb, rx, lo, hi);
      stage_row<TO, T, true>(cb, a.npsi, a.nzpsi, iy, wy, dwy, lo, hi, t, cy, cdy);
    }
    // its first ordering point also orders the staged rows; the epilogue reads
    // back only this thread's outputs (elements t + e T), so the last stage
    // needs no closing ordering point
    fresnel_1d_from<N, E, kMaxRadix, false, OwnOutputs>(h, s, t, twr, tab);
    const int p = a.pidx ? a.pidx[b] : 0;
    if (p != pcur) {
      flush(pcur);
      pcur = p;
    }
    const float2* prow = (a.prb_p ? a.prb_p[p] : a.prb + (int64_t)p * a.ndist * a.nz * N) + ((int64_t)j * a.nz + ty) * N;
    float s1 = 0.f, s2 = 0.f;
    // the probe value of the next iteration is loaded one iteration ahead
    float2 pnext = prow[dct_nat<N>(t + pair_order<E>(0) * T)];
#pragma unroll 1
    for (int k = 0; k < E; ++k) {
      const int e = pair_order<E>(k);
      const int i = dct_nat<N>(t + e * T);  // natural element of the thread's output e
      const float2 pv = pnext;
      if (k + 1 < E) pnext = prow[dct_nat<N>(t + pair_order<E>(k + 1) * T)];
      const float2 y12 = s[fpad(t + e * T)];
      TO v, d1, d2;
      if constexpr (STAGED)
        row_gather<TO, true>(cy, cdy, lo, hi, bs_pos(i, N, a.npsi, mb, rx), v, d1, d2);
      else
        s_gather<TO, true>(cb, a.npsi, a.nzpsi, bs_pos(i, N, a.npsi, mb, rx), iy, wy, dwy, v, d1, d2);
      const float2 ex = expi_fast(to_c(v));
      acc[i] = cadd(acc[i], cmulc(y12, ex));                // y12 conj(x12)
      const float2 q = cmulc(cmulc(y12, pv), ex);           // y12 conj(prb) conj(x12)
      const TO g = from_c<TO>(make_float2(q.y, -q.x));      // * (-i)
      s1 += vredot(g, d1);
      s2 += vredot(g, d2);
      if (valid) row[i] = to_c(g);
    }
    const float2 tot = row_sum2<T>(make_float2(s1, s2), red);
    if (valid && t0 == 0) a.part_r[bj * a.nz + ty] = tot;
    __syncthreads();
  }
  flush(pcur);
}

// ---- K4: gather-form Sadj summed over distances ----------------------------
// K4: dpsi = rho_obj * sum_j Sadj_j(T_j) as a gather.  A workgroup owns a
// 64 x kSadjTY tile of coefficients: lane l of wave w computes column l of rows
// w*kSadjR .. w*kSadjR + kSadjR - 1, so its x weights serve kSadjR outputs and
// a wave's rows are wave-uniform.  For a distance with magnification >= 1
// every candidate bracket has fewer than 8 entries; the tile's y weights
// (kSadjTY rows x 8 candidates) are then evaluated once into LDS and read as
// broadcasts.  Smaller magnifications take the per-thread candidate walk.
// Either way each output sums over detector rows in ascending order and the
// distances in order.
constexpr int kSadjR = 4, kSadjTY = 4 * kSadjR, kSadjK = 8;

template <typename TO>
__global__ __launch_bounds__(256) void bh_sadj_k(BhArgs a, TO* __restrict__ dpsi, int tiles_x,
                                                 int tiles_y) {
  constexpr int TX = 64, R = kSadjR, TY = kSadjTY, K = kSadjK;
  __shared__ float s_wy[TY][K];
  __shared__ int s_ylo[TY];
  const int64_t bid = xcd_tile(blockIdx.x, gridDim.x);
  const int txt = (int)(bid % tiles_x);
  const int64_t rest = bid / tiles_x;
  const int tyt = (int)(rest % tiles_y);
  const int64_t b = rest / tiles_y;
  const int t = threadIdx.x;
  const int ix = txt * TX + (t & 63);
  const int r0 = (t >> 6) * R;  // first tile row of this wave
  const bool xin = ix < a.npsi;
  TO acc[R];
#pragma unroll
  for (int rr = 0; rr < R; ++rr) acc[rr] = vzero<TO>();
  for (int j = 0; j < a.ndist; ++j) {
    const int64_t bj = b * a.ndist + j;
    const float mb = a.mag[bj], ry = a.pos[2 * bj], rx = a.pos[2 * bj + 1];
    const float2* g = a.T + bj * (int64_t)a.nz * a.n;
    if (mb >= 1.0f) {  // uniform over the workgroup (one angle)
      __syncthreads();  // the previous distance's table has been read
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
        // the R outputs' candidate detector rows overlap: walk their union in
        // ascending order, form each row's x sum once and add it to every
        // output that weights it (per output the same terms in the same order)
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
          const float2* gr = g + (int64_t)ty * a.n + xlo;
          TO rowv = vzero<TO>();
#pragma unroll
          for (int kk = 0; kk < K; ++kk)
            if (wxs[kk] != 0.0f) rowv = vfma(wxs[kk], from_c<TO>(gr[kk]), rowv);
#pragma unroll
          for (int rr = 0; rr < R; ++rr)
            if (wv[rr] != 0.0f) acc[rr] = vfma(wv[rr], rowv, acc[rr]);
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
          TO rowv = vzero<TO>();
          for (int tx = xlo; tx <= xhi; ++tx) {
            const float wxv = bs_reach(bs_pos(tx, a.n, a.npsi, mb, rx), ix);
            if (wxv != 0.0f) rowv = vfma(wxv, from_c<TO>(g[(int64_t)ty * a.n + tx]), rowv);
          }
          acc[rr] = vfma(wyv, rowv, acc[rr]);
        }
      }
    }
  }
  if (!xin) return;
  const float rho = a.scales[b].y;
#pragma unroll
  for (int rr = 0; rr < R; ++rr) {
    const int iy = tyt * TY + r0 + rr;
    if (iy < a.nzpsi) dpsi[(b * a.nzpsi + iy) * (int64_t)a.npsi + ix] = vfma(rho, acc[rr], vzero<TO>());
  }
}

// ---- small reductions -------------------------------------------------------
__global__ __launch_bounds__(256) void bh_pos_final_k(const float2* __restrict__ part_r, int nz,
                                                      int ndist, const float4* __restrict__ scales,
                                                      float* __restrict__ gpos) {
  const int64_t bj = blockIdx.x;
  double s1 = 0.0, s2 = 0.0;
  for (int ty = threadIdx.x; ty < nz; ty += 256) {
    const float2 v = part_r[bj * nz + ty];
    s1 += (double)v.x;
    s2 += (double)v.y;
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
    const float rho = scales[bj / ndist].z;
    gpos[2 * bj + 0] = (float)r1[0] * rho;
    gpos[2 * bj + 1] = (float)r2[0] * rho;
  }
}

__global__ __launch_bounds__(256) void bh_sum_k(const double* __restrict__ part, int64_t count,
                                                double scale, double* __restrict__ out) {
  __shared__ double r[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < count; i += 256) s += part[i];
  r[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) r[threadIdx.x] += r[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = r[0] * scale;
}

// ---- host side --------------------------------------------------------------

struct BhLayout {
  int64_t c = 0, T = 0, dpsi = 0, part_r = 0, part_obj = 0, total = 0;
};

// Column pass of the fused gradient: 16 elements per thread at N = 2048 (two
// columns per 256-thread workgroup), else 8.
template <int N>
constexpr int bh_col_e() {
  return N >= 2048 ? 16 : 8;
}

// Column tile width C of K2 (C * 8 contiguous bytes per row): bounded by 256
// threads and the 160 KiB of LDS of a CU.  The pass does two operators per
// load and store, so the strided access matters less than for a single
// operator; at 2048 (16 elements per thread) C = 2 and C = 4 measure the
// same, 32 elements with C = 8 and 8 elements with C = 4 are slower, and
// C = 1 (8-byte rows) takes 5.5 ms against 3.6.
template <int NZ, int C>
constexpr bool col_tile_fits() {
  return C * (NZ / bh_col_e<NZ>()) <= 256 && (int64_t)2 * C * col_stride<NZ, C>() * 8 + 4096 <= 160 * 1024;
}

template <int NZ>
static int col_tiles_c(int64_t n) {
  constexpr int CM = col_tile_fits<NZ, 16>() ? 16 : col_tile_fits<NZ, 8>() ? 8 : col_tile_fits<NZ, 4>() ? 4
                     : col_tile_fits<NZ, 2>() ? 2 : 1;
  for (int c = CM; c > 1; c >>= 1)
    if (n % c == 0) return c;
  return 1;
}

static int cols_per_tile(int64_t nz, int64_t n) {
  switch (nz) {
    case 8: return col_tiles_c<8>(n);
    case 16: return col_tiles_c<16>(n);
    case 32: return col_tiles_c<32>(n);
    case 64: return col_tiles_c<64>(n);
    case 128: return col_tiles_c<128>(n);
    case 256: return col_tiles_c<256>(n);
    case 512: return col_tiles_c<512>(n);
    case 1024: return col_tiles_c<1024>(n);
    case 2048: return col_tiles_c<2048>(n);
    default: return 1;
  }
}

static BhLayout bh_layout(const htamd_prop_plan* pp, const htamd_shift_plan* sp, int64_t nang) {
  BhLayout L;
  const int64_t osz = sp->is_complex ? 8 : 4;
  const int64_t plane = sp->nzpsi * sp->npsi;
  const int64_t field = pp->n * pp->nz;
  const int64_t ntiles = nang * pp->ndist * (pp->n / cols_per_tile(pp->nz, pp->n));
  int64_t off = 0;
  L.c = off;
  off += align256(nang * plane * osz);
  L.T = off;
  off += align256(nang * pp->ndist * field * 8);
  L.dpsi = off;
  off += align256(nang * plane * osz);
  L.part_r = off;
  off += align256(nang * pp->ndist * pp->nz * 8);
  L.part_obj = off;
  off += align256(ntiles * 8);
  L.total = off;
  return L;
}

// LDS of one workgroup: the statically sized buffers plus the staged object
// rows (one row of npsi values per detector row, two in the adjoint pass);
// rows are staged when the staged part fits 64 KiB and the total the 160 KiB
// of a CU.
constexpr int64_t kRowDynBudget = 65536;
constexpr int64_t kRowLdsBudget = 160 * 1024;

// Row-pass elements per thread (one row per 256 threads at N = 2048; 16
// elements with radix-16 stages measured slower: fewer waves per CU).
template <int N>
constexpr int bh_row_e() {
  return 8;
}

template <int N, typename TO>
static hipError_t launch_rows(bool adjoint, const BhArgs& a, hipStream_t st) {
  constexpr int E = bh_row_e<N>();
  constexpr int ROWS = kRowThreads / (N / E);
  const int64_t units = (int64_t)a.ndist * a.nz;
  const int64_t grid = (units + ROWS - 1) / ROWS;
  const int64_t stat = (int64_t)ROWS * (fft_buf<N>() + (adjoint ? N : 0)) * 8 + 64;
  const int64_t dyn = (int64_t)ROWS * a.npsi * (int64_t)sizeof(TO) * (adjoint ? 2 : 1);
  const bool staged = dyn <= kRowDynBudget && stat + dyn <= kRowLdsBudget;
  if (adjoint) {
    if (staged)
      hipLaunchKernelGGL((bh_adj_rows_k<N, E, TO, true>), dim3((unsigned)grid), dim3(kRowThreads), (size_t)dyn,
                         st, a);
    else
      hipLaunchKernelGGL((bh_adj_rows_k<N, E, TO, false>), dim3((unsigned)grid), dim3(kRowThreads), 0, st, a);
  } else {
    if (staged)
      hipLaunchKernelGGL((bh_fwd_rows_k<N, E, TO, true>), dim3((unsigned)grid), dim3(kRowThreads), (size_t)dyn,
                         st, a);
    else
      hipLaunchKernelGGL((bh_fwd_rows_k<N, E, TO, false>), dim3((unsigned)grid), dim3(kRowThreads), 0, st, a);
  }
  return hipGetLastError();
}

template <typename TO>
static hipError_t rows_dispatch(bool adjoint, const BhArgs& a, hipStream_t st) {
  switch (a.n) {
    case 8: return launch_rows<8, TO>(adjoint, a, st);
    case 16: return launch_rows<16, TO>(adjoint, a, st);
    case 32: return launch_rows<32, TO>(adjoint, a, st);
    case 64: return launch_rows<64, TO>(adjoint, a, st);
    case 128: return launch_rows<128, TO>(adjoint, a, st);
    case 256: return launch_rows<256, TO>(adjoint, a, st);
    case 512: return launch_rows<512, TO>(adjoint, a, st);
    case 1024: return launch_rows<1024, TO>(adjoint, a, st);
    case 2048: return launch_rows<2048, TO>(adjoint, a, st);
    default: return hipErrorInvalidValue;
  }
}


template <int NZ>
static hipError_t launch_cols(const BhArgs& a, hipStream_t st) {
  constexpr int E = bh_col_e<NZ>();
  const int C = cols_per_tile(NZ, a.n);
  const int tpf = a.n / C;
  const int64_t grid = a.nang * a.ndist * tpf;
  switch (C) {
#define HTAMD_BHC(CC)                                                                                 \
  case CC:                                                                                            \
    if constexpr (col_tile_fits<NZ, CC>())                                                            \
      hipLaunchKernelGGL((bh_cols_k<NZ, E, CC>), dim3((unsigned)grid), dim3(CC * (NZ / E)), 0, st, a, \
                         tpf);                                                                        \
    else                                                                                              \
      return hipErrorInvalidValue;                                                                    \
    break;
    HTAMD_BHC(1) HTAMD_BHC(2) HTAMD_BHC(4) HTAMD_BHC(8) HTAMD_BHC(16)
#undef HTAMD_BHC
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

static hipError_t cols_dispatch(const BhArgs& a, hipStream_t st) {
  switch (a.nz) {
    case 8: return launch_cols<8>(a, st);
    case 16: return launch_cols<16>(a, st);
    case 32: return launch_cols<32>(a, st);
    case 64: return launch_cols<64>(a, st);
    case 128: return launch_cols<128>(a, st);
    case 256: return launch_cols<256>(a, st);
    case 512: return launch_cols<512>(a, st);
    case 1024: return launch_cols<1024>(a, st);
    case 2048: return launch_cols<2048>(a, st);
    default: return hipErrorInvalidValue;
  }
}

// proj_p / gproj_p: optional per-angle addresses of the projections and of
// their gradients (NULL: proj / gproj are [nang, nzpsi, npsi] arrays)
template <typename TO>
static hipError_t bh_run(const htamd_shift_plan* sp, BhArgs a, const void* proj, void* gproj,
                         float* gpos, double* objective, const BhLayout& L, char* ws, hipStream_t st,
                         const StageEvents& sev, const void* const* proj_p = nullptr,
                         void* const* gproj_p = nullptr) {
  hipError_t e;
  // coefficients of the projections (coeff, shift.py:74-82)
  sev.mark(0, st);
  PlaneTab<TO> tin;
  tin.in = (const TO* const*)proj_p;
  if ((e = bs_launch_prefilter<TO>((const TO*)proj, (TO*)(ws + L.c), a.nang, a.nzpsi, a.npsi, sp->taps,
                                   st, tin)) != hipSuccess)
    return e;
  sev.mark(1, st);
  sev.mark(2, st);
  if ((e = rows_dispatch<TO>(false, a, st)) != hipSuccess) return e;
  sev.mark(3, st);
  sev.mark(4, st);
  if ((e = cols_dispatch(a, st)) != hipSuccess) return e;
  sev.mark(5, st);
  sev.mark(6, st);
  if ((e = rows_dispatch<TO>(true, a, st)) != hipSuccess) return e;
  sev.mark(7, st);
  const int tx = (a.npsi + 63) / 64, ty = (a.nzpsi + kSadjTY - 1) / kSadjTY;
  TO* dpsi = (TO*)(ws + L.dpsi);
  sev.mark(8, st);
  hipLaunchKernelGGL((bh_sadj_k<TO>), dim3((unsigned)(a.nang * tx * ty)), dim3(256), 0, st, a, dpsi, tx, ty);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  sev.mark(9, st);
  sev.mark(10, st);
  PlaneTab<TO> tout;
  tout.out = (TO* const*)gproj_p;
  if ((e = bs_launch_prefilter<TO>(dpsi, (TO*)gproj, a.nang, a.nzpsi, a.npsi, sp->taps, st, tout)) != hipSuccess)
    return e;
  sev.mark(11, st);
  sev.mark(12, st);
  hipLaunchKernelGGL(bh_pos_final_k, dim3((unsigned)(a.nang * a.ndist)), dim3(256), 0, st, a.part_r, a.nz,
                     a.ndist, a.scales, gpos);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (objective) {
    const int64_t ntiles = a.nang * a.ndist * (a.n / cols_per_tile(a.nz, a.n));
    hipLaunchKernelGGL(bh_sum_k, dim3(1), dim3(256), 0, st, a.part_obj, ntiles, 1.0, objective);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  sev.mark(13, st);
  return hipSuccess;
}

}  // namespace htamd

using namespace htamd;

extern "C" {

int htamd_bh_plan_create(const htamd_prop_plan* prop, const htamd_shift_plan* shift,
                         htamd_bh_plan** plan) {
  if (!plan || !prop || !shift) return fail(HTAMD_ERR_ARG, "bh_plan_create: NULL argument");
  *plan = nullptr;
  if (prop->n != shift->n || prop->nz != shift->nz)
    return fail(HTAMD_ERR_ARG, "bh_plan_create: propagator and shift plans disagree on the detector size");
  if ((shift->npsi % 2) || (shift->nzpsi % 2))
    return fail(HTAMD_ERR_UNSUPPORTED, "bh_plan_create: object grid sizes must be even");
  if (!fft_size_supported(prop->n) || !fft_size_supported(prop->nz))
    return fail(HTAMD_ERR_UNSUPPORTED,
                "bh_plan_create: the fused passes take power-of-two detector sizes in [8, 2048]; other sizes run "
                "on the per-operator kernels");
  auto* p = new htamd_bh_plan();
  p->prop = prop;
  p->shift = shift;
  *plan = p;
  return HTAMD_OK;
}

int htamd_bh_plan_destroy(htamd_bh_plan* p) {
  delete p;
  return HTAMD_OK;
}

int htamd_bh_set_stage_events(htamd_bh_plan* p, void** events, int count) {
  if (!p || count < 0 || (count > 0 && !events)) return fail(HTAMD_ERR_ARG, "bh_set_stage_events: bad arguments");
  p->events.ev = (hipEvent_t*)events;
  p->events.count = count;
  return HTAMD_OK;
}

int htamd_bh_workspace_size(const htamd_bh_plan* p, int64_t nangles, int64_t* bytes) {
  if (!p || !bytes || nangles < 0) return fail(HTAMD_ERR_ARG, "bh_workspace_size: bad arguments");
  *bytes = bh_layout(p->prop, p->shift, nangles).total;
  return HTAMD_OK;
}

// Either the packed arrays (proj, prb, data, gproj, gprb) or, with tab != NULL,
// the address table tab = [proj x nangles | data x nangles | gproj x nangles |
// prb x nprobes | gprb x nprobes] (device memory).
static int bh_gradient_impl(const htamd_bh_plan* p, int64_t nangles, const void* proj, const float* pos,
                            const float* mag, const void* prb, const int* probe_index, const float* data,
                            const void* const* tab, int64_t nprobes, const float* angle_scales,
                            const float* rho_prb, int accumulate_prb, void* gproj, float* gpos, void* gprb,
                            double* objective, void* workspace, int64_t workspace_bytes, void* stream) {
  if (!p) return fail(HTAMD_ERR_ARG, "bh_gradient: plan is NULL");
  if (nangles < 0) return fail(HTAMD_ERR_ARG, "bh_gradient: negative angle count");
  if (nangles == 0) return HTAMD_OK;
  if (tab ? (nprobes < 1 || !pos || !mag || !angle_scales || !rho_prb || !gpos)
          : (!proj || !pos || !mag || !prb || !data || !angle_scales || !rho_prb || !gproj || !gpos || !gprb))
    return fail(HTAMD_ERR_ARG, "bh_gradient: NULL buffer");
  if (tab && nprobes > 1 && !probe_index)
    return fail(HTAMD_ERR_ARG, "bh_gradient_tab: several probes need a probe index");
  const BhLayout L = bh_layout(p->prop, p->shift, nangles);
  if (!workspace || workspace_bytes < L.total)
    return fail(HTAMD_ERR_ARG, "bh_gradient: workspace of " + std::to_string(L.total) + " bytes required");
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != p->prop->device) return fail(HTAMD_ERR_ARG, "bh_gradient: plan belongs to another device");
  char* ws = (char*)workspace;
  BhArgs a;
  a.n = (int)p->prop->n;
  a.nz = (int)p->prop->nz;
  a.npsi = (int)p->shift->npsi;
  a.nzpsi = (int)p->shift->nzpsi;
  a.ndist = (int)p->prop->ndist;
  a.nang = nangles;
  a.c = ws + L.c;
  a.pos = pos;
  a.mag = mag;
  a.prb = (const float2*)prb;
  a.pidx = probe_index;
  a.data = data;
  const void* const* proj_p = nullptr;
  void* const* gproj_p = nullptr;
  if (tab) {
    proj_p = tab;
    a.data_p = (const float* const*)(tab + nangles);
    gproj_p = (void* const*)(tab + 2 * nangles);
    a.prb_p = (const float2* const*)(tab + 3 * nangles);
    a.gprb_p = (float2* const*)(tab + 3 * nangles + nprobes);
  }
  a.scales = (const float4*)angle_scales;
  a.T = (float2*)(ws + L.T);
  a.gprb = (float2*)gprb;
  a.rho_prb = rho_prb;
  a.accumulate_prb = accumulate_prb;
  a.part_r = (float2*)(ws + L.part_r);
  a.part_obj = (double*)(ws + L.part_obj);
  a.twx = p->prop->tw_x;
  a.twy = p->prop->tw_y;
  a.tabx = p->prop->tab_x;
  a.taby = p->prop->tab_y;
  hipError_t e = p->shift->is_complex
      ? bh_run<float2>(p->shift, a, proj, gproj, gpos, objective, L, ws, (hipStream_t)stream, p->events, proj_p,
                       gproj_p)
      : bh_run<float>(p->shift, a, proj, gproj, gpos, objective, L, ws, (hipStream_t)stream, p->events, proj_p,
                      gproj_p);
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, "bh_gradient");
}

int htamd_bh_gradient(const htamd_bh_plan* p, int64_t nangles, const void* proj, const float* pos,
                      const float* mag, const void* prb, const int* probe_index, const float* data,
                      const float* angle_scales, const float* rho_prb, int accumulate_prb, void* gproj,
                      float* gpos, void* gprb, double* objective, void* workspace,
                      int64_t workspace_bytes, void* stream) {
  return bh_gradient_impl(p, nangles, proj, pos, mag, prb, probe_index, data, nullptr, 0, angle_scales, rho_prb,
                          accumulate_prb, gproj, gpos, gprb, objective, workspace, workspace_bytes, stream);
}

int htamd_bh_gradient_tab(const htamd_bh_plan* p, int64_t nangles, int64_t nprobes, const void* const* tables,
                          const float* pos, const float* mag, const int* probe_index, const float* angle_scales,
                          const float* rho_prb, int accumulate_prb, float* gpos, double* objective, void* workspace,
                          int64_t workspace_bytes, void* stream) {
  if (!tables) return fail(HTAMD_ERR_ARG, "bh_gradient_tab: NULL address table");
  return bh_gradient_impl(p, nangles, nullptr, pos, mag, nullptr, probe_index, nullptr, tables, nprobes,
                          angle_scales, rho_prb, accumulate_prb, nullptr, gpos, nullptr, objective, workspace,
                          workspace_bytes, stream);
}

}  // extern "C"
