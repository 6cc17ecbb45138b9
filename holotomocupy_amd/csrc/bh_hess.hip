// Fused BH bilinear form of the holotomography data term on gfx950.
//
// Computes what Rec.hessian_cascade computes for a chunk of angles
// (src/holotomocupy/rec_mpi.py:250-303, cascade 517-813): with the point
// x = (prb, proj, pos) and directions y, z,
//   c = coeff(proj), cy = coeff(proj_y), cz = coeff(proj_z)
//   x22 = S(c),  y22 = dS(c; cy, dr_y),  z22 = dS(c; cz, dr_z),
//   w22 = d2S(c; cy, dr_y, cz, dr_z)            (cuda_kernels.py:377-450, pairing kept)
//   e = exp(i x22), y12 = i e y22, z12 = i e z22, w12 = e (i w22 - y22 z22)
//   x0 = D(p e), y0 = D(py e + p y12), z0 = D(pz e + p z12),
//   w0 = D(py z12 + pz y12 + p w12)
//   H = 2/N sum [(1 - d/|x0|) Re(conj(y0) z0)
//                + d/|x0|^3 Re(conj(x0) y0) Re(conj(x0) z0) + (1 - d/|x0|) Re(conj(x0) w0)]
// (rec_mpi.py:539-569 for the last line).  With z = y the z fields are the y
// fields and w0 = D(2 py y12 + p w12), as the reference's y_is_z branch.
//
// Passes, per (angle, distance):
//   H1 chain:   staged-row gathers of c, cy, cz with the y weights and their
//               derivatives, exp and the probe products: the four (three,
//               five) input fields                                    -> fields
//   H1 rows:    D_x of every field in place (the operator's row pass,
//               one strided batch per distance)
//   H2 columns: D_y of each field and the pointwise sum; one FP64 partial
//               per column tile, summed in a fixed order
// The shift terms are sums over the same 16 taps as S: along y they are
// combined once per (angle, row) into five LDS rows
//   R0 = sum wy c,  Ay = sum wy cy - dry_y sum dwy c,  Az (same with z),
//   W1 = cross sum dwy c - dry_x sum wy cy - drz_x sum wy cz,
//   W0 = dry_y drz_y sum d2wy c - dry_y sum dwy cy - drz_y sum dwy cz
// so that per pixel  x22 = sum wx R0,  y22 = sum wx Ay - dry_x sum dwx R0,
//   w22 = dry_x drz_x sum d2wx R0 + sum dwx W1 + sum wx W0.
#include <string>

#include "bh_common.hpp"

namespace htamd {

struct BhhArgs {
  int n = 0, nz = 0, npsi = 0, nzpsi = 0, ndist = 0;
  int64_t nang = 0;
  const void* c = nullptr;      // prefiltered projections, [nang, nzpsi, npsi] (object type)
  const void* cy = nullptr;     // prefiltered direction y projections
  const void* cz = nullptr;     // prefiltered direction z projections (== cy when z is y)
  const float* pos = nullptr;   // [nang, ndist, 2] (y, x)
  const float* dy = nullptr;    // direction y positions
  const float* dz = nullptr;    // direction z positions
  const float* mag = nullptr;   // [nang, ndist]
  const float2* prb = nullptr;  // [ndist, nz, n]
  const float2* py = nullptr;
  const float2* pz = nullptr;
  const float* data = nullptr;  // [nang, ndist, nz, n]
  // x, y, z, w fields and the second-order field of the (z, z) form (pair only), [nang, ndist, nz, n]
  float2* F[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  double* part = nullptr;       // [column tiles]
  const float2* twx = nullptr;
  const float2* twy = nullptr;
  const float2* tabx = nullptr;
  const float2* taby = nullptr;
};

__device__ __forceinline__ float2 mul_i(float2 v) { return make_float2(-v.y, v.x); }

// ---- H1: rows ---------------------------------------------------------------
// MODE: kGeneral = H(y, z); kSame = H(y, y); kPair = H(y, z) and H(z, z) in
// one pass (the beta step of the BH loop: they share the point and z).
enum { kGeneral = 0, kSame = 1, kPair = 2 };

template <int MODE>
struct HessShape {
  static constexpr bool SAME = MODE == kSame, PAIR = MODE == kPair;
  static constexpr int NR = SAME ? 4 : PAIR ? 7 : 5;  // staged rows: R0, Ay, (Az), W1, W0, (W1zz, W0zz)
  static constexpr int NF = SAME ? 3 : PAIR ? 5 : 4;  // fields: x, y, (z), w, (wzz)
  static constexpr int NOUT = PAIR ? 2 : 1;
};

// H1 is two launches: this kernel evaluates the shift chain per pixel and
// writes the NF input fields of the row transform (natural order), and the
// operator's row pass (fresnel_rows_k) then applies D_x to all of them in
// place.  Nothing is held across transforms, so no per-pixel array is kept
// (the fused form held four or five complex values per pixel across its NF
// transforms, in scratch).  A workgroup = rpb row units (j, ty) of T = N/E
// threads; it loops over the angles.  rpb <= kRowThreads / T, chosen so the
// staged rows fit the dynamic LDS.
template <int N, int E, typename TO, int MODE>
__global__ __launch_bounds__(kRowThreads) void bhh_chain_k(BhhArgs a, int rpb) {
  using HS = HessShape<MODE>;
  constexpr bool SAME = HS::SAME, PAIR = HS::PAIR;
  constexpr int T = N / E;
  constexpr int NR = HS::NR;
  const int r = threadIdx.x / T, t0 = threadIdx.x % T;
  const int64_t nunits = (int64_t)a.ndist * a.nz;
  const int64_t unit = xcd_tile(blockIdx.x, gridDim.x) * rpb + r;
  const bool valid = unit < nunits;
  const int j = valid ? (int)(unit / a.nz) : 0, ty = valid ? (int)(unit % a.nz) : 0;
  const int64_t plane = (int64_t)a.nzpsi * a.npsi;
  TO* rows = (TO*)dyn_lds() + (int64_t)r * NR * a.npsi;
  TO* R0 = rows;
  TO* Ay = rows + a.npsi;
  TO* Az = SAME ? Ay : rows + 2 * a.npsi;
  TO* W1 = rows + (SAME ? 2 : 3) * a.npsi;
  TO* W0 = rows + (SAME ? 3 : 4) * a.npsi;
  TO* V1 = rows + 5 * a.npsi;  // pair: W1 and W0 of the (z, z) form
  TO* V0 = rows + 6 * a.npsi;
  const int64_t prow = ((int64_t)j * a.nz + ty) * N;
  for (int64_t b = 0; b < a.nang; ++b) {
    int t = t0;
    asm volatile("" : "+v"(t));
    const int64_t bj = b * a.ndist + j;
    const float mb = a.mag[bj], ry = a.pos[2 * bj], rx = a.pos[2 * bj + 1];
    const float yyx = a.dy[2 * bj + 1], yyy = a.dy[2 * bj];
    const float zzx = SAME ? yyx : a.dz[2 * bj + 1], zzy = SAME ? yyy : a.dz[2 * bj];
    const float cross = yyx * zzy + yyy * zzx;
    const BsPos py = bs_pos(ty, a.nz, a.nzpsi, mb, ry);
    const int iy = py.i - 1;
    float wy[4], dwy[4], d2wy[4];
    bs_w(py.d, wy);
    bs_dw(py.d, dwy);
    bs_d2w(py.d, d2wy);
    const TO* cb = (const TO*)a.c + b * plane;
    const TO* cyb = (const TO*)a.cy + b * plane;
    const TO* czb = (const TO*)a.cz + b * plane;
    int lo, hi;
    row_span<N>(a.npsi, mb, rx, lo, hi);
    // two columns in flight per thread: their loads overlap
#pragma unroll 2
    for (int x = lo + t; x < hi; x += T) {
      TO r0 = vzero<TO>(), r1 = vzero<TO>(), r2 = vzero<TO>();
      TO y0 = vzero<TO>(), y1 = vzero<TO>(), z0 = vzero<TO>(), z1 = vzero<TO>();
#pragma unroll
      for (int jy = 0; jy < 4; ++jy) {
        const int yy = iy + jy;
        if (yy < 0 || yy >= a.nzpsi) continue;
        const int64_t k = (int64_t)yy * a.npsi + x;
        const TO cv = cb[k], cyv = cyb[k];
        r0 = vfma(wy[jy], cv, r0);
        r1 = vfma(dwy[jy], cv, r1);
        r2 = vfma(d2wy[jy], cv, r2);
        y0 = vfma(wy[jy], cyv, y0);
        y1 = vfma(dwy[jy], cyv, y1);
        if constexpr (!SAME) {
          const TO czv = czb[k];
          z0 = vfma(wy[jy], czv, z0);
          z1 = vfma(dwy[jy], czv, z1);
        }
      }
      if constexpr (SAME) {
        z0 = y0;
        z1 = y1;
      }
      const int o = x - lo;
      R0[o] = r0;
      Ay[o] = vfma(-yyy, r1, y0);
      if constexpr (!SAME) Az[o] = vfma(-zzy, r1, z0);
      W1[o] = vfma(-zzx, z0, vfma(-yyx, y0, vfma(cross, r1, vzero<TO>())));
      W0[o] = vfma(-zzy, z1, vfma(-yyy, y1, vfma(yyy * zzy, r2, vzero<TO>())));
      if constexpr (PAIR) {
        V1[o] = vfma(-2.0f * zzx, z0, vfma(2.0f * zzx * zzy, r1, vzero<TO>()));
        V0[o] = vfma(-2.0f * zzy, z1, vfma(zzy * zzy, r2, vzero<TO>()));
      }
    }
    __syncthreads();
    const int64_t orow = (bj * a.nz + ty) * (int64_t)N;
    // the probe values of the next iteration are loaded one iteration ahead
    float2 pn = a.prb[prow + t], pyn = a.py[prow + t];
    float2 pzn = SAME ? make_float2(0.f, 0.f) : a.pz[prow + t];
#pragma unroll 2
    for (int e = 0; e < E; ++e) {
      const int i = t + e * T;
      const float2 p = pn, pyv = pyn, pzv = pzn;
      if (e + 1 < E) {
        pn = a.prb[prow + i + T];
        pyn = a.py[prow + i + T];
        if constexpr (!SAME) pzn = a.pz[prow + i + T];
      }
      const BsPos px = bs_pos(i, N, a.npsi, mb, rx);
      const int ix = px.i - 1;
      float wx[4], dwx[4], d2wx[4];
      bs_w(px.d, wx);
      bs_dw(px.d, dwx);
      bs_d2w(px.d, d2wx);
      TO x22 = vzero<TO>(), sx = vzero<TO>(), sxx = vzero<TO>(), ay = vzero<TO>(), az = vzero<TO>();
      TO w1 = vzero<TO>(), w0 = vzero<TO>(), v1 = vzero<TO>(), v0 = vzero<TO>();
#pragma unroll
      for (int jx = 0; jx < 4; ++jx) {
        const int xx = ix + jx;
        if (xx < lo || xx >= hi) continue;  // within the row's span: xx outside [0, npsi)
        const int o = xx - lo;
        const TO rv = R0[o];
        x22 = vfma(wx[jx], rv, x22);
        sx = vfma(dwx[jx], rv, sx);
        sxx = vfma(d2wx[jx], rv, sxx);
        ay = vfma(wx[jx], Ay[o], ay);
        if constexpr (!SAME) az = vfma(wx[jx], Az[o], az);
        w1 = vfma(dwx[jx], W1[o], w1);
        w0 = vfma(wx[jx], W0[o], w0);
        if constexpr (PAIR) {
          v1 = vfma(dwx[jx], V1[o], v1);
          v0 = vfma(wx[jx], V0[o], v0);
        }
      }
      const float2 yv = to_c(vfma(-yyx, sx, ay));
      const float2 wv = to_c(vfma(yyx * zzx, sxx, vfma(1.0f, w1, w0)));
      const float2 ex = expi_fast(to_c(x22));
      const float2 y12 = cmul(ex, mul_i(yv));
      if (valid) {
        a.F[0][orow + i] = cmul(p, ex);
        a.F[1][orow + i] = cadd(cmul(pyv, ex), cmul(p, y12));
        if constexpr (SAME) {
          const float2 yz = cmul(yv, yv);
          const float2 w12 = cmul(ex, make_float2(-wv.y - yz.x, wv.x - yz.y));
          a.F[3][orow + i] = cadd(cscale(cmul(pyv, y12), 2.0f), cmul(p, w12));
        } else {
          const float2 zv = to_c(vfma(-zzx, sx, az));
          const float2 z12 = cmul(ex, mul_i(zv));
          const float2 yz = cmul(yv, zv);
          const float2 w12 = cmul(ex, make_float2(-wv.y - yz.x, wv.x - yz.y));
          a.F[2][orow + i] = cadd(cmul(pzv, ex), cmul(p, z12));
          a.F[3][orow + i] = cadd(cadd(cmul(pyv, z12), cmul(pzv, y12)), cmul(p, w12));
          if constexpr (PAIR) {
            const float2 vv = to_c(vfma(zzx * zzx, sxx, vfma(1.0f, v1, v0)));
            const float2 zz = cmul(zv, zv);
            const float2 v12 = cmul(ex, make_float2(-vv.y - zz.x, vv.x - zz.y));
            a.F[4][orow + i] = cadd(cscale(cmul(pzv, z12), 2.0f), cmul(p, v12));
          }
        }
      }
    }
    __syncthreads();  // the next angle overwrites the staged rows
  }
}

// ---- H2: columns ------------------------------------------------------------
// The fields of a column tile are transformed one after another in two LDS
// tiles (ping-pong stages: one barrier per FFT stage).  While one field is
// transformed, the next field's tile is already loading into registers, and
// sqrt(data) loads with the first field.
template <int N, int E, int C>
struct HColPipe {
  static constexpr int T = N / E, LS = col_stride<N, C>();
  float2* A;                 // this thread's column of the tile the fields are loaded into
  const MakhoulRows<N, E> rows;
  const int q0;
  float2 nx[E];              // the next field's tile values (transfer mapping)
  __device__ __forceinline__ HColPipe(float2* lds, int tid, int n)
      : A(lds + (tid % C) * LS), rows(tid / C, tid % C, n), q0(tid / C) {}
  __device__ __forceinline__ void fetch(const float2* base) {
#pragma unroll
    for (int e = 0; e < E; ++e) nx[e] = base[rows.off(e)];
  }
  // store the fetched tile (FFT slot q holds row dct_nat(q)); the caller
  // orders it before the transform
  __device__ __forceinline__ void put() {
#pragma unroll
    for (int e = 0; e < E; ++e) A[slot_run<N, E>(q0, e)] = nx[e];
  }
};

// Transform the fetched field (the next one starts loading); epi consumes
// the result in the last stage's registers (output e of this thread is
// element t + e N/E of its column), so nothing is written back to LDS.
template <int N, int E, int C, typename EPI>
__device__ __forceinline__ void hcol_transform_pp(HColPipe<N, E, C>& pp, float2* lds, const float2* next, int t,
                                                  int cc, const FftTw<N, E>& twr, const float2* tab,
                                                  const EPI& epi) {
  constexpr int LS = col_stride<N, C>();
  pp.put();
  if (next) pp.fetch(next);
  __syncthreads();
  fresnel_1d_pp<N, E, kMaxRadix, false, EPI>(lds + cc * LS, lds + C * LS + cc * LS, t, twr, tab, epi);
}

// Consumers of the column transforms (output hooks, nothing stored).
// x0 or y0 kept in registers:
template <int E>
struct HcolKeep {
  static constexpr bool active = true, store = false, sync_after = true;
  float2 (&out)[E];
  __device__ __forceinline__ float2 operator()(int e, float2 v) const {
    out[e] = v;
    return v;
  }
};

// acc += (1 - d/|x0|) Re(conj(x0) w0):
template <int E>
struct HcolDot {
  static constexpr bool active = true, store = false, sync_after = true;
  const float2 (&xv)[E];
  const float (&dv)[E];
  double& acc;
  __device__ __forceinline__ float2 operator()(int e, float2 w) const {
    const float am = sqrtf(xv[e].x * xv[e].x + xv[e].y * xv[e].y);
    acc += (double)((1.0f - dv[e] / am) * (xv[e].x * w.x + xv[e].y * w.y));
    return w;
  }
};

// equal form, y0 = z0: acc += (1 - d/|x0|) |y0|^2 + d/|x0|^3 Re(conj(x0) y0)^2
template <int E>
struct HcolSame {
  static constexpr bool active = true, store = false, sync_after = true;
  const float2 (&xv)[E];
  const float (&dv)[E];
  double& acc;
  __device__ __forceinline__ float2 operator()(int e, float2 yv) const {
    const float am = sqrtf(xv[e].x * xv[e].x + xv[e].y * xv[e].y);
    const float xy = xv[e].x * yv.x + xv[e].y * yv.y;
    const float t1 = (1.0f - dv[e] / am) * (yv.x * yv.x + yv.y * yv.y);
    const float t2 = dv[e] / (am * am * am) * xy * xy;
    acc += (double)(t1 + t2);
    return yv;
  }
};

// general form: acc += g Re(conj(y0) z0) + h Re(conj(x0) y0) Re(conj(x0) z0),
// and for the pair acc2 += g |z0|^2 + h Re(conj(x0) z0)^2 (g = 1 - d/|x0|,
// h = d/|x0|^3):
template <int E, bool PAIR>
struct HcolCross {
  static constexpr bool active = true, store = false, sync_after = true;
  const float2 (&xv)[E];
  const float2 (&yv)[E];
  const float (&dv)[E];
  double& acc;
  double& acc2;
  __device__ __forceinline__ float2 operator()(int e, float2 zv) const {
    const float am = sqrtf(xv[e].x * xv[e].x + xv[e].y * xv[e].y);
    const float g = 1.0f - dv[e] / am, h = dv[e] / (am * am * am);
    const float xy = xv[e].x * yv[e].x + xv[e].y * yv[e].y;
    const float xz = xv[e].x * zv.x + xv[e].y * zv.y;
    acc += (double)(g * (yv[e].x * zv.x + yv[e].y * zv.y) + h * xy * xz);
    if constexpr (PAIR) acc2 += (double)(g * (zv.x * zv.x + zv.y * zv.y) + h * xz * xz);
    return zv;
  }
};

template <int NT>
__device__ __forceinline__ double block_sum_d(double v, double* red) {
  if constexpr (NT >= 64) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double tot = 0.0;
    if (threadIdx.x == 0) {
#pragma unroll
      for (int k = 0; k < NT / 64; ++k) tot += red[k];
    }
    __syncthreads();
    return tot;
  } else {
    red[threadIdx.x] = v;
    __syncthreads();
    double tot = 0.0;
    if (threadIdx.x == 0)
      for (int k = 0; k < NT; ++k) tot += red[k];
    __syncthreads();
    return tot;
  }
}

template <int N, int E, int C, int MODE>
__global__ __launch_bounds__(C*(N / E)) void bhh_cols_k(BhhArgs a, int tiles_per_field) {
  using HS = HessShape<MODE>;
  constexpr bool SAME = HS::SAME, PAIR = HS::PAIR;
  constexpr int T = N / E, NT = C * T, LS = col_stride<N, C>();
  __shared__ float2 lds[2 * C * LS];
  __shared__ double red[NT >= 64 ? NT / 64 : NT];
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t f = tile / tiles_per_field;  // field b*ndist + j
  const int x0c = (int)(tile % tiles_per_field) * C;
  const int j = (int)(f % a.ndist);
  const int cc = threadIdx.x / T, t = threadIdx.x % T;
  FftTw<N, E> twr;
  fft_load_twiddles<N, E>(twr, t, a.twy);
  const int64_t off = f * (int64_t)N * a.n + x0c;
  const float2* tab = a.taby + (int64_t)(j * 2 + 0) * 3 * (N / 2);
  HColPipe<N, E, C> pp(lds, threadIdx.x, a.n);
  // fields in the order x, w, (w of (z, z)), y, (z)
  pp.fetch(a.F[0] + off);
  float dv[E];
#pragma unroll
  for (int e = 0; e < E; ++e) dv[e] = a.data[off + (int64_t)dct_nat<N>(t + e * T) * a.n + cc];
  float2 xv[E];  // x0 of this thread's elements
  hcol_transform_pp<N, E, C>(pp, lds, a.F[3] + off, t, cc, twr, tab, HcolKeep<E>{xv});
  double acc = 0.0, acc2 = 0.0;
  hcol_transform_pp<N, E, C>(pp, lds, PAIR ? a.F[4] + off : a.F[1] + off, t, cc, twr, tab, HcolDot<E>{xv, dv, acc});
  if constexpr (PAIR) hcol_transform_pp<N, E, C>(pp, lds, a.F[1] + off, t, cc, twr, tab, HcolDot<E>{xv, dv, acc2});
  if constexpr (SAME) {
    hcol_transform_pp<N, E, C>(pp, lds, nullptr, t, cc, twr, tab, HcolSame<E>{xv, dv, acc});
  } else {
    float2 yv[E];
    hcol_transform_pp<N, E, C>(pp, lds, a.F[2] + off, t, cc, twr, tab, HcolKeep<E>{yv});
    hcol_transform_pp<N, E, C>(pp, lds, nullptr, t, cc, twr, tab, HcolCross<E, PAIR>{xv, yv, dv, acc, acc2});
  }
  const double tot = block_sum_d<NT>(acc, red);
  if constexpr (PAIR) {
    const double tot2 = block_sum_d<NT>(acc2, red);
    if (threadIdx.x == 0) {
      a.part[2 * tile] = tot;
      a.part[2 * tile + 1] = tot2;
    }
  } else {
    if (threadIdx.x == 0) a.part[tile] = tot;
  }
}

// First stage of the partial sums: block g of nb sums tiles
// [g * chunk, min(count, (g + 1) * chunk)) in a fixed order into
// part2[g * nout + k]; bhh_sum_k then sums the nb block sums.
constexpr int kSumBlocks = 256;

__global__ __launch_bounds__(256) void bhh_sum1_k(const double* __restrict__ part, int64_t count, int nout,
                                                  int64_t chunk, double* __restrict__ part2) {
  __shared__ double r[256];
  const int64_t lo = blockIdx.x * chunk, hi = lo + chunk < count ? lo + chunk : count;
  for (int k = 0; k < nout; ++k) {
    double s = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) s += part[i * nout + k];
    r[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) r[threadIdx.x] += r[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) part2[blockIdx.x * nout + k] = r[0];
    __syncthreads();
  }
}

// out[k] (+)= scale * sum over tiles t of part[t * nout + k], fixed order
__global__ __launch_bounds__(256) void bhh_sum_k(const double* __restrict__ part, int64_t count, int nout,
                                                 double scale, int accumulate, double* __restrict__ out) {
  __shared__ double r[256];
  for (int k = 0; k < nout; ++k) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < count; i += 256) s += part[i * nout + k];
    r[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) r[threadIdx.x] += r[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = accumulate ? out[k] + r[0] * scale : r[0] * scale;
    __syncthreads();
  }
}

// ---- host side --------------------------------------------------------------
// Column pass: 8 elements per thread (x0, sqrt(data) and y0 stay in registers
// across the transforms), C columns per workgroup within 256 threads.
template <int N>
constexpr int bhh_col_e() {
  return N >= 8 ? 8 : N;
}

template <int N>
constexpr int bhh_cols_c() {
  constexpr int T = N / bhh_col_e<N>();
  return T >= 256 ? 1 : 256 / T > 4 ? 4 : 256 / T;
}

static int hess_cols_c(int64_t nz) {
  switch (nz) {
    case 8: return bhh_cols_c<8>();
    case 16: return bhh_cols_c<16>();
    case 32: return bhh_cols_c<32>();
    case 64: return bhh_cols_c<64>();
    case 128: return bhh_cols_c<128>();
    case 256: return bhh_cols_c<256>();
    case 512: return bhh_cols_c<512>();
    case 1024: return bhh_cols_c<1024>();
    case 2048: return bhh_cols_c<2048>();
    default: return 1;
  }
}

struct BhhLayout {
  int64_t c = 0, cy = 0, cz = 0, F[5] = {0, 0, 0, 0, 0}, part = 0, part2 = 0, total = 0;
};

static BhhLayout bhh_layout(const htamd_prop_plan* pp, const htamd_shift_plan* sp, int64_t nang, int mode) {
  BhhLayout L;
  const bool same = mode == kSame, pair = mode == kPair;
  const int64_t osz = sp->is_complex ? 8 : 4;
  const int64_t cbytes = align256(nang * sp->nzpsi * sp->npsi * osz);
  const int64_t fbytes = align256(nang * pp->ndist * pp->n * pp->nz * 8);
  int64_t off = 0;
  L.c = off;
  off += cbytes;
  L.cy = off;
  off += cbytes;
  L.cz = same ? L.cy : off;
  if (!same) off += cbytes;
  for (int k = 0; k < 5; ++k) {
    if ((same && k == 2) || (!pair && k == 4)) {
      L.F[k] = L.F[same && k == 2 ? 1 : 3];
      continue;
    }
    L.F[k] = off;
    off += fbytes;
  }
  const int C = hess_cols_c(pp->nz);
  const int64_t ntiles = nang * pp->ndist * (pp->n / C);
  L.part = off;
  off += align256(ntiles * 8 * (pair ? 2 : 1));
  L.part2 = off;
  off += align256(kSumBlocks * 8 * (pair ? 2 : 1));
  L.total = off;
  return L;
}

// Row pass: 8 elements per thread as the gradient's row passes; the staged
// rows of a workgroup stay within 64 KiB of dynamic LDS.
constexpr int64_t kHessDynBudget = 65536;

template <int N, typename TO, int MODE>
static hipError_t hess_rows_m(const BhhArgs& a, hipStream_t st) {
  constexpr int E = N >= 8 ? 8 : N;
  constexpr int T = N / E, ROWS = kRowThreads / T;
  const int64_t per_row = (int64_t)HessShape<MODE>::NR * a.npsi * (int64_t)sizeof(TO);
  int rpb = ROWS;
  while (rpb > 1 && rpb * per_row > kHessDynBudget) rpb >>= 1;
  if (rpb * per_row > kHessDynBudget) return hipErrorInvalidConfiguration;
  const int64_t grid = ((int64_t)a.ndist * a.nz + rpb - 1) / rpb;
  hipLaunchKernelGGL((bhh_chain_k<N, E, TO, MODE>), dim3((unsigned)grid), dim3(rpb * T), (size_t)(rpb * per_row),
                     st, a, rpb);
  return hipGetLastError();
}

template <int N, typename TO>
static hipError_t hess_rows(int mode, const BhhArgs& a, hipStream_t st) {
  if (mode == kSame) return hess_rows_m<N, TO, kSame>(a, st);
  if (mode == kPair) return hess_rows_m<N, TO, kPair>(a, st);
  return hess_rows_m<N, TO, kGeneral>(a, st);
}

template <typename TO>
static hipError_t hess_rows_dispatch(int mode, const BhhArgs& a, hipStream_t st) {
  switch (a.n) {
    case 8: return hess_rows<8, TO>(mode, a, st);
    case 16: return hess_rows<16, TO>(mode, a, st);
    case 32: return hess_rows<32, TO>(mode, a, st);
    case 64: return hess_rows<64, TO>(mode, a, st);
    case 128: return hess_rows<128, TO>(mode, a, st);
    case 256: return hess_rows<256, TO>(mode, a, st);
    case 512: return hess_rows<512, TO>(mode, a, st);
    case 1024: return hess_rows<1024, TO>(mode, a, st);
    case 2048: return hess_rows<2048, TO>(mode, a, st);
    default: return hipErrorInvalidValue;
  }
}

template <int NZ>
static hipError_t hess_cols(int mode, const BhhArgs& a, hipStream_t st) {
  constexpr int E = bhh_col_e<NZ>(), C = bhh_cols_c<NZ>();
  if (a.n % C) return hipErrorInvalidValue;
  const int tpf = a.n / C;
  const dim3 grid((unsigned)(a.nang * a.ndist * tpf)), block(C * (NZ / E));
  if (mode == kSame)
    hipLaunchKernelGGL((bhh_cols_k<NZ, E, C, kSame>), grid, block, 0, st, a, tpf);
  else if (mode == kPair)
    hipLaunchKernelGGL((bhh_cols_k<NZ, E, C, kPair>), grid, block, 0, st, a, tpf);
  else
    hipLaunchKernelGGL((bhh_cols_k<NZ, E, C, kGeneral>), grid, block, 0, st, a, tpf);
  return hipGetLastError();
}

static hipError_t hess_cols_dispatch(int mode, const BhhArgs& a, hipStream_t st) {
  switch (a.nz) {
    case 8: return hess_cols<8>(mode, a, st);
    case 16: return hess_cols<16>(mode, a, st);
    case 32: return hess_cols<32>(mode, a, st);
    case 64: return hess_cols<64>(mode, a, st);
    case 128: return hess_cols<128>(mode, a, st);
    case 256: return hess_cols<256>(mode, a, st);
    case 512: return hess_cols<512>(mode, a, st);
    case 1024: return hess_cols<1024>(mode, a, st);
    case 2048: return hess_cols<2048>(mode, a, st);
    default: return hipErrorInvalidValue;
  }
}

template <typename TO>
static hipError_t bhh_run(const htamd_shift_plan* sp, int mode, BhhArgs a, const void* proj, const void* projy,
                          const void* projz, char* ws, const BhhLayout& L, double scale, int accumulate,
                          double* result, hipStream_t st) {
  hipError_t e;
  if ((e = bs_launch_prefilter<TO>((const TO*)proj, (TO*)(ws + L.c), a.nang, a.nzpsi, a.npsi, sp->taps, st)) !=
      hipSuccess)
    return e;
  if ((e = bs_launch_prefilter<TO>((const TO*)projy, (TO*)(ws + L.cy), a.nang, a.nzpsi, a.npsi, sp->taps, st)) !=
      hipSuccess)
    return e;
  if (mode != kSame &&
      (e = bs_launch_prefilter<TO>((const TO*)projz, (TO*)(ws + L.cz), a.nang, a.nzpsi, a.npsi, sp->taps, st)) !=
          hipSuccess)
    return e;
  if ((e = hess_rows_dispatch<TO>(mode, a, st)) != hipSuccess) return e;
  // D_x of the NF input fields, one strided batch per distance: the used field
  // slots are consecutive (bhh_layout), field (k, b, j) at F[0] + ((k nang + b)
  // ndist + j) nz n
  const int nf = mode == kSame ? 3 : mode == kPair ? 5 : 4;
  const int64_t field = (int64_t)a.n * a.nz;
  const int64_t slot = a.nang * a.ndist * field;
  const int used[5] = {0, 1, mode == kSame ? 3 : 2, 3, 4};
  for (int k = 0; k < nf; ++k)
    if (a.F[used[k]] != a.F[0] + k * slot) return hipErrorInvalidValue;  // slots not consecutive
  for (int j = 0; j < a.ndist; ++j) {
    float2* f0 = a.F[0] + j * field;
    if ((e = launch_fresnel_rows(a.n, a.nz, f0, a.ndist * field, f0, a.ndist * field, nf * a.nang, a.twx,
                                 a.tabx + (int64_t)(j * 2 + 0) * 3 * (a.n / 2), st)) != hipSuccess)
      return e;
  }
  if ((e = hess_cols_dispatch(mode, a, st)) != hipSuccess) return e;
  const int64_t ntiles = a.nang * a.ndist * (a.n / hess_cols_c(a.nz));
  const int nout = mode == kPair ? 2 : 1;
  const int64_t chunk = (ntiles + kSumBlocks - 1) / kSumBlocks;
  const int64_t nb = (ntiles + chunk - 1) / chunk;  // <= kSumBlocks
  double* part2 = (double*)(ws + L.part2);
  hipLaunchKernelGGL(bhh_sum1_k, dim3((unsigned)nb), dim3(256), 0, st, a.part, ntiles, nout, chunk, part2);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(bhh_sum_k, dim3(1), dim3(256), 0, st, part2, nb, nout, scale, accumulate, result);
  return hipGetLastError();
}

}  // namespace htamd

using namespace htamd;

extern "C" {

int htamd_bh_hessian_workspace_size(const htamd_bh_plan* p, int64_t nangles, int same_directions,
                                    int64_t* bytes) {
  if (!p || !bytes || nangles < 0 || same_directions < 0 || same_directions > 2)
    return fail(HTAMD_ERR_ARG, "bh_hessian_workspace_size: bad arguments");
  *bytes = bhh_layout(p->prop, p->shift, nangles, same_directions).total;
  return HTAMD_OK;
}

static int bh_hessian_impl(const htamd_bh_plan* p, int mode, int64_t nangles, const void* proj, const float* pos,
                           const float* mag, const void* prb, const void* proj_y, const float* pos_y,
                           const void* prb_y, const void* proj_z, const float* pos_z, const void* prb_z,
                           const float* data, double scale, double* result, int accumulate, void* workspace,
                           int64_t workspace_bytes, void* stream, const char* what) {
  const int nout = mode == kPair ? 2 : 1;
  if (nangles == 0) {
    if (!accumulate) {
      hipLaunchKernelGGL(bhh_sum_k, dim3(1), dim3(256), 0, (hipStream_t)stream, (const double*)nullptr, 0, nout,
                         0.0, 0, result);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return hip_fail(e, what);
    }
    return HTAMD_OK;
  }
  if (!proj || !pos || !mag || !prb || !proj_y || !pos_y || !prb_y || !data)
    return fail(HTAMD_ERR_ARG, std::string(what) + ": NULL buffer");
  const BhhLayout L = bhh_layout(p->prop, p->shift, nangles, mode);
  if (!workspace || workspace_bytes < L.total)
    return fail(HTAMD_ERR_ARG, std::string(what) + ": workspace of " + std::to_string(L.total) + " bytes required");
  int dev = -1;
  (void)hipGetDevice(&dev);
  if (dev != p->prop->device) return fail(HTAMD_ERR_ARG, std::string(what) + ": plan belongs to another device");
  const bool same = mode == kSame;
  char* ws = (char*)workspace;
  BhhArgs a;
  a.n = (int)p->prop->n;
  a.nz = (int)p->prop->nz;
  a.npsi = (int)p->shift->npsi;
  a.nzpsi = (int)p->shift->nzpsi;
  a.ndist = (int)p->prop->ndist;
  a.nang = nangles;
  a.c = ws + L.c;
  a.cy = ws + L.cy;
  a.cz = ws + L.cz;
  a.pos = pos;
  a.dy = pos_y;
  a.dz = same ? pos_y : pos_z;
  a.mag = mag;
  a.prb = (const float2*)prb;
  a.py = (const float2*)prb_y;
  a.pz = (const float2*)(same ? prb_y : prb_z);
  a.data = data;
  static_assert(sizeof(a.F) / sizeof(a.F[0]) == sizeof(L.F) / sizeof(L.F[0]), "one pointer per field slot");
  for (int k = 0; k < 5; ++k) a.F[k] = (float2*)(ws + L.F[k]);
  a.part = (double*)(ws + L.part);
  a.twx = p->prop->tw_x;
  a.twy = p->prop->tw_y;
  a.tabx = p->prop->tab_x;
  a.taby = p->prop->tab_y;
  hipError_t e = p->shift->is_complex
      ? bhh_run<float2>(p->shift, mode, a, proj, proj_y, proj_z, ws, L, scale, accumulate, result,
                        (hipStream_t)stream)
      : bhh_run<float>(p->shift, mode, a, proj, proj_y, proj_z, ws, L, scale, accumulate, result,
                       (hipStream_t)stream);
  if (e == hipErrorInvalidConfiguration)
    return fail(HTAMD_ERR_UNSUPPORTED, std::string(what) + ": the staged object rows of one row pass exceed the LDS");
  return e == hipSuccess ? HTAMD_OK : hip_fail(e, what);
}

int htamd_bh_hessian(const htamd_bh_plan* p, int64_t nangles, const void* proj, const float* pos, const float* mag,
                     const void* prb, const void* proj_y, const float* pos_y, const void* prb_y, const void* proj_z,
                     const float* pos_z, const void* prb_z, const float* data, double scale, double* result,
                     int accumulate, void* workspace, int64_t workspace_bytes, void* stream) {
  if (!p) return fail(HTAMD_ERR_ARG, "bh_hessian: plan is NULL");
  if (nangles < 0 || !result) return fail(HTAMD_ERR_ARG, "bh_hessian: bad arguments");
  const bool same = proj_z == nullptr && pos_z == nullptr && prb_z == nullptr;
  if (!same && (!proj_z || !pos_z || !prb_z))
    return fail(HTAMD_ERR_ARG, "bh_hessian: give all three parts of direction z, or none (z = y)");
  return bh_hessian_impl(p, same ? kSame : kGeneral, nangles, proj, pos, mag, prb, proj_y, pos_y, prb_y, proj_z,
                         pos_z, prb_z, data, scale, result, accumulate, workspace, workspace_bytes, stream,
                         "bh_hessian");
}

int htamd_bh_hessian_pair(const htamd_bh_plan* p, int64_t nangles, const void* proj, const float* pos,
                          const float* mag, const void* prb, const void* proj_y, const float* pos_y,
                          const void* prb_y, const void* proj_z, const float* pos_z, const void* prb_z,
                          const float* data, double scale, double* result, int accumulate, void* workspace,
                          int64_t workspace_bytes, void* stream) {
  if (!p) return fail(HTAMD_ERR_ARG, "bh_hessian_pair: plan is NULL");
  if (nangles < 0 || !result || (nangles > 0 && (!proj_z || !pos_z || !prb_z)))
    return fail(HTAMD_ERR_ARG, "bh_hessian_pair: bad arguments");
  return bh_hessian_impl(p, kPair, nangles, proj, pos, mag, prb, proj_y, pos_y, prb_y, proj_z, pos_z, prb_z, data,
                         scale, result, accumulate, workspace, workspace_bytes, stream, "bh_hessian_pair");
}

}  // extern "C"
