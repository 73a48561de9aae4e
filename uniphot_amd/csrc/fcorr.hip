// Correlated-noise variance factor F_corr of the post-fit error model
// (tract7dt/fit/grid_gpu/error_model.py:933-1113 raw-sky estimator,
// :1116-1302 residual estimator).  The device passes run over "maps": 2-D
// float64 images packed back to back (map k starts at off[k], h x w row-major),
// each with a byte mask.  A raw-sky map is a band's central crop; a residual
// map is one (grid, band) window of the confusion pass; a stamp map is a PSF
// stamp zero-padded to 2T x 2T, whose circular autocorrelation at lags below T
// is the linear one.
//
//  fc_load_kernel       v = float64(src), g0 = finite(v) & good
//  fc_hist/scan_kernel  exact k-th smallest of v (or |v - c|) over g0 by an
//                       8-pass, 8-bit radix select on order-preserving keys:
//                       the median and the MAD with numpy's even-count rule
//                       taken on the host from two ranks
//  fc_clean_kernel      struct = !g0 | outlier(v); clean = no struct pixel
//                       within |dy| + |dx| <= 2 (two dilations by the
//                       4-neighbour cross, outside pixels not struct)
//  fc_tile_stats_kernel per 256-px tile: masked count, sum, sum of squared
//                       deviations from the tile mean
//  fc_mask_kernel       x = keep(tile) & clean ? v - mean(tile) : 0, m likewise
//  fc_autocov_kernel    A(d) = sum_i x_i x_{i+d}, K(d) = sum_i m_i m_{i+d}
//                       (circular) for dy in [0, L], dx in [-L, L] (dy = 0:
//                       dx >= 0; the other half follows from A(-d) = A(d)):
//                       8-row x 64-column pixel tiles staged in LDS with their
//                       circular halo, each thread holding 8 adjacent lags of
//                       up to two lag runs; per-workgroup partial sums reduced
//                       in a fixed order by fc_autocov_reduce_kernel.
// A is float64; the mask products are 0/1 and each thread's partial count
// stays far below 2^24, so K accumulates exactly in float32.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"

namespace uph {

constexpr int kFcRows = 8, kFcCols = 64, kFcLagRun = 8, kFcXu = 8, kFcBlock = 512;
constexpr int kFcMaxL = 56;

__device__ __forceinline__ uint64_t fc_key(double d) {
  const uint64_t u = (uint64_t)__double_as_longlong(d);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double fc_unkey(uint64_t k) {
  const uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)u);
}

__global__ __launch_bounds__(UPH_BLOCK) void fc_fill_kernel(uint32_t* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)UPH_BLOCK + threadIdx.x; i < n;
       i += (size_t)gridDim.x * UPH_BLOCK)
    p[i] = 0u;
}

__global__ __launch_bounds__(UPH_BLOCK) void fc_load_kernel(
    const int64_t* map_off, const int32_t* map_hw, const float* src, const int64_t* src_off,
    const int32_t* src_stride, const uint8_t* good_u8, const float* good_f32, double* v,
    uint8_t* g0, unsigned long long* n_good) {
  const int k = blockIdx.y;
  const int h = map_hw[2 * k], w = map_hw[2 * k + 1];
  const int64_t n = (int64_t)h * w, o = map_off[k], so = src_off[k];
  const int st = src_stride[k];
  unsigned cnt = 0;
  for (int64_t e = blockIdx.x * (int64_t)UPH_BLOCK + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * UPH_BLOCK) {
    const int64_t y = e / w, x = e % w, si = so + y * st + x;
    const double d = (double)src[si];
    const bool gd = good_u8 ? good_u8[si] != 0 : good_f32[si] > 0.5f;
    const bool ok = gd && isfinite(d);
    v[o + e] = d;
    g0[o + e] = ok ? 1 : 0;
    cnt += ok ? 1u : 0u;
  }
  cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(n_good + k, (unsigned long long)cnt);
}

// Query q selects rank state[2q + 1] of map q_map[q]: its keys are those of v
// (q_center NaN) or of |v - q_center|.  state[2q] holds the key prefix.
__global__ __launch_bounds__(UPH_BLOCK) void fc_hist_kernel(
    const int32_t* q_map, const double* q_center, const int64_t* map_off, const int32_t* map_hw,
    const double* v, const uint8_t* g0, const uint64_t* state, uint32_t* hist, int shift) {
  __shared__ uint32_t h[256];
  const int q = blockIdx.y, k = q_map[q];
  for (int i = threadIdx.x; i < 256; i += UPH_BLOCK) h[i] = 0u;
  __syncthreads();
  const int64_t n = (int64_t)map_hw[2 * k] * map_hw[2 * k + 1], o = map_off[k];
  const double c = q_center[q];
  const bool tr = !isnan(c);
  const uint64_t prefix = state[2 * q];
  const uint64_t hi = shift >= 56 ? 0ull : (~0ull << (shift + 8));
  for (int64_t e = blockIdx.x * (int64_t)UPH_BLOCK + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * UPH_BLOCK) {
    if (!g0[o + e]) continue;
    const double d = v[o + e];
    const uint64_t key = fc_key(tr ? fabs(d - c) : d);
    if (((key ^ prefix) & hi) == 0ull) atomicAdd(&h[(key >> shift) & 255u], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += UPH_BLOCK)
    if (h[i]) atomicAdd(hist + (size_t)q * 256 + i, h[i]);
}

__global__ __launch_bounds__(64) void fc_scan_kernel(int n_q, uint64_t* state, uint32_t* hist,
                                                     int shift) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  if (q >= n_q) return;
  uint32_t* hq = hist + (size_t)q * 256;
  int64_t r = (int64_t)state[2 * q + 1];
  int digit = 255;
  for (int i = 0; i < 256; ++i) {
    const int64_t c = hq[i];
    if (r < c) {
      digit = i;
      break;
    }
    r -= c;
  }
  for (int i = 0; i < 256; ++i) hq[i] = 0u;
  state[2 * q] |= (uint64_t)digit << shift;
  state[2 * q + 1] = (uint64_t)r;
}

__global__ __launch_bounds__(64) void fc_decode_kernel(int n_q, const uint64_t* state,
                                                       double* out) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  if (q < n_q) out[q] = fc_unkey(state[2 * q]);
}

__global__ __launch_bounds__(64) void fc_init_state_kernel(int n_q, const int64_t* rank,
                                                           uint64_t* state) {
  const int q = blockIdx.x * 64 + threadIdx.x;
  if (q >= n_q) return;
  state[2 * q] = 0ull;
  state[2 * q + 1] = (uint64_t)rank[q];
}

// outlier test of the raw-sky estimator (v > med + 3 mad | v < med - 5 mad) or
// of the residual one (|v - med| > 3 mad), written as the reference writes them
__device__ __forceinline__ bool fc_struct(const double* v, const uint8_t* g0, int64_t i, double med,
                                          double hi, double lo, double cut, bool sym) {
  if (!g0[i]) return true;
  const double d = v[i];
  return sym ? (fabs(d - med) > cut) : (d > hi || d < lo);
}

__global__ __launch_bounds__(UPH_BLOCK) void fc_clean_kernel(
    const int64_t* map_off, const int32_t* map_hw, const double* v, const uint8_t* g0,
    const double* med, const double* mad, double clip, double low, int sym, uint8_t* clean) {
  const int k = blockIdx.y;
  const int h = map_hw[2 * k], w = map_hw[2 * k + 1];
  const int64_t n = (int64_t)h * w, o = map_off[k];
  const double md = med[k], ma = mad[k];
  const double hi = md + clip * ma, lo = md - low * ma, cut = clip * ma;
  for (int64_t e = blockIdx.x * (int64_t)UPH_BLOCK + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * UPH_BLOCK) {
    const int y = (int)(e / w), x = (int)(e % w);
    bool s = false;
#pragma unroll
    for (int dy = -2; dy <= 2; ++dy) {
      const int yy = y + dy;
      if (yy < 0 || yy >= h) continue;
      const int r = 2 - (dy < 0 ? -dy : dy);
      for (int dx = -r; dx <= r; ++dx) {
        const int xx = x + dx;
        if (xx < 0 || xx >= w) continue;
        s = s || fc_struct(v, g0, o + (int64_t)yy * w + xx, md, hi, lo, cut, sym != 0);
      }
    }
    clean[o + e] = s ? 0 : 1;
  }
}

// one workgroup per tile: (count, sum, sum of squared deviations) of the
// masked pixels of a ts x ts tile (clipped at the map edge)
__global__ __launch_bounds__(UPH_BLOCK) void fc_tile_stats_kernel(
    const int32_t* tile_map, const int32_t* tile_yx, int ts, const int64_t* map_off,
    const int32_t* map_hw, const double* v, const uint8_t* mask, double* stats) {
  __shared__ double red[UPH_BLOCK / 64];
  __shared__ double cnt_s[UPH_BLOCK / 64];
  const int t = blockIdx.x, k = tile_map[t];
  const int w = map_hw[2 * k + 1], h = map_hw[2 * k];
  const int y0 = tile_yx[2 * t], x0 = tile_yx[2 * t + 1];
  const int th = min(ts, h - y0), tw = min(ts, w - x0);
  const int64_t o = map_off[k];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double s = 0.0, c = 0.0;
  for (int e = threadIdx.x; e < th * tw; e += UPH_BLOCK) {
    const int64_t i = o + (int64_t)(y0 + e / tw) * w + x0 + e % tw;
    if (mask[i]) {
      s += v[i];
      c += 1.0;
    }
  }
  s = wave_sum(s);
  c = wave_sum(c);
  if (lane == 0) {
    red[wv] = s;
    cnt_s[wv] = c;
  }
  __syncthreads();
  double S = 0.0, C = 0.0;
  for (int i = 0; i < UPH_BLOCK / 64; ++i) {
    S += red[i];
    C += cnt_s[i];
  }
  __syncthreads();
  const double mean = C > 0.0 ? S / C : 0.0;
  double m2 = 0.0;
  for (int e = threadIdx.x; e < th * tw; e += UPH_BLOCK) {
    const int64_t i = o + (int64_t)(y0 + e / tw) * w + x0 + e % tw;
    if (mask[i]) {
      const double d = v[i] - mean;
      m2 = fma(d, d, m2);
    }
  }
  m2 = wave_sum(m2);
  if (lane == 0) red[wv] = m2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double M2 = 0.0;
    for (int i = 0; i < UPH_BLOCK / 64; ++i) M2 += red[i];
    stats[3 * t] = C;
    stats[3 * t + 1] = S;
    stats[3 * t + 2] = M2;
  }
}

__global__ __launch_bounds__(UPH_BLOCK) void fc_mask_kernel(
    const int64_t* map_off, const int32_t* map_hw, const int32_t* tile_base, int ts,
    const uint8_t* keep, const double* tile_mean, double* v, uint8_t* mask) {
  const int k = blockIdx.y;
  const int h = map_hw[2 * k], w = map_hw[2 * k + 1];
  const int ntx = (w + ts - 1) / ts;
  const int64_t n = (int64_t)h * w, o = map_off[k];
  for (int64_t e = blockIdx.x * (int64_t)UPH_BLOCK + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * UPH_BLOCK) {
    const int y = (int)(e / w), x = (int)(e % w);
    const int t = tile_base[k] + (y / ts) * ntx + x / ts;
    const bool m = mask[o + e] && keep[t];
    v[o + e] = m ? v[o + e] - tile_mean[t] : 0.0;
    mask[o + e] = m ? 1 : 0;
  }
}

__host__ __device__ inline int fc_nr0(int L) { return (L + 1 + kFcLagRun - 1) / kFcLagRun; }
__host__ __device__ inline int fc_nr1(int L) { return (2 * L + 1 + kFcLagRun - 1) / kFcLagRun; }
__host__ __device__ inline int fc_nlag(int L) { return (L + 1) + L * (2 * L + 1); }
// LDS row stride of the staged rows: at least 64 + 2L + 8 columns, rounded up to 1
// mod 32 doubles so that lanes reading the same column of adjacent rows (adjacent
// row lags) hit distinct banks
__host__ __device__ inline int fc_sw(int L) {
  const int need = kFcCols + 2 * L + kFcLagRun;
  return (need + 30) / 32 * 32 + 1;
}
__host__ __device__ inline size_t fc_lds_bytes(int L) {
  return (size_t)(kFcRows + L) * fc_sw(L) * (sizeof(double) + sizeof(float)) +
         (size_t)kFcRows * kFcCols * (sizeof(double) + sizeof(float));
}

// grid (bpm, n_maps); partial[(k * bpm + b) * 2 * nlag + {0, nlag} + lag]
__global__ __launch_bounds__(kFcBlock) void fc_autocov_kernel(const int64_t* map_off,
                                                              const int32_t* map_hw, int L,
                                                              const double* x, const uint8_t* m,
                                                              double* partial) {
  extern __shared__ double fc_lds[];
  const int SW = fc_sw(L), SR = kFcRows + L;
  double* BX = fc_lds;
  float* BM = reinterpret_cast<float*>(BX + (size_t)SR * SW);
  double* AX = reinterpret_cast<double*>(BM + (size_t)SR * SW);
  float* AM = reinterpret_cast<float*>(AX + kFcRows * kFcCols);
  const int k = blockIdx.y, bpm = gridDim.x;
  const int h = map_hw[2 * k], w = map_hw[2 * k + 1];
  const int64_t o = map_off[k];
  const int nr0 = fc_nr0(L), nr1 = fc_nr1(L), nruns = nr0 + L * nr1, nlag = fc_nlag(L);
  const int ntx = (w + kFcCols - 1) / kFcCols, nty = (h + kFcRows - 1) / kFcRows;
  // this thread's lag runs (up to two): dy, first dx.  Adjacent runs past the
  // dy = 0 row step through dy first, so a wave's lanes read one column of
  // adjacent staged rows
  int dy[2], dx0[2];
  bool act[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int r = threadIdx.x + j * kFcBlock;
    act[j] = r < nruns;
    if (r < nr0) {
      dy[j] = 0;
      dx0[j] = r * kFcLagRun;
    } else {
      const int rr = r - nr0;
      dy[j] = act[j] ? 1 + rr % L : 0;
      dx0[j] = act[j] ? -L + (rr / L) * kFcLagRun : 0;
    }
  }
  double accA[2][kFcLagRun];
  float accK[2][kFcLagRun];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int q = 0; q < kFcLagRun; ++q) {
      accA[j][q] = 0.0;
      accK[j][q] = 0.0f;
    }
  for (int tile = blockIdx.x; tile < ntx * nty; tile += bpm) {
    const int ty0 = (tile / ntx) * kFcRows, tx0 = (tile % ntx) * kFcCols;
    __syncthreads();
    // b region: rows ty0 .. ty0 + SR - 1, columns tx0 - L .. tx0 - L + SW - 1, circular
    for (int e = threadIdx.x; e < SR * SW; e += kFcBlock) {
      const int r = e / SW, c = e % SW;
      const int gy = (ty0 + r) % h;
      int gx = (tx0 - L + c) % w;
      if (gx < 0) gx += w;
      const int64_t i = o + (int64_t)gy * w + gx;
      BX[e] = x[i];
      BM[e] = m[i] ? 1.0f : 0.0f;
    }
    // a region: the tile itself, zero outside the map
    for (int e = threadIdx.x; e < kFcRows * kFcCols; e += kFcBlock) {
      const int r = e / kFcCols, c = e % kFcCols;
      const bool in = ty0 + r < h && tx0 + c < w;
      const int64_t i = o + (int64_t)(ty0 + r) * w + tx0 + c;
      AX[e] = in ? x[i] : 0.0;
      AM[e] = (in && m[i]) ? 1.0f : 0.0f;
    }
    __syncthreads();
    for (int y = 0; y < kFcRows; ++y) {
      const double* arow = AX + y * kFcCols;
      const float* amrow = AM + y * kFcCols;
      for (int xc = 0; xc < kFcCols; xc += kFcXu) {
        double a[kFcXu];
        float am[kFcXu];
#pragma unroll
        for (int i = 0; i < kFcXu; ++i) {  // one address across the wave: broadcast
          a[i] = arow[xc + i];
          am[i] = amrow[xc + i];
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (!act[j]) continue;
          const int bo = (y + dy[j]) * SW + L + dx0[j] + xc;
          double b[kFcXu + kFcLagRun - 1];
          float bm[kFcXu + kFcLagRun - 1];
#pragma unroll
          for (int i = 0; i < kFcXu + kFcLagRun - 1; ++i) {
            b[i] = BX[bo + i];
            bm[i] = BM[bo + i];
          }
#pragma unroll
          for (int i = 0; i < kFcXu; ++i)
#pragma unroll
            for (int q = 0; q < kFcLagRun; ++q) {
              accA[j][q] = fma(a[i], b[i + q], accA[j][q]);
              accK[j][q] = fmaf(am[i], bm[i + q], accK[j][q]);
            }
        }
      }
    }
  }
  double* P = partial + ((size_t)k * bpm + blockIdx.x) * 2 * nlag;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    if (!act[j]) continue;
#pragma unroll
    for (int q = 0; q < kFcLagRun; ++q) {
      const int dx = dx0[j] + q;
      if (dx > L) continue;
      const int lag = dy[j] == 0 ? dx : (L + 1) + (dy[j] - 1) * (2 * L + 1) + dx + L;
      P[lag] = accA[j][q];
      P[nlag + lag] = (double)accK[j][q];
    }
  }
}

__global__ __launch_bounds__(UPH_BLOCK) void fc_autocov_reduce_kernel(int bpm, int nlag,
                                                                      const double* partial,
                                                                      double* out) {
  const int k = blockIdx.y;
  const int j = blockIdx.x * UPH_BLOCK + threadIdx.x;
  if (j >= 2 * nlag) return;
  const double* P = partial + (size_t)k * bpm * 2 * nlag + j;
  double s = 0.0;
  for (int b = 0; b < bpm; ++b) s += P[(size_t)b * 2 * nlag];
  out[(size_t)k * 2 * nlag + j] = s;
}

inline unsigned fc_blocks(int64_t n, int cap) {
  const int64_t b = (n + UPH_BLOCK * 16 - 1) / (UPH_BLOCK * 16);
  return (unsigned)(b < 1 ? 1 : (b > cap ? cap : b));
}

}  // namespace uph

extern "C" int uph_fcorr_load(int n_maps, int64_t max_pix, const int64_t* map_off,
                              const int32_t* map_hw, const float* src, const int64_t* src_off,
                              const int32_t* src_stride, const uint8_t* good_u8,
                              const float* good_f32, double* v, uint8_t* g0, int64_t* n_good,
                              void* stream) {
  UPH_CHECK_ARG(n_maps >= 0 && n_maps <= 65535 && max_pix >= 0,
                "uph_fcorr_load: n_maps=%d max_pix=%lld", n_maps, (long long)max_pix);
  if (n_maps == 0) return UPH_OK;
  UPH_CHECK_ARG(map_off && map_hw && src && src_off && src_stride && v && g0 && n_good &&
                    ((good_u8 != nullptr) != (good_f32 != nullptr)),
                "uph_fcorr_load: null argument (exactly one of good_u8 / good_f32)");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t nw = (size_t)n_maps * 2;  // int64 counters as 32-bit words
  hipLaunchKernelGGL(uph::fc_fill_kernel, dim3(1), dim3(UPH_BLOCK), 0, s,
                     reinterpret_cast<uint32_t*>(n_good), nw);
  hipLaunchKernelGGL(uph::fc_load_kernel, dim3(uph::fc_blocks(max_pix, 256), n_maps),
                     dim3(UPH_BLOCK), 0, s, map_off, map_hw, src, src_off, src_stride, good_u8,
                     good_f32, v, g0, reinterpret_cast<unsigned long long*>(n_good));
  return uph::hip_status(hipGetLastError(), "uph_fcorr_load launch");
}

extern "C" size_t uph_fcorr_select_work_words(int n_q) {
  return (size_t)(n_q > 0 ? n_q : 0) * (256 + 4);
}

extern "C" int uph_fcorr_select(int n_q, int64_t max_pix, const int32_t* q_map,
                                const int64_t* q_rank, const double* q_center,
                                const int64_t* map_off, const int32_t* map_hw, const double* v,
                                const uint8_t* g0, uint32_t* work, size_t work_words,
                                double* out, void* stream) {
  UPH_CHECK_ARG(n_q >= 0 && n_q <= 65535 && max_pix >= 0, "uph_fcorr_select: n_q=%d", n_q);
  if (n_q == 0) return UPH_OK;
  UPH_CHECK_ARG(q_map && q_rank && q_center && map_off && map_hw && v && g0 && work && out,
                "uph_fcorr_select: null argument");
  UPH_CHECK_ARG(work_words >= uph_fcorr_select_work_words(n_q),
                "uph_fcorr_select: work of %zu words < %zu", work_words,
                uph_fcorr_select_work_words(n_q));
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint64_t* state = reinterpret_cast<uint64_t*>(work);
  uint32_t* hist = work + (size_t)n_q * 4;
  const unsigned qb = (unsigned)((n_q + 63) / 64);
  hipLaunchKernelGGL(uph::fc_fill_kernel, dim3(64), dim3(UPH_BLOCK), 0, s, hist,
                     (size_t)n_q * 256);
  hipLaunchKernelGGL(uph::fc_init_state_kernel, dim3(qb), dim3(64), 0, s, n_q, q_rank, state);
  const unsigned hb = uph::fc_blocks(max_pix, 64);
  for (int shift = 56; shift >= 0; shift -= 8) {
    hipLaunchKernelGGL(uph::fc_hist_kernel, dim3(hb, n_q), dim3(UPH_BLOCK), 0, s, q_map,
                       q_center, map_off, map_hw, v, g0, state, hist, shift);
    hipLaunchKernelGGL(uph::fc_scan_kernel, dim3(qb), dim3(64), 0, s, n_q, state, hist, shift);
  }
  hipLaunchKernelGGL(uph::fc_decode_kernel, dim3(qb), dim3(64), 0, s, n_q, state, out);
  return uph::hip_status(hipGetLastError(), "uph_fcorr_select launch");
}

extern "C" int uph_fcorr_clean(int n_maps, int64_t max_pix, const int64_t* map_off,
                               const int32_t* map_hw, const double* v, const uint8_t* g0,
                               const double* med, const double* mad, double clip, double low,
                               int symmetric, uint8_t* clean, void* stream) {
  UPH_CHECK_ARG(n_maps >= 0 && n_maps <= 65535 && max_pix >= 0, "uph_fcorr_clean: n_maps=%d",
                n_maps);
  if (n_maps == 0) return UPH_OK;
  UPH_CHECK_ARG(map_off && map_hw && v && g0 && med && mad && clean,
                "uph_fcorr_clean: null argument");
  hipLaunchKernelGGL(uph::fc_clean_kernel, dim3(uph::fc_blocks(max_pix, 256), n_maps),
                     dim3(UPH_BLOCK), 0, static_cast<hipStream_t>(stream), map_off, map_hw, v,
                     g0, med, mad, clip, low, symmetric, clean);
  return uph::hip_status(hipGetLastError(), "uph_fcorr_clean launch");
}

extern "C" int uph_fcorr_tile_stats(int n_tiles, const int32_t* tile_map, const int32_t* tile_yx,
                                    int ts, const int64_t* map_off, const int32_t* map_hw,
                                    const double* v, const uint8_t* mask, double* stats,
                                    void* stream) {
  UPH_CHECK_ARG(n_tiles >= 0 && ts >= 1, "uph_fcorr_tile_stats: n_tiles=%d ts=%d", n_tiles, ts);
  if (n_tiles == 0) return UPH_OK;
  UPH_CHECK_ARG(tile_map && tile_yx && map_off && map_hw && v && mask && stats,
                "uph_fcorr_tile_stats: null argument");
  hipLaunchKernelGGL(uph::fc_tile_stats_kernel, dim3(n_tiles), dim3(UPH_BLOCK), 0,
                     static_cast<hipStream_t>(stream), tile_map, tile_yx, ts, map_off, map_hw, v,
                     mask, stats);
  return uph::hip_status(hipGetLastError(), "uph_fcorr_tile_stats launch");
}

extern "C" int uph_fcorr_mask(int n_maps, int64_t max_pix, const int64_t* map_off,
                              const int32_t* map_hw, const int32_t* tile_base, int ts,
                              const uint8_t* keep, const double* tile_mean, double* v,
                              uint8_t* mask, void* stream) {
  UPH_CHECK_ARG(n_maps >= 0 && n_maps <= 65535 && ts >= 1, "uph_fcorr_mask: n_maps=%d ts=%d",
                n_maps, ts);
  if (n_maps == 0) return UPH_OK;
  UPH_CHECK_ARG(map_off && map_hw && tile_base && keep && tile_mean && v && mask,
                "uph_fcorr_mask: null argument");
  hipLaunchKernelGGL(uph::fc_mask_kernel, dim3(uph::fc_blocks(max_pix, 256), n_maps),
                     dim3(UPH_BLOCK), 0, static_cast<hipStream_t>(stream), map_off, map_hw,
                     tile_base, ts, keep, tile_mean, v, mask);
  return uph::hip_status(hipGetLastError(), "uph_fcorr_mask launch");
}

extern "C" int uph_fcorr_nlag(int L) { return L >= 0 ? uph::fc_nlag(L) : 0; }

extern "C" int uph_fcorr_autocov(int n_maps, int bpm, int L, const int64_t* map_off,
                                 const int32_t* map_hw, const double* x, const uint8_t* m,
                                 double* partial, size_t partial_doubles, double* out,
                                 void* stream) {
  UPH_CHECK_ARG(n_maps >= 0 && n_maps <= 65535 && bpm >= 1 && bpm <= 65535 && L >= 0 &&
                    L <= uph::kFcMaxL,
                "uph_fcorr_autocov: n_maps=%d bpm=%d L=%d (L at most %d)", n_maps, bpm, L,
                uph::kFcMaxL);
  if (n_maps == 0) return UPH_OK;
  UPH_CHECK_ARG(map_off && map_hw && x && m && partial && out,
                "uph_fcorr_autocov: null argument");
  const int nlag = uph::fc_nlag(L);
  const size_t need = (size_t)n_maps * bpm * 2 * nlag;
  UPH_CHECK_ARG(partial_doubles >= need, "uph_fcorr_autocov: partial of %zu doubles < %zu",
                partial_doubles, need);
  const size_t lds = uph::fc_lds_bytes(L);
  UPH_CHECK_ARG(lds <= 160 * 1024, "uph_fcorr_autocov: L=%d needs %zu B of LDS", L, lds);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(uph::fc_autocov_kernel, dim3(bpm, n_maps), dim3(uph::kFcBlock), lds, s,
                     map_off, map_hw, L, x, m, partial);
  hipLaunchKernelGGL(uph::fc_autocov_reduce_kernel,
                     dim3((2 * nlag + UPH_BLOCK - 1) / UPH_BLOCK, n_maps), dim3(UPH_BLOCK), 0, s,
                     bpm, nlag, partial, out);
  return uph::hip_status(hipGetLastError(), "uph_fcorr_autocov launch");
}
