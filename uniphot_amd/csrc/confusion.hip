// Confusion-noise flux error of the post-fit error model
// (tract7dt/fit/grid_gpu/error_model.py:268-325 _confusion_band_numeric,
// driven per grid and band by compute_grid_confusion :663-930).
//
// Per (grid, band) problem, three launches:
//  1. conf_prep_kernel: residual r = good ? d - wing - model : 0 (model =
//     sum_s flux_s x unit stamp_s + the background plane), g = good, and the
//     pixel variance v = good & invvar > 0 ? 1 / invvar : 0;
//  2. conf_corr_kernel: the 'same' correlations corr(r, phi), corr(g, phi^2)
//     and corr(v, phi^2) with the band's unit-sum central PSF phi, tiled
//     through LDS like the blank-response correlation (4 adjacent outputs
//     per thread); the matched filter MF = corr(r, phi) / corr(g, phi^2) and
//     Vstat = corr(v, phi^2) / corr(g, phi^2)^2 where the kernel-weighted
//     occupancy corr(g, phi^2) exceeds 0.6 sum(phi^2) (valid), else 0;
//  3. conf_annulus_kernel: one wavefront per source: float64 sums of valid,
//     valid MF, valid MF^2 and valid Vstat over the square annulus
//     r_in < max(|dx|, |dy|) <= r_out around the source pixel (clipped to the
//     window), excess = max(0, <MF^2> - <MF>^2 - <Vstat>) where at least 50
//     valid pixels, and sigma_conf = sqrt(R excess) with the extended-template
//     factor R = 2 s^4 / ((s^2 + sp^2) sp^2), s^2 = sp^2 + s_gal^2.
// The correlations are VALU-bound (3 T^2 multiply-adds per window pixel);
// the other two launches are HBM-bound passes over the window.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"

namespace uph {

struct ConfArgs {
  int n_prob, T, bg_order, H, W, max_pix;
  const int64_t* src_off;
  const int32_t* win;  // (n_prob, 5) band, x0, y0, h, w
  const float *data, *invvar, *wing;
  const uint8_t *good, *use_wing;
  const float* stamps;  // (n_src, T, T)
  const int32_t *ix_top, *iy_top;
  const float *flux, *bg;  // (n_src), (n_prob, M)
  double bxo, bxs, byo, bys;
  const float* phi;        // (n_prob, T, T) unit sum
  const float* sumsq;      // (n_prob) sum(phi^2), float32
  const int32_t* radii;    // (n_prob, 2) r_in, r_out
  const double* sp2;       // (n_prob) Gaussian-equivalent PSF variance
  const int32_t *xi, *yi;  // (n_src) window pixel of the source
  const double* sgal2;     // (n_src) template variance
  const uint8_t* no_fit;   // (n_src)
  float* work;             // (n_prob, 6, max_pix)
  float* sigma;            // (n_src)
};

constexpr int kConfRows = 16, kConfCols = 64, kConfBlock = 256;

__host__ __device__ inline int conf_tp(int T) { return (T + 3) / 4 * 4; }
__host__ __device__ inline int conf_sx(int T) { return kConfCols + conf_tp(T) + 4; }
__host__ __device__ inline size_t conf_lds_bytes(int T) {
  return ((size_t)2 * T * conf_tp(T) + (size_t)3 * (kConfRows + T - 1) * conf_sx(T)) *
         sizeof(float);
}

template <int M>
__global__ __launch_bounds__(UPH_BLOCK) void conf_prep_kernel(ConfArgs a) {
  const int q = blockIdx.x;
  const int band = a.win[q * 5], x0 = a.win[q * 5 + 1], y0 = a.win[q * 5 + 2];
  const int h = a.win[q * 5 + 3], w = a.win[q * 5 + 4], T = a.T, TT = T * T;
  const int64_t s0 = a.src_off[q], s1 = a.src_off[q + 1];
  const size_t off = (size_t)band * a.H * a.W;
  float* R = a.work + (size_t)q * 6 * a.max_pix;
  float* G = R + a.max_pix;
  float* V = G + a.max_pix;
  const float* bgq = a.bg + (size_t)q * M;
  const bool wing = a.wing != nullptr && a.use_wing[q];
  for (int e = blockIdx.y * UPH_BLOCK + threadIdx.x; e < h * w; e += gridDim.y * UPH_BLOCK) {
    const int py = e / w, px = e % w;
    const size_t gi = off + (size_t)(y0 + py) * a.W + x0 + px;
    const bool gd = a.good[gi] != 0;
    const float iv = a.invvar[gi];
    float model = 0.0f;
    for (int64_t s = s0; s < s1; ++s) {
      const int u = px - a.ix_top[s], t = py - a.iy_top[s];
      if (u < 0 || u >= T || t < 0 || t >= T) continue;
      model = fmaf(a.flux[s], a.stamps[(size_t)s * TT + t * T + u], model);
    }
    const double X = (double(x0 + px) - a.bxo) / a.bxs, Y = (double(y0 + py) - a.byo) / a.bys;
    constexpr int m_[6] = {0, 1, 0, 2, 1, 0};
    constexpr int n_[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
    for (int m = 0; m < M; ++m) {
      double v = 1.0;
      for (int k = 0; k < m_[m]; ++k) v *= X;
      for (int k = 0; k < n_[m]; ++k) v *= Y;
      model = fmaf(bgq[m], float(v), model);
    }
    const float d = a.data[gi] - (wing ? a.wing[gi] : 0.0f);
    R[e] = gd ? d - model : 0.0f;
    G[e] = gd ? 1.0f : 0.0f;
    V[e] = (gd && iv > 0.0f) ? 1.0f / fmaxf(iv, 1e-30f) : 0.0f;
  }
}

__global__ __launch_bounds__(kConfBlock) void conf_corr_kernel(ConfArgs a) {
  extern __shared__ float4 lds4[];
  float* lds = reinterpret_cast<float*>(lds4);
  const int q = blockIdx.x;
  const int h = a.win[q * 5 + 3], w = a.win[q * 5 + 4], T = a.T, c = T / 2;
  const int TP = conf_tp(T), SX = conf_sx(T), SY = kConfRows + T - 1;
  const int tiles_x = (w + kConfCols - 1) / kConfCols, tiles_y = (h + kConfRows - 1) / kConfRows;
  float* P1 = lds;          // phi (T rows of TP)
  float* P2 = P1 + T * TP;  // phi^2
  float* A1 = P2 + T * TP;  // r tile with halo (SY rows of SX)
  float* A2 = A1 + SY * SX; // g tile
  float* A3 = A2 + SY * SX; // v tile
  float* base = a.work + (size_t)q * 6 * a.max_pix;
  const float* R = base;
  const float* G = base + a.max_pix;
  const float* V = base + 2 * (size_t)a.max_pix;
  float* MF = base + 3 * (size_t)a.max_pix;
  float* VS = base + 4 * (size_t)a.max_pix;
  float* VL = base + 5 * (size_t)a.max_pix;
  const float* ph = a.phi + (size_t)q * T * T;
  const float thr = 0.6f * a.sumsq[q];
  for (int e = threadIdx.x; e < T * TP; e += kConfBlock) {
    const int ky = e / TP, kx = e % TP;
    const float v = kx < T ? ph[ky * T + kx] : 0.0f;
    P1[e] = v;
    P2[e] = v * v;
  }
  const int ly = threadIdx.x / (kConfCols / 4), lx = 4 * (threadIdx.x % (kConfCols / 4));
  for (int tile = blockIdx.y; tile < tiles_x * tiles_y; tile += gridDim.y) {
    const int ty0 = (tile / tiles_x) * kConfRows, tx0 = (tile % tiles_x) * kConfCols;
    __syncthreads();
    for (int e = threadIdx.x; e < SY * SX; e += kConfBlock) {
      const int yy = ty0 - c + e / SX, xx = tx0 - c + e % SX;
      const bool in = yy >= 0 && yy < h && xx >= 0 && xx < w;
      A1[e] = in ? R[yy * w + xx] : 0.0f;
      A2[e] = in ? G[yy * w + xx] : 0.0f;
      A3[e] = in ? V[yy * w + xx] : 0.0f;
    }
    __syncthreads();
    float num[4] = {0, 0, 0, 0}, den[4] = {0, 0, 0, 0}, vn[4] = {0, 0, 0, 0};
    for (int ky = 0; ky < T; ++ky) {
      const float4* r1 = reinterpret_cast<const float4*>(A1 + (ly + ky) * SX + lx);
      const float4* r2 = reinterpret_cast<const float4*>(A2 + (ly + ky) * SX + lx);
      const float4* r3 = reinterpret_cast<const float4*>(A3 + (ly + ky) * SX + lx);
      const float4* k1 = reinterpret_cast<const float4*>(P1 + ky * TP);
      const float4* k2 = reinterpret_cast<const float4*>(P2 + ky * TP);
      float4 u1 = r1[0], u2 = r2[0], u3 = r3[0];
      for (int k4 = 0; k4 < TP / 4; ++k4) {
        const float4 v1 = r1[k4 + 1], v2 = r2[k4 + 1], v3 = r3[k4 + 1];
        const float4 t1 = k1[k4], t2 = k2[k4];
        const float x1[8] = {u1.x, u1.y, u1.z, u1.w, v1.x, v1.y, v1.z, v1.w};
        const float x2[8] = {u2.x, u2.y, u2.z, u2.w, v2.x, v2.y, v2.z, v2.w};
        const float x3[8] = {u3.x, u3.y, u3.z, u3.w, v3.x, v3.y, v3.z, v3.w};
        const float c1[4] = {t1.x, t1.y, t1.z, t1.w};
        const float c2[4] = {t2.x, t2.y, t2.z, t2.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            num[j] = fmaf(x1[j + t], c1[t], num[j]);
            den[j] = fmaf(x2[j + t], c2[t], den[j]);
            vn[j] = fmaf(x3[j + t], c2[t], vn[j]);
          }
        }
        u1 = v1;
        u2 = v2;
        u3 = v3;
      }
    }
    const int oy = ty0 + ly;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int ox = tx0 + lx + j;
      if (oy >= h || ox >= w) continue;
      const bool valid = den[j] > thr && isfinite(num[j]);
      const float ds = fmaxf(den[j], 1e-30f);
      MF[oy * w + ox] = valid ? num[j] / ds : 0.0f;
      VS[oy * w + ox] = valid ? vn[j] / (ds * ds) : 0.0f;
      VL[oy * w + ox] = valid ? 1.0f : 0.0f;
    }
  }
}

__global__ __launch_bounds__(UPH_BLOCK) void conf_annulus_kernel(ConfArgs a, int64_t n_src) {
  const int lane = threadIdx.x & 63;
  const int64_t s = int64_t(blockIdx.x) * (UPH_BLOCK / 64) + (threadIdx.x >> 6);
  if (s >= n_src) return;
  // the problem of source s: src_off is sorted, n_prob is small
  int lo = 0, hi = a.n_prob - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (a.src_off[mid] <= s)
      lo = mid;
    else
      hi = mid - 1;
  }
  const int q = lo;
  const int h = a.win[q * 5 + 3], w = a.win[q * 5 + 4];
  const float* base = a.work + (size_t)q * 6 * a.max_pix;
  const float* MF = base + 3 * (size_t)a.max_pix;
  const float* VS = base + 4 * (size_t)a.max_pix;
  const float* VL = base + 5 * (size_t)a.max_pix;
  const int r_in = a.radii[2 * q], r_out = a.radii[2 * q + 1];
  const int cx = a.xi[s], cy = a.yi[s];
  const int y0 = max(cy - r_out, 0), y1 = min(cy + r_out, h - 1);
  const int x0 = max(cx - r_out, 0), x1 = min(cx + r_out, w - 1);
  const int bw = x1 - x0 + 1, nbox = (y1 - y0 + 1) * bw;
  double n = 0.0, s1 = 0.0, s2 = 0.0, sv = 0.0;
  for (int e = lane; e < nbox; e += 64) {
    const int yy = y0 + e / bw, xx = x0 + e % bw;
    if (abs(yy - cy) <= r_in && abs(xx - cx) <= r_in) continue;
    const int p = yy * w + xx;
    const double vl = VL[p];
    const double mf = double(MF[p]) * vl;
    n += vl;
    s1 += mf;
    s2 += mf * double(MF[p]);
    sv += double(VS[p]) * vl;
  }
  n = wave_sum(n);
  s1 = wave_sum(s1);
  s2 = wave_sum(s2);
  sv = wave_sum(sv);
  if (lane != 0) return;
  float out = NAN;
  if (!a.no_fit[s] && n >= 50.0) {
    const double m1 = s1 / n, m2 = s2 / n, vsl = sv / n;
    const double exc = fmax(0.0, (m2 - m1 * m1) - vsl);
    const double sp2 = a.sp2[q], sg = sp2 + a.sgal2[s];
    const double R = 2.0 * sg * sg / ((sg + sp2) * sp2);
    out = float(sqrt(R * exc));
  }
  a.sigma[s] = out;
}

}  // namespace uph

extern "C" size_t uph_confusion_work_floats(int n_prob, int max_pix) {
  return size_t(6) * size_t(n_prob > 0 ? n_prob : 0) * size_t(max_pix > 0 ? max_pix : 0);
}

extern "C" int uph_confusion(int n_prob, int64_t n_src, int T, int bg_order, const int64_t* src_off,
                             const int32_t* win, int H, int W, const float* data,
                             const float* invvar, const uint8_t* good, const float* wing,
                             const uint8_t* use_wing, const float* stamps, const int32_t* ix_top,
                             const int32_t* iy_top, const float* flux, const float* bg,
                             double bg_xo, double bg_xs, double bg_yo, double bg_ys,
                             const float* phi, const float* sumsq, const int32_t* radii,
                             const double* sp2, const int32_t* xi, const int32_t* yi,
                             const double* sgal2, const uint8_t* no_fit, float* work,
                             size_t work_floats, int max_pix, float* sigma, void* stream) {
  UPH_CHECK_ARG(n_prob >= 0 && n_src >= 0 && T >= 1 && (T & 1) && bg_order >= 0 &&
                    bg_order <= 2 && H >= 1 && W >= 1 && max_pix >= 0,
                "uph_confusion: n_prob=%d n_src=%lld T=%d bg_order=%d image %d x %d", n_prob,
                (long long)n_src, T, bg_order, H, W);
  if (n_prob == 0) return UPH_OK;
  UPH_CHECK_ARG(src_off && win && data && invvar && good && stamps && ix_top && iy_top && flux &&
                    bg && phi && sumsq && radii && sp2 && xi && yi && sgal2 && no_fit && work &&
                    sigma,
                "uph_confusion: null argument");
  UPH_CHECK_ARG(work_floats >= uph_confusion_work_floats(n_prob, max_pix),
                "uph_confusion: work of %zu floats < %zu", work_floats,
                uph_confusion_work_floats(n_prob, max_pix));
  const size_t lds = uph::conf_lds_bytes(T);
  UPH_CHECK_ARG(lds <= 160 * 1024, "uph_confusion: template T=%d too large (%zu B of LDS)", T,
                lds);
  uph::ConfArgs a{n_prob, T, bg_order, H, W, max_pix, src_off, win, data, invvar, wing, good,
                  use_wing, stamps, ix_top, iy_top, flux, bg, bg_xo, bg_xs, bg_yo, bg_ys, phi,
                  sumsq, radii, sp2, xi, yi, sgal2, no_fit, work, sigma};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (bg_order == 0)
    hipLaunchKernelGGL(uph::conf_prep_kernel<1>, dim3(n_prob, 16), dim3(UPH_BLOCK), 0, s, a);
  else if (bg_order == 1)
    hipLaunchKernelGGL(uph::conf_prep_kernel<3>, dim3(n_prob, 16), dim3(UPH_BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL(uph::conf_prep_kernel<6>, dim3(n_prob, 16), dim3(UPH_BLOCK), 0, s, a);
  hipLaunchKernelGGL(uph::conf_corr_kernel, dim3(n_prob, 32), dim3(uph::kConfBlock), lds, s, a);
  if (n_src > 0) {
    const int64_t blocks = (n_src + UPH_BLOCK / 64 - 1) / (UPH_BLOCK / 64);
    hipLaunchKernelGGL(uph::conf_annulus_kernel, dim3(unsigned(blocks)), dim3(UPH_BLOCK), 0, s,
                       a, n_src);
  }
  return uph::hip_status(hipGetLastError(), "uph_confusion launch");
}
