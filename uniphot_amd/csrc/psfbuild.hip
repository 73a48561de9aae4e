// PSF model build from training stars (tract7dt/fit/grid_gpu/psf_from_stars.py
// :1478-2239, the gw_centroid path): one workgroup per star for the pixel work,
// float64 throughout except where the reference computes in float32.
//
//  pb_cutout_kernel   61-px cutout at the rounded position (dropped off the
//                     frame or on a bad pixel), sigma-clipped annulus median
//                     background (float32 values in r^2 in [rin^2, rout^2],
//                     sorted in LDS: median, 1.4826 MAD, clip, median of the
//                     kept run), cutout minus the background
//  pb_centroid_kernel Gaussian-weighted first moments of max(cutout, 0), every
//                     iteration's centre and move stored (the host stops all
//                     stars at the first iteration whose largest move is
//                     below the tolerance, as the reference does)
//  pb_stamp_kernel    recentred cutout at the rounded centroid, its
//                     background, the Fourier shift by minus the fractional
//                     part written as the separable Dirichlet kernel of the
//                     zero-padded length m (odd, so the 2-D shift is the
//                     product of two real 1-D operators), cropped and
//                     normalised to unit sum
//  pb_polyfit_kernel  coefficients = pinv(A) x stamps per pixel (float64,
//                     rounded to float32), pinv(A) from the host's SVD
//  pb_chi2_kernel     per-star sum of squared residuals against the float32
//                     coefficients
//  pb_median_kernel   pixelwise median over the accepted stamps (sorted in LDS)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"

namespace uph {

constexpr int kPbBlock = 256;
constexpr int kPbSort = 4096;   // LDS sort capacity (annulus pixels, stars per pixel)
constexpr int kPbCut = 61, kPbMaxPsf = 61, kPbMaxTerms = 21;

// ascending bitonic sort of kPbSort floats in LDS (padding +inf)
__device__ void pb_sort(float* a) {
  for (int k = 2; k <= kPbSort; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < kPbSort; i += kPbBlock) {
        const int p = i ^ j;
        if (p > i) {
          const bool up = (i & k) == 0;
          const float x = a[i], y = a[p];
          if ((x > y) == up) {
            a[i] = y;
            a[p] = x;
          }
        }
      }
      __syncthreads();
    }
}

// numpy median of the first n sorted float32 values, in float32
__device__ inline float pb_median_sorted(const float* a, int lo, int n) {
  return n & 1 ? a[lo + n / 2] : (a[lo + n / 2 - 1] + a[lo + n / 2]) * 0.5f;
}

// sigma-clipped annulus median (psf_from_stars.py:1957-1984); NaN below minpix
__device__ double pb_annulus_median(const float* img, const uint8_t* bad, int H, int W, int ix,
                                    int iy, double rin, double rout, double sig, int minpix,
                                    float* s, int* cnt) {
  const int ri = (int)ceil(rout);
  const int y0 = max(0, iy - ri), y1 = min(H, iy + ri + 1);
  const int x0 = max(0, ix - ri), x1 = min(W, ix + ri + 1);
  const int bw = x1 - x0, nb = (y1 - y0) * bw;
  if (threadIdx.x == 0) *cnt = 0;
  for (int i = threadIdx.x; i < kPbSort; i += kPbBlock) s[i] = INFINITY;
  __syncthreads();
  const double rin2 = rin * rin, rout2 = rout * rout;
  for (int e = threadIdx.x; e < nb; e += kPbBlock) {
    const int y = y0 + e / bw, x = x0 + e % bw;
    const int dy = y - iy, dx = x - ix;
    const double r2 = (double)(dy * dy + dx * dx);
    const size_t gi = (size_t)y * W + x;
    const float v = img[gi];
    if (r2 >= rin2 && r2 <= rout2 && !bad[gi] && isfinite(v)) s[atomicAdd(cnt, 1)] = v;
  }
  __syncthreads();
  const int n = *cnt;
  if (n < minpix) return NAN;
  pb_sort(s);
  const float med = pb_median_sorted(s, 0, n);
  __syncthreads();
  const float* d = s;  // the sorted values stay in s for the clipped median
  __shared__ float keep_lo, keep_hi;
  __shared__ float madf;
  // MAD: |v - med| in float32, sorted in a second buffer
  __shared__ float tmp[kPbSort];
  for (int i = threadIdx.x; i < kPbSort; i += kPbBlock) tmp[i] = i < n ? fabsf(s[i] - med) : INFINITY;
  __syncthreads();
  pb_sort(tmp);
  if (threadIdx.x == 0) madf = pb_median_sorted(tmp, 0, n);
  __syncthreads();
  const double mad = (double)madf * 1.4826;
  if (!(mad > 0.0)) return (double)med;
  if (threadIdx.x == 0) {
    keep_lo = (float)((double)med - sig * mad);
    keep_hi = (float)((double)med + sig * mad);
  }
  __syncthreads();
  // the kept values are a contiguous run of the sorted values
  int lo = 0, hi = n;  // [lo, hi)
  {
    int a = 0, b = n;
    while (a < b) {
      const int m = (a + b) >> 1;
      if (d[m] >= keep_lo) b = m; else a = m + 1;
    }
    lo = a;
    a = lo;
    b = n;
    while (a < b) {
      const int m = (a + b) >> 1;
      if (d[m] <= keep_hi) a = m + 1; else b = m;
    }
    hi = a;
  }
  if (hi - lo < minpix) return (double)med;
  return (double)pb_median_sorted(d, lo, hi - lo);
}

template <typename T>
__device__ T pb_block_sum(T v, T* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  T s = T(0);
  for (int i = 0; i < kPbBlock / 64; ++i) s += red[i];
  return s;
}

__global__ __launch_bounds__(kPbBlock) void pb_cutout_kernel(
    int n, int H, int W, const float* img, const uint8_t* bad, const int32_t* ox,
    const int32_t* oy, int half, double rin, double rout, double sig, int minpix, int do_bkg,
    uint8_t* status, double* bkg, float* cut) {
  __shared__ float srt[kPbSort];
  __shared__ int cnt, anybad;
  const int k = blockIdx.x, ix = ox[k], iy = oy[k], C = 2 * half + 1;
  if (ix - half < 0 || ix + half + 1 > W || iy - half < 0 || iy + half + 1 > H) {
    if (threadIdx.x == 0) {
      status[k] = 1;
      bkg[k] = NAN;
    }
    return;
  }
  if (threadIdx.x == 0) anybad = 0;
  __syncthreads();
  for (int e = threadIdx.x; e < C * C; e += kPbBlock)
    if (bad[(size_t)(iy - half + e / C) * W + ix - half + e % C]) anybad = 1;
  __syncthreads();
  if (anybad) {
    if (threadIdx.x == 0) {
      status[k] = 2;
      bkg[k] = NAN;
    }
    return;
  }
  const double b = do_bkg ? pb_annulus_median(img, bad, H, W, ix, iy, rin, rout, sig, minpix,
                                               srt, &cnt)
                          : (double)NAN;
  const bool sub = isfinite(b);
  const float bf = (float)b;
  for (int e = threadIdx.x; e < C * C; e += kPbBlock) {
    const float v = img[(size_t)(iy - half + e / C) * W + ix - half + e % C];
    cut[(size_t)k * C * C + e] = sub ? v - bf : v;
  }
  if (threadIdx.x == 0) {
    status[k] = 0;
    bkg[k] = b;
  }
}

__global__ __launch_bounds__(kPbBlock) void pb_centroid_kernel(int n, int half, double fwhm,
                                                               int iters, const float* cut,
                                                               double* hist) {
  __shared__ double red[kPbBlock / 64];
  const int k = blockIdx.x, C = 2 * half + 1;
  const double sg = fwhm / 2.354820045, kk = 1.0 / (2.0 * sg * sg);
  const float* c = cut + (size_t)k * C * C;
  double cx = half, cy = half;
  for (int it = 0; it < iters; ++it) {
    double s = 0.0, sx = 0.0, sy = 0.0;
    for (int e = threadIdx.x; e < C * C; e += kPbBlock) {
      const double yy = e / C, xx = e % C;
      const double dx = xx - cx, dy = yy - cy;
      const double wd = exp(-(dx * dx + dy * dy) * kk) * fmax((double)c[e], 0.0);
      s += wd;
      sx += wd * xx;
      sy += wd * yy;
    }
    s = pb_block_sum(s, red);
    sx = pb_block_sum(sx, red);
    sy = pb_block_sum(sy, red);
    const double nx = s > 0.0 ? sx / s : cx, ny = s > 0.0 ? sy / s : cy;
    if (threadIdx.x == 0) {
      double* h = hist + ((size_t)k * iters + it) * 3;
      h[0] = nx;
      h[1] = ny;
      h[2] = fmax(fabs(nx - cx), fabs(ny - cy));
    }
    cx = nx;
    cy = ny;
  }
}

// D(k + eps) of the zero-padded length m (odd): the 1-D Fourier shift kernel
__device__ inline double pb_dirichlet(int k, double eps, int m) {
  if (eps == 0.0) return (k % m == 0) ? 1.0 : 0.0;
  const double num = ((k & 1) ? -1.0 : 1.0) * sin(M_PI * eps);
  return num / ((double)m * sin(M_PI * ((double)k + eps) / (double)m));
}

__global__ __launch_bounds__(kPbBlock) void pb_stamp_kernel(
    int n, int H, int W, const float* img, const uint8_t* bad, const int32_t* ox,
    const int32_t* oy, const double* fx, const double* fy, const double* bkg0, int half,
    int psf, int pad, double rin, double rout, double sig, int minpix, int do_bkg,
    uint8_t* status, float* stamps) {
  extern __shared__ double pb_lds[];
  __shared__ float srt[kPbSort];
  __shared__ int cnt, anybad;
  __shared__ double red[kPbBlock / 64];
  const int k = blockIdx.x, C = 2 * half + 1, c0 = (C - psf) / 2, m = C + 2 * pad;
  double* cutd = pb_lds;                 // C x C
  double* Dx = cutd + C * C;             // psf x C
  double* Dy = Dx + psf * C;             // psf x C
  double* Tm = Dy + psf * C;             // C x psf: cut . Dx^T
  const double xc = fx[k], yc = fy[k];
  if (!(isfinite(xc) && isfinite(yc))) {
    if (threadIdx.x == 0) status[k] = 3;
    return;
  }
  const int ix = (int)rint(xc), iy = (int)rint(yc);
  const double dx = xc - ix, dy = yc - iy;
  if (iy - half < 0 || ix - half < 0 || iy + half + 1 > H || ix + half + 1 > W) {
    if (threadIdx.x == 0) status[k] = 1;
    return;
  }
  if (threadIdx.x == 0) anybad = 0;
  __syncthreads();
  for (int e = threadIdx.x; e < C * C; e += kPbBlock)
    if (bad[(size_t)(iy - half + e / C) * W + ix - half + e % C]) anybad = 1;
  __syncthreads();
  if (anybad) {
    if (threadIdx.x == 0) status[k] = 2;
    return;
  }
  double b = NAN;
  if (do_bkg)
    b = (ix == ox[k] && iy == oy[k])
            ? bkg0[k]
            : pb_annulus_median(img, bad, H, W, ix, iy, rin, rout, sig, minpix, srt, &cnt);
  const bool sub = isfinite(b);
  const float bf = (float)b;
  for (int e = threadIdx.x; e < C * C; e += kPbBlock) {
    const float v = img[(size_t)(iy - half + e / C) * W + ix - half + e % C];
    cutd[e] = (double)(sub ? v - bf : v);
  }
  // out(p) = in(p - s) with s = -d: D(p - q + d), p = i + c0
  for (int e = threadIdx.x; e < psf * C; e += kPbBlock) {
    const int i = e / C, q = e % C;
    Dx[e] = pb_dirichlet(i + c0 - q, dx, m);
    Dy[e] = pb_dirichlet(i + c0 - q, dy, m);
  }
  __syncthreads();
  for (int e = threadIdx.x; e < C * psf; e += kPbBlock) {
    const int r = e / psf, j = e % psf;
    double acc = 0.0;
    for (int q = 0; q < C; ++q) acc = fma(cutd[r * C + q], Dx[j * C + q], acc);
    Tm[e] = acc;
  }
  __syncthreads();
  double part = 0.0;
  float* out = stamps + (size_t)k * psf * psf;
  for (int e = threadIdx.x; e < psf * psf; e += kPbBlock) {
    const int i = e / psf, j = e % psf;
    double acc = 0.0;
    for (int r = 0; r < C; ++r) acc = fma(Dy[i * C + r], Tm[r * psf + j], acc);
    cutd[e] = acc;  // the cutout is no longer needed
    part += acc;
  }
  const double s = pb_block_sum(part, red);
  const bool ok = isfinite(s) && s > 0.0;
  for (int e = threadIdx.x; e < psf * psf; e += kPbBlock) out[e] = ok ? (float)(cutd[e] / s) : 0.0f;
  if (threadIdx.x == 0) status[k] = ok ? 0 : 4;
}

// coefs[p * nt + t] = float32(sum_s pinv[t * ns + s] * stamps[sel[s]][p])
__global__ __launch_bounds__(kPbBlock) void pb_polyfit_kernel(int ns, int nt, int P,
                                                              const int32_t* sel,
                                                              const double* pinv,
                                                              const float* stamps, float* coefs) {
  const int p = blockIdx.x * kPbBlock + threadIdx.x;
  if (p >= P) return;
  double acc[kPbMaxTerms];
  for (int t = 0; t < nt; ++t) acc[t] = 0.0;
  for (int s = 0; s < ns; ++s) {
    const double y = (double)stamps[(size_t)sel[s] * P + p];
    for (int t = 0; t < nt; ++t) acc[t] = fma(pinv[t * ns + s], y, acc[t]);
  }
  for (int t = 0; t < nt; ++t) coefs[(size_t)p * nt + t] = (float)acc[t];
}

__global__ __launch_bounds__(kPbBlock) void pb_chi2_kernel(int n, int nt, int P, const double* A,
                                                           const float* stamps,
                                                           const float* coefs, double* chi2) {
  __shared__ double red[kPbBlock / 64];
  const int s = blockIdx.x;
  double a[kPbMaxTerms];
  for (int t = 0; t < nt; ++t) a[t] = A[(size_t)s * nt + t];
  double part = 0.0;
  for (int p = threadIdx.x; p < P; p += kPbBlock) {
    double mdl = 0.0;
    for (int t = 0; t < nt; ++t) mdl = fma(a[t], (double)coefs[(size_t)p * nt + t], mdl);
    const double d = (double)stamps[(size_t)s * P + p] - mdl;
    part = fma(d, d, part);
  }
  const double tot = pb_block_sum(part, red);
  if (threadIdx.x == 0) chi2[s] = tot;
}

// med[p] = median over the selected stamps of stamps[sel][p] (float64 of float32)
__global__ __launch_bounds__(kPbBlock) void pb_median_kernel(int ns, int P, const int32_t* sel,
                                                             const float* stamps, double* med) {
  __shared__ float srt[kPbSort];
  const int p = blockIdx.x;
  for (int i = threadIdx.x; i < kPbSort; i += kPbBlock)
    srt[i] = i < ns ? stamps[(size_t)sel[i] * P + p] : INFINITY;
  __syncthreads();
  pb_sort(srt);
  if (threadIdx.x == 0)
    med[p] = ns & 1 ? (double)srt[ns / 2]
                    : ((double)srt[ns / 2 - 1] + (double)srt[ns / 2]) / 2.0;
}

inline size_t pb_stamp_lds(int C, int psf) {
  return sizeof(double) * ((size_t)C * C + 2 * (size_t)psf * C + (size_t)C * psf);
}

}  // namespace uph

extern "C" int uph_psf_cutouts(int n, int H, int W, const float* image, const uint8_t* bad,
                               const int32_t* ox, const int32_t* oy, int half, double rin,
                               double rout, double sigma, int minpix, int do_bkg,
                               uint8_t* status, double* bkg, float* cut, void* stream) {
  UPH_CHECK_ARG(n >= 0 && H >= 1 && W >= 1 && half >= 1 && 2 * half + 1 <= uph::kPbCut &&
                    minpix >= 1 && rin >= 0.0 && rout >= rin && (2 * (int)ceil(rout) + 1) *
                    (2 * (int)ceil(rout) + 1) <= uph::kPbSort,
                "uph_psf_cutouts: n=%d image %d x %d half=%d annulus %g..%g", n, H, W, half,
                rin, rout);
  if (n == 0) return UPH_OK;
  UPH_CHECK_ARG(image && bad && ox && oy && status && bkg && cut, "uph_psf_cutouts: null argument");
  hipLaunchKernelGGL(uph::pb_cutout_kernel, dim3(n), dim3(uph::kPbBlock), 0,
                     static_cast<hipStream_t>(stream), n, H, W, image, bad, ox, oy, half, rin,
                     rout, sigma, minpix, do_bkg, status, bkg, cut);
  return uph::hip_status(hipGetLastError(), "uph_psf_cutouts launch");
}

extern "C" int uph_psf_centroids(int n, int half, double fwhm, int iters, const float* cut,
                                 double* hist, void* stream) {
  UPH_CHECK_ARG(n >= 0 && half >= 1 && 2 * half + 1 <= uph::kPbCut && iters >= 1 && fwhm > 0.0,
                "uph_psf_centroids: n=%d half=%d iters=%d fwhm=%g", n, half, iters, fwhm);
  if (n == 0) return UPH_OK;
  UPH_CHECK_ARG(cut && hist, "uph_psf_centroids: null argument");
  hipLaunchKernelGGL(uph::pb_centroid_kernel, dim3(n), dim3(uph::kPbBlock), 0,
                     static_cast<hipStream_t>(stream), n, half, fwhm, iters, cut, hist);
  return uph::hip_status(hipGetLastError(), "uph_psf_centroids launch");
}

extern "C" int uph_psf_stamps(int n, int H, int W, const float* image, const uint8_t* bad,
                              const int32_t* ox, const int32_t* oy, const double* fx,
                              const double* fy, const double* bkg, int half, int psf, int pad,
                              double rin, double rout, double sigma, int minpix, int do_bkg,
                              uint8_t* status, float* stamps, void* stream) {
  const int C = 2 * half + 1;
  UPH_CHECK_ARG(n >= 0 && H >= 1 && W >= 1 && half >= 1 && C <= uph::kPbCut && psf >= 1 &&
                    psf <= C && ((C - psf) & 1) == 0 && pad >= 0 && ((C + 2 * pad) & 1) &&
                    minpix >= 1 && rin >= 0.0 && rout >= rin &&
                    (2 * (int)ceil(rout) + 1) * (2 * (int)ceil(rout) + 1) <= uph::kPbSort,
                "uph_psf_stamps: n=%d half=%d psf=%d pad=%d (odd padded length, centred crop)",
                n, half, psf, pad);
  if (n == 0) return UPH_OK;
  UPH_CHECK_ARG(image && bad && ox && oy && fx && fy && bkg && status && stamps,
                "uph_psf_stamps: null argument");
  const size_t lds = uph::pb_stamp_lds(C, psf);
  hipLaunchKernelGGL(uph::pb_stamp_kernel, dim3(n), dim3(uph::kPbBlock), lds,
                     static_cast<hipStream_t>(stream), n, H, W, image, bad, ox, oy, fx, fy, bkg,
                     half, psf, pad, rin, rout, sigma, minpix, do_bkg, status, stamps);
  return uph::hip_status(hipGetLastError(), "uph_psf_stamps launch");
}

extern "C" int uph_psf_polyfit(int ns, int nt, int P, const int32_t* sel, const double* pinv,
                               const float* stamps, float* coefs, void* stream) {
  UPH_CHECK_ARG(ns >= 1 && nt >= 1 && nt <= uph::kPbMaxTerms && P >= 1,
                "uph_psf_polyfit: ns=%d nt=%d P=%d", ns, nt, P);
  UPH_CHECK_ARG(sel && pinv && stamps && coefs, "uph_psf_polyfit: null argument");
  hipLaunchKernelGGL(uph::pb_polyfit_kernel, dim3((P + uph::kPbBlock - 1) / uph::kPbBlock),
                     dim3(uph::kPbBlock), 0, static_cast<hipStream_t>(stream), ns, nt, P, sel,
                     pinv, stamps, coefs);
  return uph::hip_status(hipGetLastError(), "uph_psf_polyfit launch");
}

extern "C" int uph_psf_chi2(int n, int nt, int P, const double* A, const float* stamps,
                            const float* coefs, double* chi2, void* stream) {
  UPH_CHECK_ARG(n >= 0 && nt >= 1 && nt <= uph::kPbMaxTerms && P >= 1,
                "uph_psf_chi2: n=%d nt=%d P=%d", n, nt, P);
  if (n == 0) return UPH_OK;
  UPH_CHECK_ARG(A && stamps && coefs && chi2, "uph_psf_chi2: null argument");
  hipLaunchKernelGGL(uph::pb_chi2_kernel, dim3(n), dim3(uph::kPbBlock), 0,
                     static_cast<hipStream_t>(stream), n, nt, P, A, stamps, coefs, chi2);
  return uph::hip_status(hipGetLastError(), "uph_psf_chi2 launch");
}

extern "C" int uph_psf_pixel_median(int ns, int P, const int32_t* sel, const float* stamps,
                                    double* med, void* stream) {
  UPH_CHECK_ARG(ns >= 1 && ns <= uph::kPbSort && P >= 1,
                "uph_psf_pixel_median: ns=%d (at most %d) P=%d", ns, uph::kPbSort, P);
  UPH_CHECK_ARG(sel && stamps && med, "uph_psf_pixel_median: null argument");
  hipLaunchKernelGGL(uph::pb_median_kernel, dim3(P), dim3(uph::kPbBlock), 0,
                     static_cast<hipStream_t>(stream), ns, P, sel, stamps, med);
  return uph::hip_status(hipGetLastError(), "uph_psf_pixel_median launch");
}
