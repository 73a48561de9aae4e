// Per-(source, band) PSF diagnostics behind the band flags
// (tract7dt/fit/grid_gpu/band_flags.py): the PSF model stamp at a fitted
// position, its peak and minimum (negative-lobe zone, :115-137) and the
// fraction of its positive part that falls on pixels without data
// (partial coverage, :140-196).
//
// One wavefront per (band, source) item, items band-major so neighbouring
// waves read the same band's coefficient planes (L2-resident: 15 planes of
// 55 x 55 floats).  The stamp is never stored: each lane evaluates its
// pixels, keeps running max / min and float64 sums of the positive part,
// total and on no-data pixels, and the wave reduces them.  HBM traffic is
// the no-data mask under each stamp (S^2 bytes) plus 16 B written per item.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"

namespace uph {

constexpr int kFlagWaves = UPH_BLOCK / 64;

__global__ __launch_bounds__(UPH_BLOCK) void psf_flags_kernel(
    int n, int nb, int nt, int T, const int32_t* __restrict__ size,
    const float* __restrict__ coefs, const double* __restrict__ geo, const double* __restrict__ xs,
    const double* __restrict__ ys, int H, int W, const uint8_t* __restrict__ nodata,
    const uint8_t* __restrict__ has_nodata, float* __restrict__ peak, float* __restrict__ vmin,
    double* __restrict__ uncovered) {
  const int lane = threadIdx.x & 63;
  const long item = long(blockIdx.x) * kFlagWaves + (threadIdx.x >> 6);
  if (item >= long(n) * nb) return;
  const int b = int(item / n);
  const int s = int(item - long(b) * n);
  const double x = xs[s], y = ys[s];
  const long out = long(b) * n + s;
  if (!isfinite(x) || !isfinite(y)) {
    // never flagged: no lobe, nothing uncovered
    if (lane == 0) {
      peak[out] = NAN;
      vmin[out] = NAN;
      uncovered[out] = 0.0;
    }
    return;
  }
  // basis in float64, rounded to float32 as the reference's numpy
  // evaluation does (psf_polynomial.py:262-268)
  const double X = (x - geo[b * 4 + 0]) / geo[b * 4 + 1];
  const double Y = (y - geo[b * 4 + 2]) / geo[b * 4 + 3];
  double bd[21];
  psf_basis<double>(X, Y, nt, bd, nullptr, nullptr);
  float bf[21];
  for (int t = 0; t < nt; ++t) bf[t] = float(bd[t]);

  const int S = size[b];
  const int o = (T - S) / 2;  // native box inside the padded T x T plane
  const int h = S / 2;
  const double rx = rint(x), ry = rint(y);
  const bool on = rx >= 0.0 && rx < double(W) && ry >= 0.0 && ry < double(H);
  const bool want_cov = has_nodata != nullptr && has_nodata[b] != 0 && nodata != nullptr;
  const int ix = on ? int(rx) : 0, iy = on ? int(ry) : 0;
  const uint8_t* nd = want_cov ? nodata + long(b) * H * W : nullptr;
  const float* cb = coefs + long(b) * nt * T * T;

  float pk = -INFINITY, mn = INFINITY;
  double tot = 0.0, cov = 0.0;
  int hit = 0;
  for (int p = lane; p < S * S; p += 64) {
    const int r = p / S, c = p - r * S;
    const int q = (o + r) * T + (o + c);
    float v = 0.0f;
    for (int t = 0; t < nt; ++t) v = fmaf(cb[long(t) * T * T + q], bf[t], v);
    pk = fmaxf(pk, v);
    mn = fminf(mn, v);
    if (want_cov && on) {
      const double vp = v > 0.0f ? double(v) : 0.0;
      const int yy = iy - h + r, xx = ix - h + c;
      const bool miss = yy < 0 || yy >= H || xx < 0 || xx >= W || nd[long(yy) * W + xx] != 0;
      tot += vp;
      if (miss) {
        cov += vp;
        hit = 1;
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    pk = fmaxf(pk, __shfl_xor(pk, off, 64));
    mn = fminf(mn, __shfl_xor(mn, off, 64));
    hit |= __shfl_xor(hit, off, 64);
  }
  tot = wave_sum(tot);
  cov = wave_sum(cov);
  if (lane == 0) {
    peak[out] = pk;
    vmin[out] = mn;
    double u = 0.0;
    if (want_cov) {
      if (!on)
        u = 1.0;  // a position off the image
      else if (hit)
        u = tot > 0.0 ? cov / tot : 1.0;
    }
    uncovered[out] = u;
  }
}

}  // namespace uph

extern "C" int uph_band_psf_flags(int n, int n_bands, int n_terms, int T, const int32_t* size,
                                  const float* coefs, const double* geo, const double* x,
                                  const double* y, int H, int W, const uint8_t* nodata,
                                  const uint8_t* has_nodata, float* peak, float* vmin,
                                  double* uncovered, void* stream) {
  UPH_CHECK_ARG(n >= 0 && n_bands >= 1 && n_terms >= 1 && n_terms <= 21 && T >= 1 && H >= 0 &&
                    W >= 0,
                "uph_band_psf_flags: n=%d n_bands=%d n_terms=%d T=%d image %d x %d", n, n_bands,
                n_terms, T, H, W);
  if (n == 0) return UPH_OK;
  UPH_CHECK_ARG(size && coefs && geo && x && y && peak && vmin && uncovered,
                "uph_band_psf_flags: null argument");
  UPH_CHECK_ARG((nodata == nullptr) == (has_nodata == nullptr),
                "uph_band_psf_flags: nodata and has_nodata go together");
  const long items = long(n) * n_bands;
  const long blocks = (items + uph::kFlagWaves - 1) / uph::kFlagWaves;
  UPH_CHECK_ARG(blocks <= 0x7fffffffL, "uph_band_psf_flags: %ld items exceed one launch", items);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(uph::psf_flags_kernel, dim3(unsigned(blocks)), dim3(UPH_BLOCK), 0, s, n,
                     n_bands, n_terms, T, size, coefs, geo, x, y, H, W, nodata, has_nodata, peak,
                     vmin, uncovered);
  return uph::hip_status(hipGetLastError(), "uph_band_psf_flags launch");
}
