// Direct render backend of the unit-flux stamp renderer
// (tract7dt/fit/grid_gpu/source_model.py:218-256 and the "direct" branch of
// render_source_stamp_at, :298-303): STAR = the PSF polynomial at (x, y);
// EXP = the pixel-centre 6-Gaussian exponential profile (unit sum) convolved
// directly with the PSF stamp ('same' mode) and renormalised to unit box
// sum; both shifted by the separable Lanczos-3 filter.  DEV and SERSIC have
// no direct form (the host rejects them).
//
// One workgroup per stamp (grid-stride when the stamps live in scratch).
// The PSF, the profile and the convolution sit in LDS up to T = 63 (3 T^2
// floats = 47.6 KB); larger stamps use the workgroup's slice of `scratch`.
// The convolution is T^4 multiply-adds per EXP stamp (9.2 M at T = 55),
// VALU-bound; each thread owns output pixels and sums over the PSF rows
// that overlap its row.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"

namespace uph {

constexpr int kDirectLdsMaxT = 63;

// Exponential-profile mixture (amplitudes normalised to unit sum in float32
// and variances in units of Re^2): Tractor's mixture_profiles constants as
// the reference ships them (source_model.py:59-75).
__constant__ const float kExpAmpRaw[6] = {2.34853813e-03f, 3.07995260e-02f, 2.23364214e-01f,
                                          1.17949102e00f,  4.33873750e00f,  5.99820770e00f};
__constant__ const float kExpVar[6] = {1.20078965e-03f, 8.84526493e-03f, 3.91463084e-02f,
                                       1.39976817e-01f, 4.60962500e-01f, 1.50159566e00f};

__device__ __forceinline__ float exp_amp(int k) {
  float s = 0.0f;
  for (int i = 0; i < 6; ++i) s += kExpAmpRaw[i];
  return kExpAmpRaw[k] / s;
}

__global__ __launch_bounds__(UPH_BLOCK) void direct_stamps_kernel(
    int n, int T, int nt, const float* __restrict__ coefs, const float* __restrict__ geo_all,
    const int32_t* __restrict__ src_band, const float* __restrict__ xs_,
    const float* __restrict__ ys_, const float* __restrict__ gx0, const float* __restrict__ gy0,
    const int32_t* __restrict__ stype, const float* __restrict__ logre,
    const float* __restrict__ ee1, const float* __restrict__ ee2, float* __restrict__ out,
    float* __restrict__ scratch) {
  extern __shared__ float lds[];
  __shared__ float taps[14];
  __shared__ float bas[21];
  __shared__ float red[UPH_WAVES];
  __shared__ float cov[6][3];  // per component inverse covariance terms
  __shared__ float nrm[6];
  const int tid = threadIdx.x, TT = T * T, h = T / 2;
  float* P = scratch ? scratch + size_t(blockIdx.x) * 3 * TT : lds;
  float* G = P + TT;
  float* R = G + TT;
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    const int b = src_band[s];
    const float* geo = geo_all + 4 * b;
    const float x = xs_[s], y = ys_[s];
    const float xg = x - gx0[s], yg = y - gy0[s];
    const float dx = xg - rintf(xg), dy = yg - rintf(yg);
    const bool is_exp = stype && stype[s] == 1;
    if (tid == 0) lanczos_taps(dx, taps, (float*)nullptr);
    if (tid == 64) lanczos_taps(dy, taps + 7, (float*)nullptr);
    if (tid == 128) {
      const float X = (x - geo[0]) / geo[1], Y = (y - geo[2]) / geo[3];
      psf_basis(X, Y, nt, bas, (float*)nullptr, (float*)nullptr);
    }
    if (tid == 192 && is_exp) {
      // ellipse of the profile (render_exp_profile_stamp_jax, float32)
      const float re = fminf(fmaxf(expf(logre[s]), 0.1f), 1000.0f);
      const float e1 = ee1[s], e2 = ee2[s];
      const float ee = hypotf(e1, e2);
      const float e = 1.0f - expf(-ee);
      const float inv_ba = fminf((1.0f + e) / fmaxf(1.0f - e, 1e-4f), 1000.0f);
      const float th = 0.5f * atan2f(e2, e1);
      const float ct = cosf(th), st = sinf(th);
      const float a = re * ct / inv_ba, bb = re * st, c = -re * st / inv_ba, d = re * ct;
      for (int k = 0; k < 6; ++k) {
        const float v = kExpVar[k];
        const float c00 = v * (a * a + bb * bb), c01 = v * (a * c + bb * d),
                    c11 = v * (c * c + d * d);
        const float det = fmaxf(c00 * c11 - c01 * c01, 1e-12f);
        cov[k][0] = c11 / det;
        cov[k][1] = -c01 / det;
        cov[k][2] = c00 / det;
        nrm[k] = exp_amp(k) / (2.0f * 3.14159265358979f * sqrtf(det));
      }
    }
    __syncthreads();
    const float* C = coefs + size_t(b) * nt * TT;
    for (int e = tid; e < TT; e += UPH_BLOCK) {
      float v = 0.0f;
      for (int t = 0; t < nt; ++t) v = fmaf(C[size_t(t) * TT + e], bas[t], v);
      P[e] = v;
    }
    float* raw = P;
    if (is_exp) {
      // intrinsic profile, normalised to unit sum where positive
      float part = 0.0f;
      for (int e = tid; e < TT; e += UPH_BLOCK) {
        const float yy = float(e / T - h), xx = float(e % T - h);
        float v = 0.0f;
        for (int k = 0; k < 6; ++k) {
          const float r2 = cov[k][0] * xx * xx + 2.0f * cov[k][1] * xx * yy + cov[k][2] * yy * yy;
          v += nrm[k] * expf(-0.5f * fminf(r2, 200.0f));
        }
        G[e] = v;
        part += v;
      }
      const float tot = block_sum(part, red);  // synchronises G and P
      if (tot > 0.0f)
        for (int e = tid; e < TT; e += UPH_BLOCK) G[e] = G[e] / tot;
      __syncthreads();
      // 'same' convolution: R[i, j] = sum_k,l G[k, l] P[i + h - k, j + h - l]
      float csum = 0.0f;
      for (int e = tid; e < TT; e += UPH_BLOCK) {
        const int i = e / T, j = e % T;
        const int k0 = max(0, i + h - T + 1), k1 = min(T - 1, i + h);
        const int l0 = max(0, j + h - T + 1), l1 = min(T - 1, j + h);
        float v = 0.0f;
        for (int k = k0; k <= k1; ++k) {
          const float* gr = G + k * T;
          const float* pr = P + (i + h - k) * T + (j + h);
          for (int l = l0; l <= l1; ++l) v = fmaf(gr[l], pr[-l], v);
        }
        R[e] = v;
        csum += v;
      }
      const float ctot = block_sum(csum, red);
      if (ctot > 0.0f)
        for (int e = tid; e < TT; e += UPH_BLOCK) R[e] = R[e] / ctot;
      raw = R;
    }
    __syncthreads();
    float Lx[7], Ly[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      Lx[k] = taps[k];
      Ly[k] = taps[7 + k];
    }
    // horizontal pass into the free buffer, vertical pass to the output
    float* hbuf = is_exp ? G : R;
    for (int e = tid; e < TT; e += UPH_BLOCK) {
      const int r = e / T, c = e % T;
      float v = 0.0f;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        const int cc = c + k - 3;
        if (cc >= 0 && cc < T) v = fmaf(Lx[k], raw[r * T + cc], v);
      }
      hbuf[e] = v;
    }
    __syncthreads();
    float* o = out + size_t(s) * TT;
    for (int e = tid; e < TT; e += UPH_BLOCK) {
      const int r = e / T, c = e % T;
      float v = 0.0f;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        const int rr = r + k - 3;
        if (rr >= 0 && rr < T) v = fmaf(Ly[k], hbuf[rr * T + c], v);
      }
      o[e] = v;
    }
    __syncthreads();
  }
}

}  // namespace uph

extern "C" size_t uph_render_direct_scratch_bytes(int T, int n_blocks) {
  if (T <= uph::kDirectLdsMaxT || T <= 0 || n_blocks <= 0) return 0;
  return size_t(3) * T * T * sizeof(float) * size_t(n_blocks);
}

extern "C" int uph_render_stamps_direct(int n, int T, int n_terms, const float* psf_coefs,
                                        const float* psf_geo, const int32_t* src_band,
                                        const float* x, const float* y, const float* gx0,
                                        const float* gy0, const int32_t* stype,
                                        const float* logre, const float* ee1, const float* ee2,
                                        float* out, void* scratch, size_t scratch_bytes,
                                        void* stream) {
  UPH_CHECK_ARG(n >= 0 && T >= 1 && (T & 1) && n_terms >= 1 && n_terms <= 21,
                "uph_render_stamps_direct: n=%d T=%d (odd) n_terms=%d", n, T, n_terms);
  if (n == 0) return UPH_OK;
  UPH_CHECK_ARG(psf_coefs && psf_geo && src_band && x && y && gx0 && gy0 && out,
                "uph_render_stamps_direct: null argument");
  UPH_CHECK_ARG(stype == nullptr || (logre && ee1 && ee2),
                "uph_render_stamps_direct: stype given without the EXP shapes");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (T <= uph::kDirectLdsMaxT) {
    const size_t lds = size_t(3) * T * T * sizeof(float);
    hipLaunchKernelGGL(uph::direct_stamps_kernel, dim3(n), dim3(UPH_BLOCK), lds, s, n, T,
                       n_terms, psf_coefs, psf_geo, src_band, x, y, gx0, gy0, stype, logre, ee1,
                       ee2, out, nullptr);
  } else {
    const size_t per = uph_render_direct_scratch_bytes(T, 1);
    UPH_CHECK_ARG(scratch && scratch_bytes >= per,
                  "uph_render_stamps_direct: T=%d needs scratch of %zu bytes per workgroup, "
                  "got %zu",
                  T, per, scratch_bytes);
    const size_t blocks = std::min<size_t>(size_t(n), scratch_bytes / per);
    hipLaunchKernelGGL(uph::direct_stamps_kernel, dim3(unsigned(blocks)), dim3(UPH_BLOCK), 0, s,
                       n, T, n_terms, psf_coefs, psf_geo, src_band, x, y, gx0, gy0, stype, logre,
                       ee1, ee2, out, static_cast<float*>(scratch));
  }
  return uph::hip_status(hipGetLastError(), "uph_render_stamps_direct launch");
}
