// Extended-source (EXP/DEV/SERSIC) rendering: Sersic mixture lookup,
// ellipse covariance terms (forward-mode derivatives), block-cooperative
// radix-2 FFTs, and the analytic-Fourier-transform galaxy stamp
//   G = irfft2(rfft2(pad(PSF)) * F)[:T, :T] + core * PSF,
//   F(v, w) = sum_k a_k exp(-2 pi^2 var_k (cxx v^2 + cyy w^2 + 2 cxy v w))
// on the (pad, pad/2 + 1) frequency grid, pad = smallest power of two
// >= T + T//2.  Algorithm followed: tract7dt/fit/grid_gpu/sersic_mog.py
// (mog_params_at :87-100, _ellipse_cov_terms :108-150,
// render_mog_stamp_fft :153-194, fft_pad_for :197-207) and
// source_model.py:264-328.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "uph_device.hpp"
#include "uph_island.hpp"

namespace uph {

constexpr double kNMin = 0.29, kNMax = 6.3;  // Sersic index sigmoid range

__host__ __device__ inline int fft_pad_for(int T) {
  int n = 1;
  while (n < T + T / 2) n <<= 1;
  return n;
}

__host__ __device__ inline int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return l;
}

// Galaxy workspace layout (elements of Real).
struct GalLayout {
  size_t P, Px, Py, G, Hh, SP, SPx, SPy, Z, RB, F, total;
  int pad, H;
};

__host__ __device__ inline GalLayout galaxy_layout(int T) {
  GalLayout L;
  L.pad = fft_pad_for(T);
  L.H = L.pad / 2 + 1;
  const size_t TT = (size_t)T * T, spec = (size_t)L.pad * L.H * 2;
  size_t o = 0;
  L.P = o; o += TT;
  L.Px = o; o += TT;
  L.Py = o; o += TT;
  L.G = o; o += 7 * TT;   // G, then tangents of x, y, logre, ee1, ee2, n
  L.Hh = o; o += 7 * TT;  // horizontal Lanczos pass of the above
  L.SP = o; o += spec;
  L.SPx = o; o += spec;
  L.SPy = o; o += spec;
  L.Z = o; o += spec;
  L.RB = o; o += (size_t)T * L.pad * 2;
  L.F = o; o += (size_t)3 * L.pad * L.H;  // F, F1 = sum a v E, Fn
  L.total = o;
  return L;
}

__host__ __device__ inline size_t galaxy_scratch_bytes(int T, size_t real_bytes) {
  return align256(galaxy_layout(T).total * real_bytes);
}

// ---------------------------------------------------------------------------
// Sersic table block (float64): [n_knots, n_min, n_max, K, knots (n_knots),
// amps (n_knots, K), log_vars (n_knots, K), core (n_knots)].  Linear
// interpolation in the segment found by searchsorted(side="right") - 1.
// ---------------------------------------------------------------------------
__device__ inline void mog_lookup(const double* tab, double n, double* a, double* v, double& core,
                                  double* da, double* dv, double& dcore) {
  const int nk = int(tab[0]), K = int(tab[3]);
  const double nmin = tab[1], nmax = tab[2];
  const double* kn = tab + 4;
  const double* A = kn + nk;
  const double* LV = A + (size_t)nk * K;
  const double* CO = LV + (size_t)nk * K;
  const double nn = fmin(fmax(n, nmin), nmax);
  const bool clipped = nn != n;
  int lo = 0, hi = nk;  // first index with kn > nn
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (kn[mid] <= nn)
      lo = mid + 1;
    else
      hi = mid;
  }
  int i = lo - 1;
  i = i < 0 ? 0 : (i > nk - 2 ? nk - 2 : i);
  const double n0 = kn[i], n1 = kn[i + 1];
  const double t = n1 > n0 ? (nn - n0) / (n1 - n0) : 0.0;
  const double dt = (!clipped && n1 > n0) ? 1.0 / (n1 - n0) : 0.0;
  for (int j = 0; j < K; ++j) {
    const double a0 = A[(size_t)i * K + j], a1 = A[(size_t)(i + 1) * K + j];
    const double l0 = LV[(size_t)i * K + j], l1 = LV[(size_t)(i + 1) * K + j];
    a[j] = (1.0 - t) * a0 + t * a1;
    v[j] = exp((1.0 - t) * l0 + t * l1);
    da[j] = dt * (a1 - a0);
    dv[j] = v[j] * dt * (l1 - l0);
  }
  core = (1.0 - t) * CO[i] + t * CO[i + 1];
  dcore = dt * (CO[i + 1] - CO[i]);
}

// Forward-mode dual number in (logre, ee1, ee2).
struct Dual3 {
  double v, d[3];
};
__device__ inline Dual3 dconst(double c) { return Dual3{c, {0.0, 0.0, 0.0}}; }
__device__ inline Dual3 operator+(Dual3 a, Dual3 b) {
  return Dual3{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__device__ inline Dual3 operator-(Dual3 a, Dual3 b) {
  return Dual3{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__device__ inline Dual3 operator*(Dual3 a, Dual3 b) {
  return Dual3{a.v * b.v,
               {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1],
                a.d[2] * b.v + a.v * b.d[2]}};
}
__device__ inline Dual3 operator/(Dual3 a, Dual3 b) {
  const double q = a.v / b.v, ib = 1.0 / b.v;
  return Dual3{q,
               {(a.d[0] - q * b.d[0]) * ib, (a.d[1] - q * b.d[1]) * ib,
                (a.d[2] - q * b.d[2]) * ib}};
}
__device__ inline Dual3 dexp(Dual3 a) {
  const double e = exp(a.v);
  return Dual3{e, {e * a.d[0], e * a.d[1], e * a.d[2]}};
}
__device__ inline Dual3 dsqrt(Dual3 a) {
  const double s = sqrt(a.v), h = 0.5 / s;
  return Dual3{s, {h * a.d[0], h * a.d[1], h * a.d[2]}};
}

// (cxx, cyy, cxy) of EllipseESoft with their derivatives in (logre, ee1, ee2);
// clips (Re to [0.1, 1000], axis-ratio inverse to <= 1000) freeze the
// clipped quantity.  dt[q][p] = d term_q / d param_p.
__device__ inline void ellipse_terms(double lr, double e1, double e2, double* t, double (*dt)[3]) {
  if (e1 == 0.0 && e2 == 0.0) e1 = 1e-4;
  const Dual3 L{lr, {1.0, 0.0, 0.0}}, E1{e1, {0.0, 1.0, 0.0}}, E2{e2, {0.0, 0.0, 1.0}};
  Dual3 re = dexp(L);
  if (re.v <= 0.1 || re.v >= 1000.0) re = dconst(fmin(fmax(re.v, 0.1), 1000.0));
  const Dual3 ee = dsqrt(E1 * E1 + E2 * E2);
  const Dual3 e = dconst(1.0) - dexp(dconst(0.0) - ee);
  Dual3 den = dconst(1.0) - e;
  if (den.v <= 1e-4) den = dconst(1e-4);
  Dual3 ib = (dconst(1.0) + e) / den;
  if (ib.v >= 1000.0) ib = dconst(1000.0);
  const Dual3 s2 = dconst(1.0) / (ib * ib);
  const Dual3 C = E1 / ee, S = E2 / ee;
  const Dual3 h = re * re * dconst(0.5);
  const Dual3 up = dconst(1.0) + s2, dn = dconst(1.0) - s2;
  const Dual3 cxx = h * (up - C * dn), cyy = h * (up + C * dn), cxy = h * S * dn;
  t[0] = cxx.v;
  t[1] = cyy.v;
  t[2] = cxy.v;
  for (int p = 0; p < 3; ++p) {
    dt[0][p] = cxx.d[p];
    dt[1][p] = cyy.d[p];
    dt[2][p] = cxy.d[p];
  }
}

// ---------------------------------------------------------------------------
// Block-cooperative in-place radix-2 FFT of nl lines of length n (complex
// Real pairs).  Element i of line l sits at buf[2 * (l * ls + i * es)].
// sign = -1 forward, +1 inverse (unnormalised).
// ---------------------------------------------------------------------------
template <typename Real>
__device__ __forceinline__ void sincospi_r(Real x, Real* s, Real* c) {
  if constexpr (sizeof(Real) == 4)
    sincospif(x, s, c);
  else
    sincospi(x, s, c);
}

// LDS capacity of the line-batch FFT (complex elements) and the largest
// transform length it serves.
constexpr int kFftLdsComplex = 6144;
constexpr int kFftMaxN = 512;

template <typename Real>
__device__ void fft_lines_lds(Real* buf, int n, int logn, int nl, int ls, int es, int sign) {
  // Lines are processed in batches that fit LDS: load (bit-reversed),
  // log2(n) radix-2 stages in LDS with a shared twiddle table, store.
  constexpr int CAP = kFftLdsComplex * 4 / (int)sizeof(Real);
  __shared__ Real lb[2 * CAP];
  __shared__ Real tw[kFftMaxN];  // n/2 complex twiddles exp(sign 2 pi i j / n)
  const int tid = threadIdx.x;
  for (int j = tid; j < n / 2; j += UPH_BLOCK) {
    Real sn, cs;
    sincospi_r(Real(2 * j) / Real(n), &sn, &cs);
    tw[2 * j] = cs;
    tw[2 * j + 1] = Real(sign) * sn;
  }
  const int chunk = CAP / n;
  const bool col = (ls == 1);  // adjacent lines adjacent in memory
  for (int l0 = 0; l0 < nl; l0 += chunk) {
    const int lc = min(chunk, nl - l0);
    for (int e = tid; e < lc * n; e += UPH_BLOCK) {
      const int line = col ? e % lc : e / n;
      const int i = col ? e / lc : e % n;
      const int j = (int)(__brev((unsigned)i) >> (32 - logn));
      const Real* src = buf + 2 * ((size_t)(l0 + line) * ls + (size_t)i * es);
      lb[2 * (line * n + j)] = src[0];
      lb[2 * (line * n + j) + 1] = src[1];
    }
    __syncthreads();
    const int nh = n >> 1;
    for (int s = 1; s <= logn; ++s) {
      const int half = 1 << (s - 1);
      const int tstep = n >> s;
      for (int e = tid; e < lc * nh; e += UPH_BLOCK) {
        const int line = e / nh, k = e % nh;
        const int j = k & (half - 1);
        const int i0 = line * n + ((k >> (s - 1)) << s) + j, i1 = i0 + half;
        const Real wr = tw[2 * j * tstep], wi = tw[2 * j * tstep + 1];
        const Real qr = lb[2 * i1], qi = lb[2 * i1 + 1];
        const Real br = wr * qr - wi * qi, bi = wr * qi + wi * qr;
        const Real ar = lb[2 * i0], ai = lb[2 * i0 + 1];
        lb[2 * i0] = ar + br;
        lb[2 * i0 + 1] = ai + bi;
        lb[2 * i1] = ar - br;
        lb[2 * i1 + 1] = ai - bi;
      }
      __syncthreads();
    }
    for (int e = tid; e < lc * n; e += UPH_BLOCK) {
      const int line = col ? e % lc : e / n;
      const int i = col ? e / lc : e % n;
      Real* dst = buf + 2 * ((size_t)(l0 + line) * ls + (size_t)i * es);
      dst[0] = lb[2 * (line * n + i)];
      dst[1] = lb[2 * (line * n + i) + 1];
    }
    __syncthreads();
  }
}

// Block-cooperative radix-2 FFT working directly on the (global) buffer:
// in-place bit reversal, then log2(n) stages with a barrier each.
template <typename Real>
__device__ void fft_lines_block(Real* buf, int n, int logn, int nl, int ls, int es, int sign) {
  const int tid = threadIdx.x;
  for (int e = tid; e < nl * n; e += UPH_BLOCK) {
    const int l = e / n, i = e % n;
    const int j = (int)(__brev((unsigned)i) >> (32 - logn));
    if (i < j) {
      Real* p = buf + 2 * ((size_t)l * ls + (size_t)i * es);
      Real* q = buf + 2 * ((size_t)l * ls + (size_t)j * es);
      const Real a0 = p[0], a1 = p[1];
      p[0] = q[0];
      p[1] = q[1];
      q[0] = a0;
      q[1] = a1;
    }
  }
  __syncthreads();
  const int nh = n >> 1;
  for (int s = 1; s <= logn; ++s) {
    const int half = 1 << (s - 1);
    for (int e = tid; e < nl * nh; e += UPH_BLOCK) {
      const int l = e / nh, k = e % nh;
      const int j = k & (half - 1);
      const int i0 = ((k >> (s - 1)) << s) + j, i1 = i0 + half;
      Real sn, cs;
      sincospi_r(Real(j) / Real(half), &sn, &cs);
      const Real wr = cs, wi = Real(sign) * sn;
      Real* p = buf + 2 * ((size_t)l * ls + (size_t)i0 * es);
      Real* q = buf + 2 * ((size_t)l * ls + (size_t)i1 * es);
      const Real br = wr * q[0] - wi * q[1], bi = wr * q[1] + wi * q[0];
      const Real ar = p[0], ai = p[1];
      p[0] = ar + br;
      p[1] = ai + bi;
      q[0] = ar - br;
      q[1] = ai - bi;
    }
    __syncthreads();
  }
}

// Wave-level FFT of length n = 64 E: each wavefront owns G lines at a
// time, lane l holding elements l + 64 e (e < E) of each in registers.
// Bit-reversed gather on load, then decimation-in-time stages: spans 1..32
// across lanes (shuffles), spans 64.. inside the lane; natural-order store.
// The G lines' loads are issued together (one memory latency per batch,
// and adjacent lines of a column pass share cache lines), then the G
// transforms interleave.  No LDS allocation and no barrier inside a batch;
// the 4 waves take batches round-robin.  Twiddles depend only on the lane
// (and e for in-lane spans) and are computed once per call.
template <typename Real, int E, int G>
__device__ void fft_lines_wave(Real* buf, int nl, int ls, int es, int sign) {
  constexpr int LOGE = E == 1 ? 0 : (E == 2 ? 1 : (E == 4 ? 2 : 3));
  constexpr int LOGN = 6 + LOGE;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  Real cw[6][2];            // cross-lane stage twiddles (span h = 1 .. 32)
  Real iw[LOGE + 1][E][2];  // in-lane stage twiddles (span 64 he)
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int h = 1 << s;
    Real sn, cs;
    sincospi_r(Real(lane & (h - 1)) / Real(h), &sn, &cs);
    cw[s][0] = cs;
    cw[s][1] = Real(sign) * sn;
  }
#pragma unroll
  for (int t = 0; t < LOGE; ++t) {
    const int he = 1 << t;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = lane + 64 * (e & (he - 1));
      Real sn, cs;
      sincospi_r(Real(j) / Real(64 * he), &sn, &cs);
      iw[t][e][0] = cs;
      iw[t][e][1] = Real(sign) * sn;
    }
  }
  for (int l0 = wid * G; l0 < nl; l0 += UPH_WAVES * G) {
    Real re[G][E], im[G][E];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (l0 + g < nl) {  // wave-uniform
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = lane + 64 * e;
          const int src = (int)(__brev((unsigned)i) >> (32 - LOGN));
          const Real* p = buf + 2 * ((size_t)(l0 + g) * ls + (size_t)src * es);
          re[g][e] = p[0];
          im[g][e] = p[1];
        }
      }
    }
#pragma unroll
    for (int s = 0; s < 6; ++s) {
      const int h = 1 << s;
      const Real wr = cw[s][0], wi = cw[s][1];
      const bool upper = lane & h;
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (l0 + g >= nl) continue;  // wave-uniform
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const Real pr = __shfl_xor(re[g][e], h, 64), pim = __shfl_xor(im[g][e], h, 64);
          if (!upper) {
            re[g][e] = re[g][e] + (wr * pr - wi * pim);
            im[g][e] = im[g][e] + (wr * pim + wi * pr);
          } else {
            const Real tr = wr * re[g][e] - wi * im[g][e], ti = wr * im[g][e] + wi * re[g][e];
            re[g][e] = pr - tr;
            im[g][e] = pim - ti;
          }
        }
      }
    }
#pragma unroll
    for (int t = 0; t < LOGE; ++t) {
      const int he = 1 << t;
#pragma unroll
      for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          if (e & he) continue;
          const int f = e + he;
          const Real wr = iw[t][e][0], wi = iw[t][e][1];
          const Real br = wr * re[g][f] - wi * im[g][f], bi = wr * im[g][f] + wi * re[g][f];
          const Real ar = re[g][e], ai = im[g][e];
          re[g][e] = ar + br;
          im[g][e] = ai + bi;
          re[g][f] = ar - br;
          im[g][f] = ai - bi;
        }
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (l0 + g < nl) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
          const int i = lane + 64 * e;
          Real* p = buf + 2 * ((size_t)(l0 + g) * ls + (size_t)i * es);
          p[0] = re[g][e];
          p[1] = im[g][e];
        }
      }
    }
  }
  __syncthreads();
}

// In-place FFT of nl lines (natural order in and out).  kLds selects the
// LDS line-batch variant (used by the grid stamp renderer); otherwise the
// wave-level register FFT (n = 128, 256, 512) or the block FFT.
template <typename Real, bool kLds>
__device__ void fft_lines(Real* buf, int n, int logn, int nl, int ls, int es, int sign) {
  if constexpr (kLds) {
    fft_lines_lds(buf, n, logn, nl, ls, es, sign);
  } else {
    __syncthreads();  // lines may have been written by other waves
    if (n == 128)
      fft_lines_wave<Real, 2, 8>(buf, nl, ls, es, sign);
    else if (n == 256)
      fft_lines_wave<Real, 4, 4>(buf, nl, ls, es, sign);
    else if (n == 512)
      fft_lines_wave<Real, 8, 2>(buf, nl, ls, es, sign);
    else
      fft_lines_block(buf, n, logn, nl, ls, es, sign);
  }
}

// rfft2 of the (T, T) real plane R zero-padded to (pad, pad): spectrum S of
// (pad, H = pad/2 + 1) complex, row-major.  RB: (T, pad) complex scratch.
template <typename Real, bool kLds>
__device__ void rfft2_pad(const Real* R, int T, int pad, int logp, Real* RB, Real* S) {
  const int tid = threadIdx.x, H = pad / 2 + 1;
  for (int e = tid; e < T * pad; e += UPH_BLOCK) {
    const int r = e / pad, c = e % pad;
    RB[2 * e] = c < T ? R[r * T + c] : Real(0);
    RB[2 * e + 1] = Real(0);
  }
  __syncthreads();
  fft_lines<Real, kLds>(RB, pad, logp, T, pad, 1, -1);
  for (int e = tid; e < pad * H; e += UPH_BLOCK) {
    const int r = e / H, k = e % H;
    S[2 * e] = r < T ? RB[2 * (r * pad + k)] : Real(0);
    S[2 * e + 1] = r < T ? RB[2 * (r * pad + k) + 1] : Real(0);
  }
  __syncthreads();
  fft_lines<Real, kLds>(S, pad, logp, H, 1, H, -1);
}

// Real output O (T, T) = irfft2(Z, (pad, pad))[:T, :T]; Z is destroyed.
template <typename Real, bool kLds>
__device__ void irfft2_crop(Real* Z, int T, int pad, int logp, Real* RB, Real* O) {
  const int tid = threadIdx.x, H = pad / 2 + 1;
  fft_lines<Real, kLds>(Z, pad, logp, H, 1, H, +1);
  for (int e = tid; e < T * pad; e += UPH_BLOCK) {
    const int r = e / pad, c = e % pad;
    Real re, im;
    if (c < H) {
      re = Z[2 * (r * H + c)];
      im = Z[2 * (r * H + c) + 1];
    } else {
      re = Z[2 * (r * H + pad - c)];
      im = -Z[2 * (r * H + pad - c) + 1];
    }
    RB[2 * e] = re;
    RB[2 * e + 1] = im;
  }
  __syncthreads();
  fft_lines<Real, kLds>(RB, pad, logp, T, pad, 1, +1);
  const Real norm = Real(1) / (Real(pad) * Real(pad));
  for (int e = tid; e < T * T; e += UPH_BLOCK) {
    const int r = e / T, c = e % T;
    O[e] = RB[2 * (r * pad + c)] * norm;
  }
  __syncthreads();
}

// Z = S * W, W real on the (pad, H) grid.
template <typename Real>
__device__ void spec_scale(const Real* S, const Real* W, int n, Real* Z) {
  for (int e = threadIdx.x; e < n; e += UPH_BLOCK) {
    Z[2 * e] = S[2 * e] * W[e];
    Z[2 * e + 1] = S[2 * e + 1] * W[e];
  }
  __syncthreads();
}

// Per-member galaxy parameters shared by the block (set by one thread).
template <typename Real>
struct GalParams {
  double a[15], v[15], da[15], dv[15], core, dcore, t[3], dt[3][3], n, dndu;
  int K;
};

// F, F1 = sum a v E and Fn = sum (da E + a E expo dv) on the (pad, H) grid.
template <typename Real>
__device__ void fill_fourier(const GalParams<Real>& gp, int pad, bool shape, bool nder, Real* F) {
  const int H = pad / 2 + 1, n = pad * H;
  const double c2 = -2.0 * 9.869604401089358;  // -2 pi^2
  for (int e = threadIdx.x; e < n; e += UPH_BLOCK) {
    const int m = e / H, k = e % H;
    const double v = double(k) / pad;
    const double w = double(m < pad / 2 ? m : m - pad) / pad;
    const double ex = c2 * (gp.t[0] * v * v + gp.t[1] * w * w + 2.0 * gp.t[2] * v * w);
    double f = 0.0, f1 = 0.0, fn = 0.0;
    for (int j = 0; j < gp.K; ++j) {
      const double E = exp(gp.v[j] * ex);
      f += gp.a[j] * E;
      f1 += gp.a[j] * gp.v[j] * E;
      fn += gp.da[j] * E + gp.a[j] * E * ex * gp.dv[j];
    }
    F[e] = Real(f);
    if (shape) F[n + e] = Real(f1);
    if (nder) F[2 * n + e] = Real(fn);
  }
  __syncthreads();
}

// Z = S * F1 * d(expo)/d(param p), param p in {logre, ee1, ee2}.
template <typename Real>
__device__ void spec_shape(const Real* S, const Real* F1, const GalParams<Real>& gp, int p,
                           int pad, Real* Z) {
  const int H = pad / 2 + 1, n = pad * H;
  const double c2 = -2.0 * 9.869604401089358;
  for (int e = threadIdx.x; e < n; e += UPH_BLOCK) {
    const int m = e / H, k = e % H;
    const double v = double(k) / pad;
    const double w = double(m < pad / 2 ? m : m - pad) / pad;
    const double dex =
        c2 * (gp.dt[0][p] * v * v + gp.dt[1][p] * w * w + 2.0 * gp.dt[2][p] * v * w);
    const Real s = Real(double(F1[e]) * dex);
    Z[2 * e] = S[2 * e] * s;
    Z[2 * e + 1] = S[2 * e + 1] * s;
  }
  __syncthreads();
}

// Raw (unshifted) galaxy stamp and the raw tangent planes requested in
// `need` (bit 0 x, 1 y, 2 logre, 3 ee1, 4 ee2, 5 n).  Planes in ws at
// GalLayout offsets: G + 0 (model), G + (1 + j) T^2 for parameter j.
// Requires P (and Px, Py for x/y) already in ws and gp filled.
template <typename Real, bool kLds>
__device__ void galaxy_raw(const GalParams<Real>& gp, int T, int need, Real* ws) {
  const GalLayout L = galaxy_layout(T);
  const int pad = L.pad, logp = ilog2(pad), n = pad * L.H, TT = T * T;
  const bool shape = need & (4 | 8 | 16), nder = need & 32;
  fill_fourier(gp, pad, shape, nder, ws + L.F);
  rfft2_pad<Real, kLds>(ws + L.P, T, pad, logp, ws + L.RB, ws + L.SP);
  Real* Gm = ws + L.G;
  spec_scale(ws + L.SP, ws + L.F, n, ws + L.Z);
  irfft2_crop<Real, kLds>(ws + L.Z, T, pad, logp, ws + L.RB, Gm);
  const Real core = Real(gp.core);
  for (int e = threadIdx.x; e < TT; e += UPH_BLOCK) Gm[e] += core * ws[L.P + e];
  __syncthreads();
  for (int j = 0; j < 2; ++j) {
    if (!(need & (1 << j))) continue;
    Real* src = ws + (j == 0 ? L.Px : L.Py);
    Real* spec = ws + (j == 0 ? L.SPx : L.SPy);
    rfft2_pad<Real, kLds>(src, T, pad, logp, ws + L.RB, spec);
    spec_scale(spec, ws + L.F, n, ws + L.Z);
    Real* o = Gm + (size_t)(1 + j) * TT;
    irfft2_crop<Real, kLds>(ws + L.Z, T, pad, logp, ws + L.RB, o);
    for (int e = threadIdx.x; e < TT; e += UPH_BLOCK) o[e] += core * src[e];
    __syncthreads();
  }
  for (int p = 0; p < 3; ++p) {
    if (!(need & (4 << p))) continue;
    spec_shape(ws + L.SP, ws + L.F + n, gp, p, pad, ws + L.Z);
    irfft2_crop<Real, kLds>(ws + L.Z, T, pad, logp, ws + L.RB, Gm + (size_t)(3 + p) * TT);
  }
  if (nder) {
    spec_scale(ws + L.SP, ws + L.F + 2 * n, n, ws + L.Z);
    Real* o = Gm + (size_t)6 * TT;
    irfft2_crop<Real, kLds>(ws + L.Z, T, pad, logp, ws + L.RB, o);
    const Real dcore = Real(gp.dcore);
    for (int e = threadIdx.x; e < TT; e += UPH_BLOCK) o[e] += dcore * ws[L.P + e];
    __syncthreads();
  }
}

// Z = S * F1 * d(expo)/d(param p) with the ellipse-term derivatives given
// as dt9[3 i + p] (the layout of GalParams::dt).
template <typename Real>
__device__ void spec_shape_dt(const Real* S, const Real* F1, const double* dt9, int p, int pad,
                              Real* Z) {
  const int H = pad / 2 + 1, n = pad * H;
  const double c2 = -2.0 * 9.869604401089358;
  const double d0 = dt9[p], d1 = dt9[3 + p], d2 = dt9[6 + p];
  for (int e = threadIdx.x; e < n; e += UPH_BLOCK) {
    const int m = e / H, k = e % H;
    const double v = double(k) / pad;
    const double w = double(m < pad / 2 ? m : m - pad) / pad;
    const double dex = c2 * (d0 * v * v + d1 * w * w + 2.0 * d2 * v * w);
    const Real s = Real(double(F1[e]) * dex);
    Z[2 * e] = S[2 * e] * s;
    Z[2 * e + 1] = S[2 * e + 1] * s;
  }
  __syncthreads();
}

// galaxy_raw with the Fourier mixtures computed beforehand: F, F1 and Fn on
// the (pad, H) grid at spec, spec + n, spec + 2n (fill_fourier's layout),
// the core terms and the ellipse-term derivatives given.  Same planes, same
// operations in the same order as galaxy_raw.
template <typename Real, bool kLds>
__device__ void galaxy_raw_spec(const Real* spec, double core_d, double dcore_d,
                                const double* dt9, int T, int need, Real* ws) {
  const GalLayout L = galaxy_layout(T);
  const int pad = L.pad, logp = ilog2(pad), n = pad * L.H, TT = T * T;
  const bool nder = need & 32;
  rfft2_pad<Real, kLds>(ws + L.P, T, pad, logp, ws + L.RB, ws + L.SP);
  Real* Gm = ws + L.G;
  spec_scale(ws + L.SP, spec, n, ws + L.Z);
  irfft2_crop<Real, kLds>(ws + L.Z, T, pad, logp, ws + L.RB, Gm);
  const Real core = Real(core_d);
  for (int e = threadIdx.x; e < TT; e += UPH_BLOCK) Gm[e] += core * ws[L.P + e];
  __syncthreads();
  for (int j = 0; j < 2; ++j) {
    if (!(need & (1 << j))) continue;
    Real* src = ws + (j == 0 ? L.Px : L.Py);
    Real* sp = ws + (j == 0 ? L.SPx : L.SPy);
    rfft2_pad<Real, kLds>(src, T, pad, logp, ws + L.RB, sp);
    spec_scale(sp, spec, n, ws + L.Z);
    Real* o = Gm + (size_t)(1 + j) * TT;
    irfft2_crop<Real, kLds>(ws + L.Z, T, pad, logp, ws + L.RB, o);
    for (int e = threadIdx.x; e < TT; e += UPH_BLOCK) o[e] += core * src[e];
    __syncthreads();
  }
  for (int p = 0; p < 3; ++p) {
    if (!(need & (4 << p))) continue;
    spec_shape_dt(ws + L.SP, spec + n, dt9, p, pad, ws + L.Z);
    irfft2_crop<Real, kLds>(ws + L.Z, T, pad, logp, ws + L.RB, Gm + (size_t)(3 + p) * TT);
  }
  if (nder) {
    spec_scale(ws + L.SP, spec + 2 * n, n, ws + L.Z);
    Real* o = Gm + (size_t)6 * TT;
    irfft2_crop<Real, kLds>(ws + L.Z, T, pad, logp, ws + L.RB, o);
    const Real dcore = Real(dcore_d);
    for (int e = threadIdx.x; e < TT; e += UPH_BLOCK) o[e] += dcore * ws[L.P + e];
    __syncthreads();
  }
}

}  // namespace uph
