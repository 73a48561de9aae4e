// Member pre-render for the float64 island kernels (gfx950).
//
// The float64 evaluation places pre-rendered member planes exactly as the
// float32 one does, so the (island, band) evaluation launch holds no
// renderer and no FFT workspace.  Two launches per evaluation pass:
//   * spectra: one workgroup per extended member computes, once for all
//     bands, the Sersic mixture lookup, the ellipse terms and the Fourier
//     mixtures F, F1 = sum a v E and Fn on the (pad, pad/2 + 1) grid, in
//     float64, into the member's spectra block;
//   * planes: (member, band) items from a work queue, extended members
//     first.  A star's planes are the band PSF at its position and its x/y
//     derivatives; an extended member's are the PSF-convolved mixture and the
//     tangents of its live parameters (block-cooperative float64 FFTs in the
//     workgroup's scratch).  Both then get the separable Lanczos-3 sub-pixel
//     shift, and only the part of the stamp inside the island window is
//     stored (min(T, B)^2 doubles per plane).
// The arithmetic of each plane is the in-place float64 render's, operation
// for operation.  Algorithm followed: tract7dt/fit/grid_gpu/sersic_mog.py
// (render_mog_stamp_fft :153-194, _ellipse_cov_terms :108-150),
// source_model.py:264-328 and island_solver.py:441-562.
#include <hip/hip_runtime.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"
#include "uph_galaxy.hpp"
#include "uph_island.hpp"
#include "uph_prerender.hpp"

namespace uph {

// Doubles of an extended member's spectra block: F, F1, Fn (pad (pad/2+1)
// each), then core, dcore and the nine ellipse-term derivatives.
__host__ __device__ inline size_t spectra64_doubles(int T) {
  const int pad = fft_pad_for(T);
  return (size_t)3 * pad * (pad / 2 + 1) + 12;
}

size_t prerender64_scratch(int T) { return galaxy_scratch_bytes(T, sizeof(double)); }
size_t prerender64_spectra_doubles(int T) { return spectra64_doubles(T); }

constexpr int kLdsRows = 64;  // rows of the LDS line buffer of the pad-128 tier (T, box)
// doubles of the planes kernel's LDS buffer (line buffer + per-wave lines),
// which star items use for their box planes
constexpr size_t kLdsDoubles = (size_t)2 * (kLdsRows * 65 + UPH_WAVES * 128);

// Stamp tiers whose transforms run as wave-held lines with only the raw
// box the shift reads: pad 128 (line buffers in LDS, T <= kLdsRows) and pad
// 256 (line buffers in the workgroup's scratch).  Their spectra are stored
// frequency-column-major (k * pad + m), the others row-major
// (m * (pad / 2 + 1) + k, fill_fourier's layout).
__host__ __device__ inline bool lds_tier(int T) { return fft_pad_for(T) == 128 && T <= kLdsRows; }
__host__ __device__ inline bool box_tier(int T) {
  return lds_tier(T) || fft_pad_for(T) == 256;
}

// fill_fourier's F, F1 and Fn in the column-major layout of a box tier.
__device__ void fill_fourier_kmajor(const GalParams<double>& gp, int pad, bool grads, double* F) {
  const int n = pad * (pad / 2 + 1);
  const double c2 = -2.0 * 9.869604401089358;  // -2 pi^2
  for (int e = threadIdx.x; e < n; e += UPH_BLOCK) {
    const int k = e / pad, m = e % pad;
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
    F[e] = f;
    if (grads) {
      F[n + e] = f1;
      F[2 * n + e] = fn;
    }
  }
  __syncthreads();
}

__device__ __forceinline__ const float* pass_params(const IslandArgsPre& a) {
  return a.mode == kSplitEvalInit ? a.p0 : (a.mode == kSplitEvalTry ? a.p_try : a.p_out);
}

__device__ __forceinline__ bool pass_skips(const IslandArgsPre& a, int isl) {
  return (a.mode == kSplitEvalGrad || a.mode == kSplitEvalTry) && a.conv_out[isl];
}

// ---------------------------------------------------------------------------
// Spectra: one workgroup per extended member.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(UPH_BLOCK) void pre64_spectra_kernel(IslandArgsPre a) {
  __shared__ GalParams<double> gp;
  __shared__ int s_T, s_skip;
  const int g = blockIdx.x, tid = threadIdx.x;
  const int isl = a.pre_isl[g];
  if (pass_skips(a, isl)) return;  // uniform
  const size_t slot = (size_t)a.mem_off[isl] + a.pre_slot[g];
  const float* p6 = pass_params(a) + slot * 6;
  if (tid == 0) {
    const int T = a.tier_T[a.shape[4 * (size_t)isl + 2]];
    s_T = T;
    int skip = 0;
    if (a.pre_spec[g] < 0 || (size_t)a.pre_spec[g] + spectra64_doubles(T) > a.fspec_elems) {
      atomicOr(a.err, 1);
      skip = 1;
    } else {
      double n = double(a.n_frozen[slot]);
      gp.dndu = 0.0;
      if (a.free_n[slot]) {
        const double u = double(p6[5]);
        n = kNMin + (kNMax - kNMin) / (1.0 + exp(-u));
        gp.dndu = (n - kNMin) * (kNMax - n) / (kNMax - kNMin);
      }
      gp.n = n;
      gp.K = int(a.sersic[3]);
      mog_lookup(a.sersic, n, gp.a, gp.v, gp.core, gp.da, gp.dv, gp.dcore);
      ellipse_terms(double(p6[2]), double(p6[3]), double(p6[4]), gp.t, gp.dt);
    }
    s_skip = skip;
  }
  __syncthreads();
  if (s_skip) return;
  const int T = s_T, pad = fft_pad_for(T);
  const int n = pad * (pad / 2 + 1);
  double* F = a.fspec64 + a.pre_spec[g];
  const bool grads = split_grads(a.mode);
  if (box_tier(T))
    fill_fourier_kmajor(gp, pad, grads, F);
  else
    fill_fourier(gp, pad, grads, grads, F);
  if (tid == 0) {
    double* tail = F + 3 * (size_t)n;
    tail[0] = gp.core;
    tail[1] = gp.dcore;
    for (int i = 0; i < 9; ++i) tail[2 + i] = gp.dt[i / 3][i % 3];
    tail[11] = gp.dndu;
  }
}

// ---------------------------------------------------------------------------
// Planes: (member, band) items.
// ---------------------------------------------------------------------------
struct MemberWindow {
  int ix, iy;              // stamp top-left in the island window
  int sx0, sx1, sy0, sy1;  // window-covered box in stamp coordinates
  int rx0, rx1, ry0, ry1;  // raw-plane box the 7-tap shift reads
};

// The member's stamp against its island window (the placement of the
// evaluation kernel: rint(x - x0) - T / 2).  False when nothing lands.
__device__ inline bool member_window(float xf, float yf, float x0f, float y0f, int T, int B,
                                     MemberWindow& w) {
  const double xg = double(xf - x0f), yg = double(yf - y0f);
  w.ix = int(rint(xg)) - T / 2;
  w.iy = int(rint(yg)) - T / 2;
  const int wx0 = max(0, w.ix), wx1 = min(B, w.ix + T);
  const int wy0 = max(0, w.iy), wy1 = min(B, w.iy + T);
  if (wx1 <= wx0 || wy1 <= wy0) return false;
  w.sx0 = wx0 - w.ix;
  w.sx1 = wx1 - w.ix;
  w.sy0 = wy0 - w.iy;
  w.sy1 = wy1 - w.iy;
  w.rx0 = max(0, w.sx0 - 3);
  w.rx1 = min(T, w.sx1 + 3);
  w.ry0 = max(0, w.sy0 - 3);
  w.ry1 = min(T, w.sy1 + 3);
  return true;
}

// PSF polynomial (and its x/y derivatives with `grad`) over the box rows
// [ry0, ry0 + RH) x columns [rx0, rx0 + RW) of the T x T stamp into P, Px,
// Py (row stride RW): 4 box pixels per thread per round, the coefficient
// loads of all 4 issued before their FMAs (term order per pixel, as
// psf_poly_at).  Loads past the box or the last term read an in-bounds
// coefficient and are zeroed, so none is predicated.
// NT > 0: the term count known at compile time, 3 pixels per round with
// all their 3 NT loads in flight together; NT = 0: any nt, 4 pixels per
// round in groups of 8 terms.
template <int NT>
__device__ __forceinline__ void psf_box64_t(const float* Cf, int T, int nt, const double* bas,
                                            bool grad, double xs, double ys, int ry0, int rx0,
                                            int RH, int RW, double* P, double* Px, double* Py) {
  const int n = RH * RW;
  const size_t TT = (size_t)T * T;
  if constexpr (NT > 0) {
    constexpr int kPx = 3;  // pixels per round: 3 NT loads in flight per thread
    for (int e0 = threadIdx.x; e0 < n; e0 += kPx * UPH_BLOCK) {
      int off[kPx];
      bool ok[kPx];
#pragma unroll
      for (int j = 0; j < kPx; ++j) {
        const int e = e0 + j * UPH_BLOCK;
        ok[j] = e < n;
        off[j] = ok[j] ? (ry0 + e / RW) * T + rx0 + e % RW : 0;
      }
      float cv[kPx][NT];
#pragma unroll
      for (int j = 0; j < kPx; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) cv[j][t] = Cf[(size_t)t * TT + off[j]];
      double v[kPx], vx[kPx], vy[kPx];
#pragma unroll
      for (int j = 0; j < kPx; ++j) v[j] = vx[j] = vy[j] = 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int j = 0; j < kPx; ++j) {
          const double c = double(cv[j][t]);
          v[j] = fma(c, bas[t], v[j]);
          if (grad) {
            vx[j] = fma(c, bas[21 + t], vx[j]);
            vy[j] = fma(c, bas[42 + t], vy[j]);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < kPx; ++j) {
        if (!ok[j]) continue;
        const int e = e0 + j * UPH_BLOCK;
        P[e] = v[j];
        if (grad) {
          Px[e] = vx[j] / xs;
          Py[e] = vy[j] / ys;
        }
      }
    }
    return;
  }
  for (int e0 = threadIdx.x; e0 < n; e0 += 4 * UPH_BLOCK) {
    int off[4];
    bool ok[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int e = e0 + j * UPH_BLOCK;
      ok[j] = e < n;
      off[j] = ok[j] ? (ry0 + e / RW) * T + rx0 + e % RW : 0;
    }
    double v[4] = {0, 0, 0, 0}, vx[4] = {0, 0, 0, 0}, vy[4] = {0, 0, 0, 0};
    for (int t0 = 0; t0 < nt; t0 += 8) {
      float cv[4][8];
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int t = min(t0 + u, nt - 1);
          const float c = Cf[(size_t)t * TT + off[j]];
          cv[j][u] = (ok[j] && t0 + u < nt) ? c : 0.0f;
        }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        if (t0 + u >= nt) break;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double c = double(cv[j][u]);
          v[j] = fma(c, bas[t0 + u], v[j]);
          if (grad) {
            vx[j] = fma(c, bas[21 + t0 + u], vx[j]);
            vy[j] = fma(c, bas[42 + t0 + u], vy[j]);
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (!ok[j]) continue;
      const int e = e0 + j * UPH_BLOCK;
      P[e] = v[j];
      if (grad) {
        Px[e] = vx[j] / xs;
        Py[e] = vy[j] / ys;
      }
    }
  }
}

__device__ __forceinline__ void psf_box64(const float* Cf, int T, int nt, const double* bas,
                                          bool grad, double xs, double ys, int ry0, int rx0,
                                          int RH, int RW, double* P, double* Px, double* Py) {
  if (nt == 15)
    psf_box64_t<15>(Cf, T, nt, bas, grad, xs, ys, ry0, rx0, RH, RW, P, Px, Py);
  else if (nt == 10)
    psf_box64_t<10>(Cf, T, nt, bas, grad, xs, ys, ry0, rx0, RH, RW, P, Px, Py);
  else
    psf_box64_t<0>(Cf, T, nt, bas, grad, xs, ys, ry0, rx0, RH, RW, P, Px, Py);
}

// Star through LDS: the box planes (P, and Px, Py with `grad`) in sl, then
// the horizontal and vertical Lanczos passes fused: each thread takes 4
// vertically adjacent window pixels of one column, forms the horizontal
// sums of the 10 box rows they read (the same FMA chains as the separate
// horizontal pass) and accumulates them in row order (as the vertical
// pass), so the planes are those of star_planes64 bit for bit.
__device__ void star_planes64_lds(const IslandArgsPre& a, const float* coefs, int T, int b,
                                  const MemberWindow& w, bool grad, const double* taps,
                                  const double* bas, double xs, double ys, double* sl,
                                  double* blk, int side) {
  const int RW = w.rx1 - w.rx0, RH = w.ry1 - w.ry0, SW = w.sx1 - w.sx0, SH = w.sy1 - w.sy0;
  const size_t PP = (size_t)side * side;
  double* P = sl;
  double* Px = sl + RH * RW;
  double* Py = sl + 2 * RH * RW;
  psf_box64(coefs + (size_t)b * a.nt * T * T, T, a.nt, bas, grad, xs, ys, w.ry0, w.rx0, RH, RW,
            P, Px, Py);
  __syncthreads();
  double Lx[7], dLx[7], Ly[7], dLy[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    Lx[k] = taps[k];
    Ly[k] = taps[14 + k];
    dLx[k] = grad ? taps[7 + k] : 0.0;
    dLy[k] = grad ? taps[21 + k] : 0.0;
  }
  const int ngrp = (SH + 3) / 4;
  for (int q = threadIdx.x; q < ngrp * SW; q += UPH_BLOCK) {
    const int cc = q % SW, r0 = w.sy0 + 4 * (q / SW), c = w.sx0 + cc;
    double m[4] = {0, 0, 0, 0}, tx[4] = {0, 0, 0, 0}, ty[4] = {0, 0, 0, 0};
    for (int rr = r0 - 3; rr <= r0 + 6; ++rr) {
      if (rr < w.ry0 || rr >= w.ry1) continue;
      double h = 0, hx = 0, hy = 0;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        const int cx = c + k - 3;
        if (cx < w.rx0 || cx >= w.rx1) continue;
        const int o = (rr - w.ry0) * RW + cx - w.rx0;
        const double v = P[o];
        h = fma(Lx[k], v, h);
        if (grad) {
          hx = fma(Lx[k], Px[o], fma(dLx[k], v, hx));
          hy = fma(Lx[k], Py[o], hy);
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = rr - (r0 + j) + 3;
        if (k < 0 || k > 6) continue;
        m[j] = fma(Ly[k], h, m[j]);
        if (grad) {
          tx[j] = fma(Ly[k], hx, tx[j]);
          ty[j] = fma(Ly[k], hy, fma(dLy[k], h, ty[j]));
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = r0 + j;
      if (r >= w.sy1) break;
      const size_t so = (size_t)(r - w.sy0) * side + cc;
      blk[so] = m[j];
      if (grad) {
        blk[PP + so] = tx[j];
        blk[2 * PP + so] = ty[j];
      }
    }
  }
  __syncthreads();
}

// Star: PSF polynomial over the shift-read box (and its x/y derivatives),
// horizontal then vertical Lanczos pass, window part into planes 0 (model),
// 1 (d/dx), 2 (d/dy) of stride PP.  ws holds 6 boxes of T^2 doubles.
__device__ void star_planes64(const IslandArgsPre& a, const float* coefs, int T, int b,
                              const MemberWindow& w, bool grad, const double* taps,
                              const double* bas, double xs, double ys, double* ws, double* blk,
                              int side) {
  const int tid = threadIdx.x, nt = a.nt, TT = T * T;
  const int RW = w.rx1 - w.rx0, RH = w.ry1 - w.ry0, SW = w.sx1 - w.sx0, SH = w.sy1 - w.sy0;
  const size_t PP = (size_t)side * side;
  const float* Cf = coefs + (size_t)b * nt * TT;
  double* P = ws;
  double* Px = ws + TT;
  double* Py = ws + 2 * TT;
  double* H0 = ws + 3 * TT;
  double* Hx = ws + 4 * TT;
  double* Hy = ws + 5 * TT;
  for (int e = tid; e < RH * RW; e += UPH_BLOCK) {
    const int r = w.ry0 + e / RW, c = w.rx0 + e % RW;
    double s, sx, sy;
    psf_poly_at(Cf + r * T + c, (size_t)TT, nt, bas, grad, s, sx, sy);
    P[e] = s;
    if (grad) {
      Px[e] = sx / xs;
      Py[e] = sy / ys;
    }
  }
  __syncthreads();
  double Lx[7], dLx[7], Ly[7], dLy[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    Lx[k] = taps[k];
    Ly[k] = taps[14 + k];
    dLx[k] = grad ? taps[7 + k] : 0.0;
    dLy[k] = grad ? taps[21 + k] : 0.0;
  }
  for (int e = tid; e < RH * SW; e += UPH_BLOCK) {
    const int rr = e / SW, c = w.sx0 + e % SW;
    double h = 0, hx = 0, hy = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int cc = c + k - 3;
      if (cc < w.rx0 || cc >= w.rx1) continue;
      const int o = rr * RW + cc - w.rx0;
      const double v = P[o];
      h = fma(Lx[k], v, h);
      if (grad) {
        hx = fma(Lx[k], Px[o], fma(dLx[k], v, hx));
        hy = fma(Lx[k], Py[o], hy);
      }
    }
    H0[e] = h;
    if (grad) {
      Hx[e] = hx;
      Hy[e] = hy;
    }
  }
  __syncthreads();
  for (int e = tid; e < SH * SW; e += UPH_BLOCK) {
    const int r = w.sy0 + e / SW, cc = e % SW;
    double m = 0, tx = 0, ty = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int rr = r + k - 3;
      if (rr < w.ry0 || rr >= w.ry1) continue;
      const int o = (rr - w.ry0) * SW + cc;
      const double h = H0[o];
      m = fma(Ly[k], h, m);
      if (grad) {
        tx = fma(Ly[k], Hx[o], tx);
        ty = fma(Ly[k], Hy[o], fma(dLy[k], h, ty));
      }
    }
    const size_t so = (size_t)(e / SW) * side + cc;
    blk[so] = m;
    if (grad) {
      blk[PP + so] = tx;
      blk[2 * PP + so] = ty;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Box-tier transforms.  One wave owns an N = 64 E point line (E = 2: pad 128,
// E = 4: pad 256): lane l holds elements l + 64 e in registers.  Radix-2
// decimation in time on bit-reversed input: spans 1..32 across lanes
// (__shfl_xor), spans 64.. inside the lane.  Row transforms of real data take
// two rows per complex line (forward: split through conj(Z[N - k]); inverse:
// Hermitian rows a + i b give real a in the real and real b in the
// imaginary part).
// ---------------------------------------------------------------------------
template <int E>
struct LineE {
  static constexpr int N = 64 * E, H = N / 2 + 1, LOGE = E == 2 ? 1 : 2, LOGN = 6 + LOGE;
};

template <int E>
struct TwN {
  double c[6], s[6];    // cross-lane stage h = 2^q: pi (lane & (h - 1)) / h
  double ic[E - 1], is[E - 1];  // in-lane stage of span 64 he: pi (lane + 64 i) / (64 he)
                                // at [he - 1 + i], i < he (the he distinct values)
};

template <int E>
__device__ inline void twN(TwN<E>& t) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int h = 1 << q;
    sincospi(double(lane & (h - 1)) / double(h), &t.s[q], &t.c[q]);
  }
#pragma unroll
  for (int q = 0; q < LineE<E>::LOGE; ++q) {
    const int he = 1 << q;
#pragma unroll
    for (int i = 0; i < he; ++i)
      sincospi(double(lane + 64 * i) / double(64 * he), &t.is[he - 1 + i], &t.ic[he - 1 + i]);
  }
}

template <int E>
__device__ __forceinline__ int brevN(int i) {
  return (int)(__brev((unsigned)i) >> (32 - LineE<E>::LOGN));
}

// Lane l ^ H's value of v (H = 1 .. 32) without the LDS crossbar: quad
// permutes for 1 and 2, a quad reversal then a half-row mirror for 4 (lane
// 7 - ((l & 7) ^ 3) = l ^ 4 inside 8 lanes), a row rotation by 8 for 8, and
// the gfx950 row / half-wave swaps for 16 and 32 (the swap of a register
// with a copy of itself leaves the partner's value in one of the two
// results, by lane).
template <int H>
__device__ __forceinline__ int lane_xor(int v) {
  if constexpr (H == 1) {
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // quad_perm [1, 0, 3, 2]
  } else if constexpr (H == 2) {
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm [2, 3, 0, 1]
  } else if constexpr (H == 4) {
    const int r = __builtin_amdgcn_mov_dpp(v, 0x1B, 0xF, 0xF, false);  // quad_perm [3, 2, 1, 0]
    return __builtin_amdgcn_mov_dpp(r, 0x141, 0xF, 0xF, false);         // row_half_mirror
  } else if constexpr (H == 8) {
    return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (H == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return (threadIdx.x & 16) ? (int)r[0] : (int)r[1];
  } else {
    static_assert(H == 32, "lane_xor: H is a power of two below 64");
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32) ? (int)r[0] : (int)r[1];
  }
}

template <int H>
__device__ __forceinline__ double lane_xor(double v) {
  return __hiloint2double(lane_xor<H>(__double2hiint(v)), lane_xor<H>(__double2loint(v)));
}

// One cross-lane radix-2 stage of span H = 2^Q: the upper lane of each pair
// multiplies its value by the twiddle (the lower lane by 1), the pair
// exchanges products, and lower = own + partner's, upper = partner's - own.
template <int Q, int E>
__device__ __forceinline__ void cross_stage(double (&re)[E], double (&im)[E], double c, double s,
                                            double sign) {
  constexpr int H = 1 << Q;
  const bool upper = threadIdx.x & H;
  const double wr = upper ? c : 1.0, wi = upper ? sign * s : 0.0;
  const double g = upper ? -1.0 : 1.0;
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const double zr = wr * re[e] - wi * im[e], zi = wr * im[e] + wi * re[e];
    const double pr = lane_xor<H>(zr), pi = lane_xor<H>(zi);
    re[e] = fma(g, zr, pr);
    im[e] = fma(g, zi, pi);
  }
}

// In-register transform of L lines of 64 E points held in re/im[l E + e]
// (sign -1 forward, +1 inverse, unnormalised): on entry element e of line l
// is x_l[brevN(lane + 64 e)], on exit the outputs lane + 64 e in natural
// order.  The cross-lane stages of the L lines interleave.
template <int E, int L = 1>
__device__ __forceinline__ void fftN(double (&re)[L * E], double (&im)[L * E], const TwN<E>& t,
                                     double sign) {
  cross_stage<0>(re, im, t.c[0], t.s[0], sign);
  cross_stage<1>(re, im, t.c[1], t.s[1], sign);
  cross_stage<2>(re, im, t.c[2], t.s[2], sign);
  cross_stage<3>(re, im, t.c[3], t.s[3], sign);
  cross_stage<4>(re, im, t.c[4], t.s[4], sign);
  cross_stage<5>(re, im, t.c[5], t.s[5], sign);
#pragma unroll
  for (int q = 0; q < LineE<E>::LOGE; ++q) {
    const int he = 1 << q;
#pragma unroll
    for (int l = 0; l < L; ++l) {
#pragma unroll
      for (int e = 0; e < E; ++e) {
        if (e & he) continue;
        const int f = l * E + e + he, g = l * E + e;
        const int ti = he - 1 + (e & (he - 1));
        const double wr = t.ic[ti], wi = sign * t.is[ti];
        const double br = wr * re[f] - wi * im[f], bi = wr * im[f] + wi * re[f];
        const double ar = re[g], ai = im[g];
        re[g] = ar + br;
        im[g] = ai + bi;
        re[f] = ar - br;
        im[f] = ai - bi;
      }
    }
  }
}

// Forward rfft2 of the (T, T) real plane X zero-padded to N x N into S
// (H frequency columns k of N rows m, column-major: k * N + m); A is the
// line buffer (T x H complex, LDS or scratch), zw an N-complex LDS line
// per wave.
template <int E>
__device__ void rfft2_N(const double* X, int T, double2* S, double2* A, double2* zw) {
  constexpr int N = LineE<E>::N, H = LineE<E>::H;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  TwN<E> tw;
  twN<E>(tw);
  double2* z = zw + wid * N;
  int ci[E];
#pragma unroll
  for (int e = 0; e < E; ++e) ci[e] = brevN<E>(lane + 64 * e);
  // rows, two per line.  Pad 128: the next pair is loaded before this
  // pair's transform
  constexpr bool kAhead = E == 2;
  double xa[E], xb[E];
  auto load_rows = [&](int j) {
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int c = ci[e];
      xa[e] = c < T ? X[2 * j * T + c] : 0.0;
      xb[e] = (2 * j + 1 < T && c < T) ? X[(2 * j + 1) * T + c] : 0.0;
    }
  };
  if (kAhead && 2 * wid < T) load_rows(wid);
  for (int j = wid; 2 * j < T; j += UPH_WAVES) {
    const int ra = 2 * j, rb = 2 * j + 1;
    if (!kAhead) load_rows(j);
    double re[E], im[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      re[e] = xa[e];
      im[e] = xb[e];
    }
    if (kAhead && 2 * (j + UPH_WAVES) < T) load_rows(j + UPH_WAVES);
    fftN<E>(re, im, tw, -1.0);
#pragma unroll
    for (int e = 0; e < E; ++e) z[lane + 64 * e] = make_double2(re[e], im[e]);
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < H; k += 64) {
      const double2 zk = z[k], zm = z[(N - k) & (N - 1)];
      A[ra * H + k] = make_double2(0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y));
      if (rb < T) A[rb * H + k] = make_double2(0.5 * (zk.y + zm.y), 0.5 * (zm.x - zk.x));
    }
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // columns; pad 128 transforms columns k and k + UPH_WAVES together
  constexpr int L = E == 2 ? 2 : 1;
  for (int k = wid; k < H; k += L * UPH_WAVES) {
    double re[L * E], im[L * E];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int kl = k + l * UPH_WAVES;  // wave-uniform
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const double2 v =
            (kl < H && ci[e] < T) ? A[ci[e] * H + kl] : make_double2(0.0, 0.0);
        re[l * E + e] = v.x;
        im[l * E + e] = v.y;
      }
    }
    fftN<E, L>(re, im, tw, -1.0);
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int kl = k + l * UPH_WAVES;
      if (kl >= H) continue;
#pragma unroll
      for (int e = 0; e < E; ++e)
        S[kl * N + lane + 64 * e] = make_double2(re[l * E + e], im[l * E + e]);
    }
  }
  __syncthreads();
}

// Spectral weight of output plane kind `kind` at (m, k): 0 F, 1 F1 times
// the ellipse-term derivative of parameter p, 2 Fn (spec_scale /
// spec_shape_dt of galaxy_raw_spec), spectra column-major.
template <int E>
__device__ __forceinline__ double spec_weight(const double* spec, int kind, int p,
                                              const double* dt9, int m, int k) {
  constexpr int N = LineE<E>::N, n = N * LineE<E>::H;
  const int e = k * N + m;
  if (kind == 0) return spec[e];
  if (kind == 2) return spec[2 * n + e];
  const double c2 = -2.0 * 9.869604401089358;
  const double v = double(k) / N;
  const double w = double(m < N / 2 ? m : m - N) / N;
  const double dex = c2 * (dt9[p] * v * v + dt9[3 + p] * w * w + 2.0 * dt9[6 + p] * v * w);
  return double(spec[n + e]) * dex;
}

// Raw plane box [ry0, ry1) x [rx0, rx1) of irfft2(S * W)[:T, :T] (W by
// kind/p), plus `core` times the real plane Q (T x T) when Q is given, into
// out (row stride RW).  B is the line buffer (box rows x H complex, LDS or
// scratch).
template <int E>
__device__ void irfft2_box_N(const double2* S, const double* spec, int kind, int p,
                             const double* dt9, int ry0, int ry1, int rx0, int rx1,
                             const double* Q, double core, int T, double* out, double2* B) {
  constexpr int N = LineE<E>::N, H = LineE<E>::H;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  TwN<E> tw;
  twN<E>(tw);
  const int RH = ry1 - ry0, RW = rx1 - rx0;
  int mi[E];
#pragma unroll
  for (int e = 0; e < E; ++e) mi[e] = brevN<E>(lane + 64 * e);
  // columns: inverse along m, keep the box rows.  Pad 128 transforms
  // columns k and k + UPH_WAVES together (both loaded before either
  // transform); pad 256 has no registers to spare for it
  constexpr int L = E == 2 ? 2 : 1;
  for (int k = wid; k < H; k += L * UPH_WAVES) {
    double re[L * E], im[L * E];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int kl = k + l * UPH_WAVES;  // wave-uniform
#pragma unroll
      for (int e = 0; e < E; ++e) {
        double2 v = make_double2(0.0, 0.0);
        double w = 0.0;
        if (kl < H) {
          v = S[kl * N + mi[e]];
          w = spec_weight<E>(spec, kind, p, dt9, mi[e], kl);
        }
        re[l * E + e] = v.x * w;
        im[l * E + e] = v.y * w;
      }
    }
    fftN<E, L>(re, im, tw, 1.0);
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int kl = k + l * UPH_WAVES;
      if (kl >= H) continue;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int m = lane + 64 * e;
        if (m >= ry0 && m < ry1)
          B[(m - ry0) * H + kl] = make_double2(re[l * E + e], im[l * E + e]);
      }
    }
  }
  __syncthreads();
  // rows of the box, two per line: the Hermitian extension of row a plus i
  // times that of row b; the inverse of that line holds row a's real
  // output in its real part and row b's in its imaginary part
  const double norm = 1.0 / (double(N) * double(N));
  for (int j = wid; 2 * j < RH; j += UPH_WAVES) {
    const int ia = 2 * j, ib = 2 * j + 1;
    const bool hb = ib < RH;
    double re[E], im[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int c = mi[e];
      const bool lo = c < H;
      const int cc = lo ? c : N - c;
      const double2 a = B[ia * H + cc];
      const double2 b = hb ? B[ib * H + cc] : make_double2(0.0, 0.0);
      // the DC and Nyquist bins enter with their real parts only, as in a
      // complex-to-real transform (which a pair would otherwise mix)
      const bool edge = cc == 0 || cc == N / 2;
      const double ai = edge ? 0.0 : (lo ? a.y : -a.y);
      const double bi = edge ? 0.0 : (lo ? b.y : -b.y);
      re[e] = a.x - bi;
      im[e] = ai + b.x;
    }
    fftN<E>(re, im, tw, 1.0);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int c = lane + 64 * e;
      if (c < rx0 || c >= rx1) continue;
      const int ra = ry0 + ia, rb = ry0 + ib;
      double va = re[e] * norm;
      if (Q) va += core * Q[ra * T + c];
      out[ia * RW + (c - rx0)] = va;
      if (hb) {
        double vb = im[e] * norm;
        if (Q) vb += core * Q[rb * T + c];
        out[ib * RW + (c - rx0)] = vb;
      }
    }
  }
  __syncthreads();
}

// Shift of the raw planes into the window part: horizontal then vertical
// Lanczos-3 pass (with the tap derivatives on the x / y tangents), planes
// 0 (model) and 1 + j for the parameters j in `need`.  Raw plane q is read
// at G + q * gq + (row - gy0) * gs + (col - gx0) for rows in [ry0, ry1) and
// columns in [rx0, rx1).
__device__ void shift_planes64(const double* G, size_t gq, int gs, int gy0, int gx0, int T,
                               const MemberWindow& w, int need, const double* taps, double dndu,
                               double* Hh, double* blk, int side) {
  const int tid = threadIdx.x, TT = T * T;
  const bool grads = need != 0;
  const int RH = w.ry1 - w.ry0, SW = w.sx1 - w.sx0, SH = w.sy1 - w.sy0;
  const size_t PP = (size_t)side * side;
  double Lx[7], dLx[7], Ly[7], dLy[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    Lx[k] = taps[k];
    Ly[k] = taps[14 + k];
    dLx[k] = grads ? taps[7 + k] : 0.0;
    dLy[k] = grads ? taps[21 + k] : 0.0;
  }
  for (int e = tid; e < RH * SW; e += UPH_BLOCK) {
    const int rr = w.ry0 + e / SW, c = w.sx0 + e % SW;
    double h[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int cc = c + k - 3;
      if (cc < 0 || cc >= T) continue;
      const size_t o = (size_t)(rr - gy0) * gs + (cc - gx0);
      h[0] = fma(Lx[k], G[o], h[0]);
      if (need & 1) h[1] = fma(Lx[k], G[gq + o], fma(dLx[k], G[o], h[1]));
#pragma unroll
      for (int j = 1; j < 6; ++j)
        if (need & (1 << j)) h[1 + j] = fma(Lx[k], G[(size_t)(1 + j) * gq + o], h[1 + j]);
    }
    Hh[e] = h[0];
#pragma unroll
    for (int j = 0; j < 6; ++j)
      if (need & (1 << j)) Hh[(size_t)(1 + j) * TT + e] = h[1 + j];
  }
  __syncthreads();
  for (int e = tid; e < SH * SW; e += UPH_BLOCK) {
    const int rw = w.sy0 + e / SW, cc = e % SW;
    double v[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int rr = rw + k - 3;
      if (rr < w.ry0 || rr >= w.ry1) continue;
      const int o = (rr - w.ry0) * SW + cc;
      const double h0 = Hh[o];
      v[0] = fma(Ly[k], h0, v[0]);
      if (need & 2) v[2] = fma(Ly[k], Hh[(size_t)2 * TT + o], fma(dLy[k], h0, v[2]));
#pragma unroll
      for (int j = 0; j < 6; ++j)
        if (j != 1 && (need & (1 << j)))
          v[1 + j] = fma(Ly[k], Hh[(size_t)(1 + j) * TT + o], v[1 + j]);
    }
    const size_t so = (size_t)(e / SW) * side + cc;
    blk[so] = v[0];
#pragma unroll
    for (int j = 0; j < 6; ++j)
      if (need & (1 << j)) blk[(size_t)(1 + j) * PP + so] = j == 5 ? v[6] * dndu : v[1 + j];
  }
  __syncthreads();
}

// PSF planes of an extended member over the whole stamp: P, and dP/dx,
// dP/dy when the x or y tangent is needed, at the layout's P, Px, Py.
__device__ void galaxy_psf64(const IslandArgsPre& a, const float* coefs, int T, int b, int need,
                             const double* bas, double xs, double ys, double* ws) {
  const int nt = a.nt, TT = T * T;
  const bool gxy = need & 3;
  const GalLayout L = galaxy_layout(T);
  const float* Cf = coefs + (size_t)b * nt * TT;
  for (int e = threadIdx.x; e < TT; e += UPH_BLOCK) {
    double v, vx, vy;
    psf_poly_at(Cf + e, (size_t)TT, nt, bas, gxy, v, vx, vy);
    ws[L.P + e] = v;
    if (gxy) {
      ws[L.Px + e] = vx / xs;
      ws[L.Py + e] = vy / ys;
    }
  }
  __syncthreads();
}

// Extended member of a box tier: the convolved mixture and its raw tangents
// over the box the shift reads only, by wave-held line transforms (line
// buffers in LDS for pad 128, in the workgroup's scratch for pad 256), then
// the shift of the window part into planes 0 (model) and 1 + j for the
// parameters j in `need`.
template <int E>
__device__ void galaxy_planes64_box(const IslandArgsPre& a, const float* coefs, int T, int b,
                                    const MemberWindow& w, int need, const double* spec,
                                    const double* taps, const double* bas, double xs, double ys,
                                    double* ws, double* blk, int side, double2* lds,
                                    double2* zw) {
  const int TT = T * T;
  const GalLayout L = galaxy_layout(T);
  galaxy_psf64(a, coefs, T, b, need, bas, xs, ys, ws);
  const int n = L.pad * L.H;
  const double* tail = spec + 3 * (size_t)n;
  const double core = tail[0], dcore = tail[1], dndu = tail[11];
  const double* dt9 = tail + 2;
  // raw-plane boxes (RH x RW, row stride RW) at ws + L.G + q T^2; the
  // spectrum in the (16-byte aligned) spectra region; for pad 256 the line
  // buffer in the next spectra region
  const int RW = w.rx1 - w.rx0;
  double2* S = reinterpret_cast<double2*>(ws + (L.SP + 1) / 2 * 2);
  double2* A = E == 2 ? lds : reinterpret_cast<double2*>(ws + (L.SPx + 1) / 2 * 2);
  double* G = ws + L.G;
  rfft2_N<E>(ws + L.P, T, S, A, zw);
  irfft2_box_N<E>(S, spec, 0, 0, dt9, w.ry0, w.ry1, w.rx0, w.rx1, ws + L.P, core, T, G, A);
  for (int p = 0; p < 3; ++p)
    if (need & (4 << p))
      irfft2_box_N<E>(S, spec, 1, p, dt9, w.ry0, w.ry1, w.rx0, w.rx1, nullptr, 0.0, T,
                      G + (size_t)(3 + p) * TT, A);
  if (need & 32)
    irfft2_box_N<E>(S, spec, 2, 0, dt9, w.ry0, w.ry1, w.rx0, w.rx1, ws + L.P, dcore, T,
                    G + (size_t)6 * TT, A);
  for (int j = 0; j < 2; ++j) {
    if (!(need & (1 << j))) continue;
    const double* Q = ws + (j == 0 ? L.Px : L.Py);
    rfft2_N<E>(Q, T, S, A, zw);
    irfft2_box_N<E>(S, spec, 0, 0, dt9, w.ry0, w.ry1, w.rx0, w.rx1, Q, core, T,
                    G + (size_t)(1 + j) * TT, A);
  }
  shift_planes64(G, TT, RW, w.ry0, w.rx0, T, w, need, taps, dndu, ws + L.Hh, blk, side);
}

// Extended member of a larger tier: galaxy_raw_spec's block transforms in
// the workgroup's scratch over the whole stamp, then the shift.
__device__ void galaxy_planes64_block(const IslandArgsPre& a, const float* coefs, int T, int b,
                                      const MemberWindow& w, int need, const double* spec,
                                      const double* taps, const double* bas, double xs,
                                      double ys, double* ws, double* blk, int side) {
  const int TT = T * T;
  const GalLayout L = galaxy_layout(T);
  galaxy_psf64(a, coefs, T, b, need, bas, xs, ys, ws);
  const int n = L.pad * L.H;
  const double* tail = spec + 3 * (size_t)n;
  galaxy_raw_spec<double, false>(spec, tail[0], tail[1], tail + 2, T, need, ws);
  shift_planes64(ws + L.G, TT, T, 0, 0, T, w, need, taps, tail[11], ws + L.Hh, blk, side);
}

// Item kinds of the planes kernels: stars and the extended members of the
// LDS tier; extended members of tier 111 (pad 256); of larger tiers.
constexpr int kItemLds = 0, kItemPad256 = 1, kItemBlock = 2;

// One (member, band) item of kind kKind; an item of another kind is left
// alone.  Block-uniform; ends after a barrier.  Items of extended members
// ([0, n_pre_ext nb)) run member-major, so the bands of one member share
// its spectra in L2; star items run band-major, so the workgroups working
// at once share one band's PSF coefficient table in L2.
template <int kKind>
__device__ void plane_item(const IslandArgsPre& a, long long item, double* taps, double* bas,
                           double2* lds, double2* zw, double* ws) {
  const int tid = threadIdx.x;
  const bool grads = split_grads(a.mode);
  const long long n_ext_items = (long long)a.n_pre_ext * a.nb;
  int g, b;
  if (item < n_ext_items) {
    g = (int)(item / a.nb);
    b = (int)(item % a.nb);
  } else {
    const long long s = item - n_ext_items, n_star = a.n_pre - a.n_pre_ext;
    b = (int)(s / n_star);
    g = a.n_pre_ext + (int)(s % n_star);
  }
  const int isl = a.pre_isl[g];
  if (pass_skips(a, isl)) return;
  const int32_t* sp = a.shape + 4 * (size_t)isl;
  const int B = sp[1], T = a.tier_T[sp[2]];
  const size_t slot = (size_t)a.mem_off[isl] + a.pre_slot[g];
  const bool star = a.stype[slot] == 0;
  const int kind = (star || lds_tier(T)) ? kItemLds : (box_tier(T) ? kItemPad256 : kItemBlock);
  if (kind != kKind) return;
  const float* coefs = a.tier_coefs[sp[2]];
  const float* p6 = pass_params(a) + slot * 6;
  MemberWindow w;
  if (!member_window(p6[0], p6[1], a.x0b[isl], a.y0b[isl], T, B, w)) return;
  int need = 0;
  if (grads)
    for (int j = 0; j < 6; ++j) need |= a.keep[slot * 6 + j] ? (1 << j) : 0;
  const int side = pre_side(T, B);
  double* blk = a.pbuf64 + a.pre_off[g] + (size_t)b * pre_planes(star) * side * side;
  const float* geo = a.psf_geo + 4 * b;
  const double xs = double(geo[1]), ys = double(geo[3]);
  const double x = double(p6[0]), y = double(p6[1]);
  const double xg = double(p6[0] - a.x0b[isl]), yg = double(p6[1] - a.y0b[isl]);
  const double dx = xg - rint(xg), dy = yg - rint(yg);
  const bool tg = star ? grads : need != 0;
  if (tid == 0) lanczos_taps(dx, taps, tg ? taps + 7 : (double*)nullptr);
  if (tid == 64) lanczos_taps(dy, taps + 14, tg ? taps + 21 : (double*)nullptr);
  if (tid == 128) {
    const double X = (x - double(geo[0])) / xs, Y = (y - double(geo[2])) / ys;
    psf_basis(X, Y, a.nt, bas, bas + 21, bas + 42);
  }
  __syncthreads();
  const double* spec = star ? nullptr : a.fspec64 + a.pre_spec[g];
  if constexpr (kKind == kItemPad256) {
    galaxy_planes64_box<4>(a, coefs, T, b, w, need, spec, taps, bas, xs, ys, ws, blk, side,
                           nullptr, zw);
  } else if constexpr (kKind == kItemBlock) {
    galaxy_planes64_block(a, coefs, T, b, w, need, spec, taps, bas, xs, ys, ws, blk, side);
  } else {
    const int nplanes = grads ? 3 : 1;
    if (star && (size_t)nplanes * (w.ry1 - w.ry0) * (w.rx1 - w.rx0) <= kLdsDoubles)
      star_planes64_lds(a, coefs, T, b, w, grads, taps, bas, xs, ys,
                        reinterpret_cast<double*>(lds), blk, side);
    else if (star)
      star_planes64(a, coefs, T, b, w, grads, taps, bas, xs, ys, ws, blk, side);
    else
      galaxy_planes64_box<2>(a, coefs, T, b, w, need, spec, taps, bas, xs, ys, ws, blk, side,
                             lds, zw);
  }
}

// Stars and the extended members of LDS tiers, from the pass's work queue
// (extended members first).
__global__ __launch_bounds__(UPH_BLOCK) void pre64_planes_kernel(IslandArgsPre a) {
  __shared__ double taps[28];
  __shared__ double bas[63];
  __shared__ long long s_item;
  // line buffer (kLdsRows x 65 complex) and one 128-complex line per wave
  __shared__ double2 lds[kLdsRows * LineE<2>::H + UPH_WAVES * LineE<2>::N];
  static_assert(sizeof(lds) == kLdsDoubles * sizeof(double), "star planes LDS capacity");
  double* ws = reinterpret_cast<double*>(a.scratch + (size_t)blockIdx.x * a.scratch_per_block);
  const long long n_items = (long long)a.n_pre * a.nb;
  for (;;) {
    __syncthreads();  // the previous item's readers of s_item, taps and bas are done
    if (threadIdx.x == 0) s_item = atomicAdd(a.pre_counter, 1);
    __syncthreads();
    const long long item = s_item;
    if (item >= n_items) break;
    plane_item<kItemLds>(a, item, taps, bas, lds, lds + kLdsRows * LineE<2>::H, ws);
  }
}

// The extended members of tier 111 (pad 256), from their own work queue
// (the counter after the planes queue; only extended-member items are
// visited).
// 3 workgroups per CU: the register cap (168) costs a few spills and buys
// 3 waves per SIMD for this latency-bound kernel
__global__ __launch_bounds__(UPH_BLOCK, 3) void pre64_planes_big_kernel(IslandArgsPre a) {
  __shared__ double taps[28];
  __shared__ double bas[63];
  __shared__ long long s_item;
  __shared__ double2 zw[UPH_WAVES * LineE<4>::N];  // one 256-complex line per wave
  double* ws = reinterpret_cast<double*>(a.scratch + (size_t)blockIdx.x * a.scratch_per_block);
  const long long n_items = (long long)a.n_pre_ext * a.nb;
  for (;;) {
    __syncthreads();  // the previous item's readers of s_item, taps and bas are done
    if (threadIdx.x == 0) s_item = atomicAdd(a.pre_counter + 1, 1);
    __syncthreads();
    const long long item = s_item;
    if (item >= n_items) break;
    plane_item<kItemPad256>(a, item, taps, bas, nullptr, zw, ws);
  }
}

// The extended members of the largest tier (pad 512, galaxy_raw_spec's
// block transforms): items strided over the grid.
__global__ __launch_bounds__(UPH_BLOCK) void pre64_planes_block_kernel(IslandArgsPre a) {
  __shared__ double taps[28];
  __shared__ double bas[63];
  double* ws = reinterpret_cast<double*>(a.scratch + (size_t)blockIdx.x * a.scratch_per_block);
  const long long n_items = (long long)a.n_pre_ext * a.nb;
  for (long long item = blockIdx.x; item < n_items; item += gridDim.x) {
    __syncthreads();  // the previous item's readers of taps and bas are done
    plane_item<kItemBlock>(a, item, taps, bas, nullptr, nullptr, ws);
  }
}

void prerender64_pass(const IslandArgsPre& a, int mode, int* counter, int nblk, hipStream_t s,
                      hipEvent_t ev_spec, hipEvent_t ev_lds) {
  IslandArgsPre p = a;
  p.mode = mode;
  p.pre_counter = counter;
  if (p.n_pre_ext > 0)
    hipLaunchKernelGGL(pre64_spectra_kernel, dim3(p.n_pre_ext), dim3(UPH_BLOCK), 0, s, p);
  if (ev_spec) (void)hipEventRecord(ev_spec, s);
  const long long items = (long long)p.n_pre * p.nb;
  const int g = (int)(items < nblk ? items : nblk);
  if (g > 0) hipLaunchKernelGGL(pre64_planes_kernel, dim3(g), dim3(UPH_BLOCK), 0, s, p);
  if (ev_lds) (void)hipEventRecord(ev_lds, s);
  bool mid = false, large = false;
  for (int t = 0; t < 3; ++t) {
    const int T = p.tier_T[t];
    mid = mid || (T > 0 && !lds_tier(T) && box_tier(T));
    large = large || (T > 0 && !box_tier(T));
  }
  const long long items_ext = (long long)p.n_pre_ext * p.nb;
  const int gb = (int)(items_ext < nblk ? items_ext : nblk);
  if (mid && gb > 0)
    hipLaunchKernelGGL(pre64_planes_big_kernel, dim3(gb), dim3(UPH_BLOCK), 0, s, p);
  if (large && gb > 0)
    hipLaunchKernelGGL(pre64_planes_block_kernel, dim3(gb), dim3(UPH_BLOCK), 0, s, p);
}

}  // namespace uph

extern "C" size_t uph_prerender64_scratch_bytes(int T) {
  return T > 0 ? uph::prerender64_scratch(T) : 0;
}

extern "C" size_t uph_prerender64_spectra_doubles(int T) {
  return T > 0 ? uph::prerender64_spectra_doubles(T) : 0;
}
