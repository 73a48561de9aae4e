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
    const float* cp = Cf + r * T + c;
    double s = 0, sx = 0, sy = 0;
    for (int t = 0; t < nt; ++t) {
      const double cv = double(cp[(size_t)t * TT]);
      s = fma(cv, bas[t], s);
      if (grad) {
        sx = fma(cv, bas[21 + t], sx);
        sy = fma(cv, bas[42 + t], sy);
      }
    }
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

// Extended member: PSF over the whole stamp, the convolved mixture and its
// raw tangents (galaxy_raw_spec), then the shift of the window part into
// planes 0 (model) and 1 + j for the parameters j in `need`.
__device__ void galaxy_planes64(const IslandArgsPre& a, const float* coefs, int T, int b,
                                const MemberWindow& w, int need, const double* spec,
                                const double* taps, const double* bas, double xs, double ys,
                                double* ws, double* blk, int side) {
  const int tid = threadIdx.x, nt = a.nt, TT = T * T;
  const bool gxy = need & 3;
  const bool grads = need != 0;
  const GalLayout L = galaxy_layout(T);
  const float* Cf = coefs + (size_t)b * nt * TT;
  for (int e = tid; e < TT; e += UPH_BLOCK) {
    double v = 0, vx = 0, vy = 0;
    for (int t = 0; t < nt; ++t) {
      const double cv = double(Cf[(size_t)t * TT + e]);
      v = fma(cv, bas[t], v);
      if (gxy) {
        vx = fma(cv, bas[21 + t], vx);
        vy = fma(cv, bas[42 + t], vy);
      }
    }
    ws[L.P + e] = v;
    if (gxy) {
      ws[L.Px + e] = vx / xs;
      ws[L.Py + e] = vy / ys;
    }
  }
  __syncthreads();
  const int n = L.pad * L.H;
  const double* tail = spec + 3 * (size_t)n;
  galaxy_raw_spec<double, false>(spec, tail[0], tail[1], tail + 2, T, need, ws);
  const double dndu = tail[11];
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
  const double* G = ws + L.G;
  double* Hh = ws + L.Hh;  // plane q of the horizontal pass at Hh + q T^2
  for (int e = tid; e < RH * SW; e += UPH_BLOCK) {
    const int rr = w.ry0 + e / SW, c = w.sx0 + e % SW;
    double h[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int cc = c + k - 3;
      if (cc < 0 || cc >= T) continue;
      const int o = rr * T + cc;
      h[0] = fma(Lx[k], G[o], h[0]);
      if (need & 1) h[1] = fma(Lx[k], G[TT + o], fma(dLx[k], G[o], h[1]));
#pragma unroll
      for (int j = 1; j < 6; ++j)
        if (need & (1 << j)) h[1 + j] = fma(Lx[k], G[(size_t)(1 + j) * TT + o], h[1 + j]);
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

__global__ __launch_bounds__(UPH_BLOCK) void pre64_planes_kernel(IslandArgsPre a) {
  __shared__ double taps[28];
  __shared__ double bas[63];
  __shared__ int s_item;
  const int tid = threadIdx.x;
  double* ws = reinterpret_cast<double*>(a.scratch + (size_t)blockIdx.x * a.scratch_per_block);
  const long long n_items = (long long)a.n_pre * a.nb;
  const float* psrc = pass_params(a);
  const bool grads = split_grads(a.mode);
  for (;;) {
    __syncthreads();  // the previous item's readers of taps/bas are done
    if (tid == 0) s_item = atomicAdd(a.pre_counter, 1);
    __syncthreads();
    const long long item = s_item;
    if (item >= n_items) break;
    const int g = (int)(item / a.nb), b = (int)(item % a.nb);
    const int isl = a.pre_isl[g];
    if (pass_skips(a, isl)) continue;
    const int32_t* sp = a.shape + 4 * (size_t)isl;
    const int B = sp[1], T = a.tier_T[sp[2]];
    const float* coefs = a.tier_coefs[sp[2]];
    const size_t slot = (size_t)a.mem_off[isl] + a.pre_slot[g];
    const float* p6 = psrc + slot * 6;
    const bool star = a.stype[slot] == 0;
    MemberWindow w;
    if (!member_window(p6[0], p6[1], a.x0b[isl], a.y0b[isl], T, B, w)) continue;  // uniform
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
    if (star)
      star_planes64(a, coefs, T, b, w, grads, taps, bas, xs, ys, ws, blk, side);
    else
      galaxy_planes64(a, coefs, T, b, w, need, a.fspec64 + a.pre_spec[g], taps, bas, xs, ys, ws,
                      blk, side);
  }
}

void prerender64_pass(const IslandArgsPre& a, int mode, int* counter, int nblk, hipStream_t s,
                      hipEvent_t ev_spec, hipEvent_t ev_planes) {
  IslandArgsPre p = a;
  p.mode = mode;
  p.pre_counter = counter;
  if (p.n_pre_ext > 0)
    hipLaunchKernelGGL(pre64_spectra_kernel, dim3(p.n_pre_ext), dim3(UPH_BLOCK), 0, s, p);
  if (ev_spec) (void)hipEventRecord(ev_spec, s);
  const long long items = (long long)p.n_pre * p.nb;
  const int g = (int)(items < nblk ? items : nblk);
  if (g > 0) hipLaunchKernelGGL(pre64_planes_kernel, dim3(g), dim3(UPH_BLOCK), 0, s, p);
  if (ev_planes) (void)hipEventRecord(ev_planes, s);
}

}  // namespace uph

extern "C" size_t uph_prerender64_scratch_bytes(int T) {
  return T > 0 ? uph::prerender64_scratch(T) : 0;
}

extern "C" size_t uph_prerender64_spectra_doubles(int T) {
  return T > 0 ? uph::prerender64_spectra_doubles(T) : 0;
}
