// Galaxy pre-render for the band-split island fit (gfx950, float32).
//
// Before each evaluation pass the raw (unshifted) planes of every extended
// member -- the PSF-convolved mixture and the tangents the pass needs -- are
// rendered here, and the evaluation launch only shifts and places them.  Two
// launches per pass:
//   * spectra: one workgroup per galaxy computes, once for all bands, the
//     Fourier mixture F + core, F1 = sum a v E and Fn + dcore on the
//     (H = pad/2 + 1, pad) grid (column-major: frequency k along the stamp
//     row, then m), and the ellipse-term derivatives;
//   * planes: (galaxy, band) items.  The PSF stamp P (and dP/dx, dP/dy) of
//     the band at the member position is transformed with wave-level FFTs:
//     rows of P -> LDS line buffer -> columns -> spectrum S (LDS for pad <=
//     128, per-workgroup global columns for pad 256); each output plane is
//     then one column pass over S times its spectrum into the LDS line
//     buffer and one Hermitian row pass out to the galaxy's plane buffer.
// The core term is folded into the spectra (irfft2(S (F + core)) =
// irfft2(S F) + core P for the cropped stamp), so P is never needed in
// pixel space after the forward transform.
// Algorithm followed: tract7dt/fit/grid_gpu/sersic_mog.py
// (render_mog_stamp_fft :153-194, _ellipse_cov_terms :108-150) and
// source_model.py:264-328 (PSF at the source position); the planes are the
// raw stamps of island_solver.py:441-562 before the Lanczos shift.
#include <hip/hip_runtime.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"
#include "uph_galaxy.hpp"
#include "uph_galaxy_pre.hpp"
#include "uph_island.hpp"

namespace uph {

constexpr int kGalBlock = 512;  // threads of the planes kernel (8 waves)
constexpr int kGalWaves = kGalBlock / 64;
// LDS line buffer of the planes kernel (complex elements): holds the
// (H, T) row/column buffer plus, for pad <= 128, the (H, pad) spectrum.
constexpr int kGalLdsComplex = 14336;

__host__ __device__ inline bool galaxy_fits(int T, int pad) {
  const int H = pad / 2 + 1;
  const int need = H * T + (pad <= 128 ? H * pad : 0);
  return pad >= 64 && pad <= 256 && need <= kGalLdsComplex;
}

bool galaxy_prerender_supports(int T) { return galaxy_fits(T, fft_pad_for(T)); }

// Float offset of the global spectrum columns in a planes workgroup's
// scratch: after P, dP/dx, dP/dy, rounded up to 64 floats (float2 aligned).
__host__ __device__ inline size_t galaxy_spectrum_offset(int T) {
  return ((size_t)3 * T * T + 63) / 64 * 64;
}

__host__ __device__ inline size_t prerender_scratch_bytes(int T) {
  const int pad = fft_pad_for(T), H = pad / 2 + 1;
  size_t f = galaxy_spectrum_offset(T);          // P, dP/dx, dP/dy
  if (pad > 128) f += (size_t)2 * H * pad;       // spectrum columns
  return align256(f * sizeof(float));
}

__host__ __device__ inline size_t spectra_floats(int T) {
  const int pad = fft_pad_for(T);
  return (size_t)3 * (pad / 2 + 1) * pad;
}

size_t galaxy_prerender_scratch_bytes(int T) { return prerender_scratch_bytes(T); }
size_t galaxy_spectra_floats(int T) { return spectra_floats(T); }

__device__ __forceinline__ int brev_bits(int i, int logn) {
  return (int)(__brev((unsigned)i) >> (32 - logn));
}

template <int E>
struct LogE {
  static constexpr int v = E == 1 ? 0 : (E == 2 ? 1 : (E == 4 ? 2 : 3));
};

// Forward-transform twiddles of this lane for a line of 64 E points.
template <int E>
struct WaveTw {
  float cw[6][2];  // cross-lane stages: this lane's scale (twiddle or 1)
  float cs[6];     // and sign (+1 lower lane, -1 upper lane of its pair)
  float iw[LogE<E>::v + 1][E][2];
};

template <int E>
__device__ __forceinline__ void wave_tw(WaveTw<E>& t) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int s = 0; s < 6; ++s) {
    const int h = 1 << s;
    const bool upper = lane & h;
    float sn, cs;
    sincospif(float(lane & (h - 1)) / float(h), &sn, &cs);
    t.cw[s][0] = upper ? cs : 1.0f;
    t.cw[s][1] = upper ? -sn : 0.0f;
    t.cs[s] = upper ? -1.0f : 1.0f;
  }
#pragma unroll
  for (int q = 0; q < LogE<E>::v; ++q) {
    const int he = 1 << q;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int j = lane + 64 * (e & (he - 1));
      float sn, cs;
      sincospif(float(j) / float(64 * he), &sn, &cs);
      t.iw[q][e][0] = cs;
      t.iw[q][e][1] = -sn;
    }
  }
}

// x of lane (l ^ H) for power-of-two H < 64 without the LDS crossbar:
// DPP quad permutes and row (half) mirrors for H <= 8, a bit-mode swizzle
// for 16, the 32-lane swap for 32.
template <int H>
__device__ __forceinline__ float lane_xor(float x) {
  const int v = __float_as_int(x);
  int r;
  if constexpr (H == 1) {
    r = __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);  // quad [1,0,3,2]
  } else if constexpr (H == 2) {
    r = __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);  // quad [2,3,0,1]
  } else if constexpr (H == 4) {
    const int t = __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);  // i ^ 7
    r = __builtin_amdgcn_update_dpp(0, t, 0x1B, 0xF, 0xF, false);             // i ^ 3
  } else if constexpr (H == 8) {
    const int t = __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);  // i ^ 15
    r = __builtin_amdgcn_update_dpp(0, t, 0x141, 0xF, 0xF, false);            // i ^ 7
  } else if constexpr (H == 16) {
    r = __builtin_amdgcn_ds_swizzle(v, (0x10 << 10) | 0x1F);  // bit mode: xor 16
  } else {
    static_assert(H == 32, "lane_xor: H in 1..32");
    const auto pr = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    r = (threadIdx.x & 32) ? int(pr[0]) : int(pr[1]);
  }
  return __int_as_float(r);
}

// One cross-lane radix-2 stage of span 2^S.  Each lane scales its own
// element by w_l (the twiddle on the upper lane of a pair, 1 on the lower
// lane), swaps the product with its partner, and keeps partner + sg * own
// (sg = +1 lower, -1 upper): the lower lane ends with a + w b, the upper
// with a - w b, for a the lower and b the upper element of the pair.
template <int S, int E>
__device__ __forceinline__ void lane_stage(float (&re)[E], float (&im)[E], const float wr,
                                           const float wi, const float sg) {
  constexpr int h = 1 << S;
  float zr[E], zi[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    zr[e] = wr * re[e] - wi * im[e];
    zi[e] = wr * im[e] + wi * re[e];
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const float qr = lane_xor<h>(zr[e]), qi = lane_xor<h>(zi[e]);
    re[e] = fmaf(sg, zr[e], qr);
    im[e] = fmaf(sg, zi[e], qi);
  }
}

// In-register forward FFT of one line held by a wavefront: on entry lane l
// holds the bit-reversed-order elements l + 64 e, on exit the natural-order
// outputs l + 64 e.  Radix-2 decimation in time: spans 1..32 across lanes
// (shuffles), spans 64.. inside the lane.
template <int E>
__device__ __forceinline__ void wave_fft(float (&re)[E], float (&im)[E], const WaveTw<E>& t) {
  lane_stage<0, E>(re, im, t.cw[0][0], t.cw[0][1], t.cs[0]);
  lane_stage<1, E>(re, im, t.cw[1][0], t.cw[1][1], t.cs[1]);
  lane_stage<2, E>(re, im, t.cw[2][0], t.cw[2][1], t.cs[2]);
  lane_stage<3, E>(re, im, t.cw[3][0], t.cw[3][1], t.cs[3]);
  lane_stage<4, E>(re, im, t.cw[4][0], t.cw[4][1], t.cs[4]);
  lane_stage<5, E>(re, im, t.cw[5][0], t.cw[5][1], t.cs[5]);
#pragma unroll
  for (int q = 0; q < LogE<E>::v; ++q) {
    const int he = 1 << q;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      if (e & he) continue;
      const int f = e + he;
      const float wr = t.iw[q][e][0], wi = t.iw[q][e][1];
      const float br = wr * re[f] - wi * im[f], bi = wr * im[f] + wi * re[f];
      const float ar = re[e], ai = im[e];
      re[e] = ar + br;
      im[e] = ai + bi;
      re[f] = ar - br;
      im[f] = ai - bi;
    }
  }
}

// ---------------------------------------------------------------------------
// Spectra: one workgroup per galaxy.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(UPH_BLOCK) void galaxy_spectra_kernel(IslandArgsPre a) {
  __shared__ GalParams<float> gp;
  __shared__ int s_skip, s_T;
  const int g = blockIdx.x, tid = threadIdx.x, mode = a.mode;
  const float* psrc = mode == kSplitEvalInit ? a.p0 : (mode == kSplitEvalTry ? a.p_try : a.p_out);
  if (tid == 0) {
    const int isl = a.gal_isl[g];
    int skip = (mode == kSplitEvalGrad || mode == kSplitEvalTry) ? int(a.conv_out[isl]) : 0;
    const int32_t* sp = a.shape + 4 * (size_t)isl;
    const int T = a.tier_T[sp[2]];
    s_T = T;
    if ((size_t)a.gal_spec[g] + spectra_floats(T) > a.fspec_elems) {
      atomicOr(a.err, 1);
      skip = 1;
    }
    if (!skip) {
      const size_t slot = (size_t)a.mem_off[isl] + a.gal_slot[g];
      const float* p6 = psrc + slot * 6;
      double n = double(a.n_frozen[slot]);
      if (a.free_n[slot]) {
        const double u = double(p6[5]);
        n = kNMin + (kNMax - kNMin) / (1.0 + exp(-u));
      }
      gp.n = n;
      gp.K = int(a.sersic[3]);
      mog_lookup(a.sersic, n, gp.a, gp.v, gp.core, gp.da, gp.dv, gp.dcore);
      ellipse_terms(double(p6[2]), double(p6[3]), double(p6[4]), gp.t, gp.dt);
      for (int i = 0; i < 9; ++i) a.gdt[(size_t)g * 9 + i] = gp.dt[i / 3][i % 3];
    }
    s_skip = skip;
  }
  __syncthreads();
  if (s_skip) return;
  const int T = s_T, pad = fft_pad_for(T), H = pad / 2 + 1, n = H * pad;
  const bool grads = mode == kSplitEvalGrad;
  float* F = a.fspec + a.gal_spec[g];
  const double c2 = -2.0 * 9.869604401089358;  // -2 pi^2
  for (int e = tid; e < n; e += UPH_BLOCK) {
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
    F[e] = float(f + gp.core);
    if (grads) {
      F[n + e] = float(f1);
      F[2 * n + e] = float(fn + gp.dcore);
    }
  }
}

// ---------------------------------------------------------------------------
// Planes: (galaxy, band) items.
// ---------------------------------------------------------------------------
struct PlaneJob {
  int g, b, T, pad, logp, H, need;
  const float* coefs;
  const float* p6;
  float* out;     // the item's 7 raw planes (T^2 each)
  float* P3;      // per-workgroup P, dP/dx, dP/dy (global scratch)
  float2* S;      // spectrum (H, pad): LDS or global scratch
  float2* A;      // LDS line buffer (H, T)
  const float* F; // the galaxy's spectra
  const double* dt;
};

// Lines of one pass handled together by a wavefront (G E = 8 complex values
// in flight per lane): the gathers of G lines are issued before any of
// their transforms, so one memory latency covers the batch.
template <int E>
struct Batch {
  static constexpr int G = 8 / E;
};

// Forward transform of the T x T real plane src into S (H columns of pad).
template <int E>
__device__ void forward_plane(const PlaneJob& j, const float* src, const WaveTw<E>& tw) {
  constexpr int G = Batch<E>::G;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int T = j.T, H = j.H, pad = j.pad, logp = j.logp;
  // rows: A[k][r] = sum_c P[r][c] e^{-2 pi i k c / pad}, k < H
  for (int r0 = wid * G; r0 < T; r0 += kGalWaves * G) {
    float re[G][E], im[G][E];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int r = r0 + g;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = brev_bits(lane + 64 * e, logp);
        re[g][e] = (r < T && c < T) ? src[r * T + c] : 0.0f;
        im[g][e] = 0.0f;
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (r0 + g < T) wave_fft<E>(re[g], im[g], tw);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int r = r0 + g;
      if (r >= T) continue;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int k = lane + 64 * e;
        if (k < H) j.A[k * T + r] = make_float2(re[g][e], im[g][e]);
      }
    }
  }
  __syncthreads();
  // columns: S[k][m] = sum_r A[k][r] e^{-2 pi i m r / pad}
  for (int k0 = wid * G; k0 < H; k0 += kGalWaves * G) {
    float re[G][E], im[G][E];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int k = k0 + g;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int r = brev_bits(lane + 64 * e, logp);
        const float2 v = (k < H && r < T) ? j.A[k * T + r] : make_float2(0.0f, 0.0f);
        re[g][e] = v.x;
        im[g][e] = v.y;
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (k0 + g < H) wave_fft<E>(re[g], im[g], tw);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int k = k0 + g;
      if (k >= H) continue;
#pragma unroll
      for (int e = 0; e < E; ++e)
        j.S[(size_t)k * pad + lane + 64 * e] = make_float2(re[g][e], im[g][e]);
    }
  }
  __syncthreads();
}

// Inverse transform of S times the spectrum of output kind `kind`
// (0: F + core, 1..3: F1 d(expo)/d(logre, ee1, ee2), 4: Fn + dcore) into
// the T x T plane dst: dst = irfft2(S X, (pad, pad))[:T, :T].
template <int E>
__device__ void inverse_plane(const PlaneJob& j, int kind, float* dst, const WaveTw<E>& tw) {
  constexpr int G = Batch<E>::G;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int T = j.T, H = j.H, pad = j.pad, logp = j.logp, n = H * pad;
  const float* X = j.F + (kind == 0 ? 0 : (kind == 4 ? 2 * n : n));
  const double c2 = -2.0 * 9.869604401089358;
  const int p = kind - 1;
  // columns: Y[k][r] = sum_m S[k][m] X[k][m] e^{+2 pi i m r / pad}, r < T
  for (int k0 = wid * G; k0 < H; k0 += kGalWaves * G) {
    float re[G][E], im[G][E];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int k = k0 + g;
      const double v = double(k) / pad;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int m = brev_bits(lane + 64 * e, logp);
        float2 sv = make_float2(0.0f, 0.0f);
        float x = 0.0f;
        if (k < H) {
          const size_t o = (size_t)k * pad + m;
          sv = j.S[o];
          x = X[o];
          if (kind >= 1 && kind <= 3) {
            const double w = double(m < pad / 2 ? m : m - pad) / pad;
            const double dex = c2 * (j.dt[0 * 3 + p] * v * v + j.dt[1 * 3 + p] * w * w +
                                     2.0 * j.dt[2 * 3 + p] * v * w);
            x = float(double(x) * dex);
          }
        }
        // inverse via conj(FFT(conj(z)))
        re[g][e] = sv.x * x;
        im[g][e] = -(sv.y * x);
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (k0 + g < H) wave_fft<E>(re[g], im[g], tw);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int k = k0 + g;
      if (k >= H) continue;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int r = lane + 64 * e;
        if (r < T) j.A[k * T + r] = make_float2(re[g][e], -im[g][e]);
      }
    }
  }
  __syncthreads();
  // rows: out[r][c] = Re sum_k X_r[k] e^{+2 pi i k c / pad} / pad^2 with
  // the Hermitian extension X_r[k] = Y[k][r] (k < H), conj(Y[pad - k][r])
  const float norm = 1.0f / (float(pad) * float(pad));
  for (int r0 = wid * G; r0 < T; r0 += kGalWaves * G) {
    float re[G][E], im[G][E];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int r = r0 + g;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = brev_bits(lane + 64 * e, logp);
        float2 y = make_float2(0.0f, 0.0f);
        if (r < T) {
          if (c < H) {
            y = j.A[c * T + r];
          } else {
            y = j.A[(pad - c) * T + r];
            y.y = -y.y;
          }
        }
        re[g][e] = y.x;
        im[g][e] = -y.y;
      }
    }
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (r0 + g < T) wave_fft<E>(re[g], im[g], tw);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int r = r0 + g;
      if (r >= T) continue;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = lane + 64 * e;
        if (c < T) dst[r * T + c] = re[g][e] * norm;
      }
    }
  }
  __syncthreads();
}

template <int E>
__device__ void render_planes(const PlaneJob& j) {
  WaveTw<E> tw;
  wave_tw<E>(tw);
  const int TT = j.T * j.T, need = j.need;
  // source P: model (plane 0), shape tangents (planes 3..5), n (plane 6)
  forward_plane<E>(j, j.P3, tw);
  inverse_plane<E>(j, 0, j.out, tw);
  for (int q = 0; q < 3; ++q)
    if (need & (4 << q)) inverse_plane<E>(j, 1 + q, j.out + (size_t)(3 + q) * TT, tw);
  if (need & 32) inverse_plane<E>(j, 4, j.out + (size_t)6 * TT, tw);
  // sources dP/dx, dP/dy: position tangents (planes 1, 2)
  for (int q = 0; q < 2; ++q) {
    if (!(need & (1 << q))) continue;
    forward_plane<E>(j, j.P3 + (size_t)(1 + q) * TT, tw);
    inverse_plane<E>(j, 0, j.out + (size_t)(1 + q) * TT, tw);
  }
}

__global__ __launch_bounds__(kGalBlock) void galaxy_planes_kernel(IslandArgsPre a) {
  __shared__ float2 lds[kGalLdsComplex];
  __shared__ float bas[63];
  __shared__ int s_item, s_skip, s_need, s_T;
  __shared__ long long s_slot;
  __shared__ const float* s_coefs;
  const int tid = threadIdx.x, mode = a.mode, nb = a.nb, nt = a.nt;
  char* base = a.scratch + (size_t)blockIdx.x * a.scratch_per_block;
  const int n_items = a.n_gal * nb;
  const float* psrc = mode == kSplitEvalInit ? a.p0 : (mode == kSplitEvalTry ? a.p_try : a.p_out);
  for (;;) {
    if (tid == 0) {
      const int item = atomicAdd(a.gal_counter, 1);
      s_item = item;
      if (item < n_items) {
        const int g = item / nb, isl = a.gal_isl[g];
        int skip = (mode == kSplitEvalGrad || mode == kSplitEvalTry) ? int(a.conv_out[isl]) : 0;
        const int32_t* sp = a.shape + 4 * (size_t)isl;
        const int T = a.tier_T[sp[2]];
        s_T = T;
        s_coefs = a.tier_coefs[sp[2]];
        const size_t slot = (size_t)a.mem_off[isl] + a.gal_slot[g];
        s_slot = (long long)slot;
        int need = 0;
        if (mode == kSplitEvalGrad)
          for (int q = 0; q < 6; ++q)
            if (a.keep[slot * 6 + q]) need |= 1 << q;
        s_need = need;
        const int pad = fft_pad_for(T);
        if (!galaxy_fits(T, pad) || prerender_scratch_bytes(T) > a.scratch_per_block ||
            (size_t)a.gal_off[g] + (size_t)nb * 7 * T * T > a.gbuf_elems ||
            (size_t)a.gal_spec[g] + spectra_floats(T) > a.fspec_elems) {
          atomicOr(a.err, 1);
          skip = 1;
        }
        s_skip = skip;
        if (!skip) {
          // PSF basis of band b at the member position
          const int b = item % nb;
          const float* geo = a.psf_geo + 4 * b;
          const float* p6 = psrc + slot * 6;
          const float X = (p6[0] - geo[0]) / geo[1], Y = (p6[1] - geo[2]) / geo[3];
          psf_basis(X, Y, nt, bas, bas + 21, bas + 42);
        }
      }
    }
    __syncthreads();
    const int item = s_item;
    if (item >= n_items) break;
    if (s_skip) {
      __syncthreads();
      continue;
    }
    const int g = item / nb, b = item % nb, T = s_T, TT = T * T, need = s_need;
    const int pad = fft_pad_for(T), H = pad / 2 + 1;
    PlaneJob j;
    j.g = g;
    j.b = b;
    j.T = T;
    j.pad = pad;
    j.logp = ilog2(pad);
    j.H = H;
    j.need = need;
    j.coefs = s_coefs;
    j.p6 = psrc + (size_t)s_slot * 6;
    j.out = static_cast<float*>(a.gbuf) + a.gal_off[g] + (size_t)b * 7 * TT;
    j.P3 = reinterpret_cast<float*>(base);
    j.A = lds;
    j.S = pad <= 128 ? lds + H * T : reinterpret_cast<float2*>(j.P3 + galaxy_spectrum_offset(T));
    j.F = a.fspec + a.gal_spec[g];
    j.dt = a.gdt + (size_t)g * 9;
    // P (and dP/dx, dP/dy) of band b at the member position
    const float* geo = a.psf_geo + 4 * b;
    const float xs = geo[1], ys = geo[3];
    const float* Cf = s_coefs + (size_t)b * nt * TT;
    const bool gxy = need & 3;
    for (int e = tid; e < TT; e += kGalBlock) {
      float v = 0.0f, vx = 0.0f, vy = 0.0f;
      // coefficient loads in groups of 8 terms, all issued before their FMAs
      for (int t0 = 0; t0 < nt; t0 += 8) {
        float cv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) cv[u] = t0 + u < nt ? Cf[(size_t)(t0 + u) * TT + e] : 0.0f;
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          if (t0 + u >= nt) break;
          v = fmaf(cv[u], bas[t0 + u], v);
          if (gxy) {
            vx = fmaf(cv[u], bas[21 + t0 + u], vx);
            vy = fmaf(cv[u], bas[42 + t0 + u], vy);
          }
        }
      }
      j.P3[e] = v;
      if (gxy) {
        j.P3[TT + e] = vx / xs;
        j.P3[2 * TT + e] = vy / ys;
      }
    }
    __syncthreads();
    if (pad == 64)
      render_planes<1>(j);
    else if (pad == 128)
      render_planes<2>(j);
    else
      render_planes<4>(j);
    __syncthreads();
  }
}

void galaxy_prerender(const IslandArgsPre& a, int mode, int* counter, int nblk, hipStream_t s,
                      hipEvent_t ev_mid) {
  IslandArgsPre p = a;
  p.mode = mode;
  p.gal_counter = counter;
  if (p.n_gal > 0)
    hipLaunchKernelGGL(galaxy_spectra_kernel, dim3(p.n_gal), dim3(UPH_BLOCK), 0, s, p);
  if (ev_mid) hipEventRecord(ev_mid, s);
  const long long items = (long long)p.n_gal * p.nb;
  const int g = (int)(items < nblk ? items : nblk);
  if (g > 0) hipLaunchKernelGGL(galaxy_planes_kernel, dim3(g), dim3(kGalBlock), 0, s, p);
}

}  // namespace uph

extern "C" size_t uph_galaxy_prerender_scratch_bytes(int T) {
  return uph::galaxy_prerender_supports(T) ? uph::prerender_scratch_bytes(T) : 0;
}

extern "C" size_t uph_galaxy_spectra_floats(int T) { return uph::spectra_floats(T); }
