// Member pre-render for the band-split island fit (gfx950, float32).
//
// Before each evaluation pass every real member of the active islands is
// rendered and shifted here, and the evaluation launch only places the
// window part of its planes.  Two launches per pass:
//   * spectra: one workgroup per extended member computes, once for all
//     bands, the Fourier mixture F + core, F1 = sum a v E and Fn + dcore on
//     the (H = pad/2 + 1, pad) grid (column-major: frequency k along the
//     stamp row, then m), and the ellipse-term derivatives;
//   * planes: (member, band) items.  A star's raw planes are the PSF stamp
//     P of the band at its position and dP/dx, dP/dy; an extended member's
//     are the PSF-convolved mixture and the tangents the pass needs, from
//     wave-level FFTs: rows of P -> LDS line buffer -> columns -> spectrum
//     S (LDS for pad <= 128, per-workgroup global columns for pad 256); each
//     output plane is one column pass over S times its spectrum into the LDS
//     line buffer and one Hermitian row pass over the rows the window box
//     needs, whose outputs get the horizontal part of the member's sub-pixel
//     shift (separable Lanczos-3 with its tap derivatives) while still in
//     registers; the vertical part then writes the member's output planes.
// The core term is folded into the spectra (irfft2(S (F + core)) =
// irfft2(S F) + core P for the cropped stamp).
// Algorithm followed: tract7dt/fit/grid_gpu/sersic_mog.py
// (render_mog_stamp_fft :153-194, _ellipse_cov_terms :108-150),
// source_model.py:264-328 (PSF at the source position) and
// island_solver.py:441-562 (member stamps and their tangents).
#include <hip/hip_runtime.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"
#include "uph_galaxy.hpp"
#include "uph_prerender.hpp"
#include "uph_island.hpp"

namespace uph {

constexpr int kPreBlock = 512;  // threads of the planes kernel (8 waves)
constexpr int kGalWaves = kPreBlock / 64;
// LDS line buffer of the planes kernel (complex elements): holds the
// (H, T) row/column buffer plus, for pad <= 128, the (H, pad) spectrum.
constexpr int kGalLdsComplex = 14336;

__host__ __device__ inline bool galaxy_fits(int T, int pad) {
  const int H = pad / 2 + 1;
  const int need = H * T + (pad <= 128 ? H * pad : 0);
  return pad >= 64 && pad <= 256 && need <= kGalLdsComplex;
}

__host__ __device__ inline size_t round64(size_t v) { return (v + 63) / 64 * 64; }

// Per-workgroup global scratch of the planes kernel for stamps whose
// planes do not fit the LDS pool (floats, 64-aligned regions): P, dP/dx,
// dP/dy | two horizontally shifted planes | spectrum columns (pad 256).
struct PreScratch {
  size_t P3, H2, S, total;
};

__host__ __device__ inline PreScratch pre_scratch(int T) {
  const int pad = fft_pad_for(T), H = pad / 2 + 1;
  const size_t TT = (size_t)T * T;
  PreScratch o;
  o.P3 = 0;
  o.H2 = round64(3 * TT);
  o.S = o.H2 + round64(2 * TT);
  o.total = o.S + (pad > 128 ? (size_t)2 * H * pad : 0);
  return o;
}

// A star item's scratch: its raw planes (P, dP/dx, dP/dy) then the
// horizontal pass, T^2 floats each, row stride T.
__host__ __device__ inline size_t star_scratch_floats(int T) {
  return round64(3 * (size_t)T * T) + 3 * (size_t)T * T;
}

__host__ __device__ inline size_t spectra_floats(int T) {
  const int pad = fft_pad_for(T);
  return (size_t)3 * (pad / 2 + 1) * pad;
}

bool prerender_supports(int T) { return galaxy_fits(T, fft_pad_for(T)); }
size_t prerender_scratch(int T) { return align256(pre_scratch(T).total * sizeof(float)); }
size_t prerender_spectra_floats(int T) { return spectra_floats(T); }

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
  // every lane of these DPP forms reads a lane of its own row, so the move
  // needs no old value and can fold into the instruction that consumes it
  if constexpr (H == 1) {
    r = __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad [1,0,3,2]
  } else if constexpr (H == 2) {
    r = __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad [2,3,0,1]
  } else if constexpr (H == 4) {
    const int t = __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true);  // i ^ 7
    r = __builtin_amdgcn_mov_dpp(t, 0x1B, 0xF, 0xF, true);             // i ^ 3
  } else if constexpr (H == 8) {
    const int t = __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, true);  // i ^ 15
    r = __builtin_amdgcn_mov_dpp(t, 0x141, 0xF, 0xF, true);            // i ^ 7
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
    // partner + sg * own as a product and a sum: sg = +-1 makes the product
    // exact, so this is the fused result, and the partner's DPP move folds
    // into the sum
#pragma clang fp contract(off)
    const float qr = lane_xor<h>(zr[e]), qi = lane_xor<h>(zi[e]);
    const float sr = sg * zr[e], si = sg * zi[e];
    re[e] = qr + sr;
    im[e] = qi + si;
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
// Taps and spectra: one workgroup per pre-rendered member.  Every member
// gets its Lanczos taps for the pass (band-independent: the sub-pixel
// offset is taken against the island window origin, where the evaluation
// kernel places the stamp at rint(x - x0) - T / 2); extended members also
// get their Fourier spectra.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(UPH_BLOCK) void prerender_spectra_kernel(IslandArgsPre a) {
  __shared__ GalParams<float> gp;
  __shared__ int s_skip, s_T;
  const int g = blockIdx.x, tid = threadIdx.x, mode = a.mode;
  const float* psrc = mode == kSplitEvalInit ? a.p0 : (mode == kSplitEvalTry ? a.p_try : a.p_out);
  const int isl = a.pre_isl[g];
  if ((mode == kSplitEvalGrad || mode == kSplitEvalTry) && a.conv_out[isl]) return;  // uniform
  const size_t slot = (size_t)a.mem_off[isl] + a.pre_slot[g];
  const float* p6 = psrc + slot * 6;
  if (tid == 0 || tid == 64) {
    // x taps on wave 0, y taps on wave 1
    const float d = tid == 0 ? p6[0] - a.x0b[isl] : p6[1] - a.y0b[isl];
    float* t = a.gtap + (size_t)g * 28 + (tid == 0 ? 0 : 14);
    lanczos_taps(d - rintf(d), t, t + 7);
  }
  if (a.pre_spec[g] < 0) return;  // star: no spectra (uniform)
  if (tid == 0) {
    int skip = 0;
    const int32_t* sp = a.shape + 4 * (size_t)isl;
    const int T = a.tier_T[sp[2]];
    s_T = T;
    if ((size_t)a.pre_spec[g] + spectra_floats(T) > a.fspec_elems) {
      atomicOr(a.err, 1);
      skip = 1;
    }
    if (!skip) {
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
  const bool grads = split_grads(mode);
  float* F = a.fspec + a.pre_spec[g];
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
// Planes: (member, band) items.
// ---------------------------------------------------------------------------
struct PlaneJob {
  int g, b, T, pad, logp, H, need;
  const float* p6;
  float* P3;      // P, dP/dx, dP/dy (LDS or the workgroup's global scratch)
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
__device__ __forceinline__ void forward_plane(const PlaneJob& j, const float* src,
                                              const WaveTw<E>& tw) {
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

// Part of a member's T x T stamp that lands in its island window: the
// window-covered box [sx0, sx1) x [sy0, sy1) in stamp coordinates, and the
// box [rx0, rx1) x [ry0, ry1) of raw-plane pixels the 7-tap shift reads
// for it (the window box grown by 3 and clipped to the stamp).
struct StampClip {
  int sx0, sx1, sy0, sy1, rx0, rx1, ry0, ry1;
};

// ---------------------------------------------------------------------------
// Planes pipeline: the horizontal Lanczos pass is applied to each output row
// of an inverse transform while the row is in registers, and only the rows
// and columns the window box needs are kept, so no raw plane is stored.
// Up to pad 128 and T = 55 the PSF planes, the line buffer, the spectrum and
// the two shifted-row buffers all live in one LDS pool; larger stamps keep
// the line buffer (and the pad-128 spectrum) in LDS and the rest in the
// workgroup's global scratch.
// ---------------------------------------------------------------------------
constexpr int kPoolFloats = 40704;  // LDS pool of the planes kernel (159 KiB)

__host__ __device__ inline bool planes_in_lds(int T, int pad) {
  const int H = pad / 2 + 1;
  return pad <= 128 && 2 * (H * T + H * pad) + 5 * T * T <= kPoolFloats;
}

// Horizontal 7-tap pass of one row held across the wave (element e of lane
// l is column l + 64 e): h[e] = sum_k L[k] v(c + k - 3) and, when `both`,
// hd[e] = sum_k dL[k] v(c + k - 3), over source columns in [c0, c1) (zero
// outside).  Neighbour columns come from the other lanes by __shfl.
template <int E>
__device__ __forceinline__ void lanczos_row(const float (&v)[E], const float* L, const float* dL,
                                            int c0, int c1, bool both, float (&h)[E],
                                            float (&hd)[E]) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int e = 0; e < E; ++e) h[e] = hd[e] = 0.0f;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const int d = k - 3;
    const int sl = (lane + d) & 63;
    const int wrap = lane + d < 0 ? -1 : (lane + d >= 64 ? 1 : 0);
    float s[E];
#pragma unroll
    for (int e = 0; e < E; ++e) s[e] = d == 0 ? v[e] : __shfl(v[e], sl, 64);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      float x = s[e];
      if (wrap < 0) x = e > 0 ? s[e > 0 ? e - 1 : 0] : 0.0f;
      if (wrap > 0) x = e + 1 < E ? s[e + 1 < E ? e + 1 : 0] : 0.0f;
      const int c = lane + 64 * e + d;
      if (c < c0 || c >= c1) x = 0.0f;
      h[e] = fmaf(L[k], x, h[e]);
      if (both) hd[e] = fmaf(dL[k], x, hd[e]);
    }
  }
}

// Column half of an inverse transform: A[k][r] = sum_m S[k][m] X[k][m]
// e^{+2 pi i m r / pad} for r < T, X the spectrum of output kind `kind`
// (0: F + core, 1..3: F1 d(expo)/d(logre, ee1, ee2), 4: Fn + dcore).
template <int E>
__device__ __forceinline__ void inverse_columns(const PlaneJob& j, int kind,
                                                const WaveTw<E>& tw) {
  constexpr int G = Batch<E>::G;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int T = j.T, H = j.H, pad = j.pad, logp = j.logp, n = H * pad;
  const float* X = j.F + (kind == 0 ? 0 : (kind == 4 ? 2 * n : n));
  const double c2 = -2.0 * 9.869604401089358;
  const int p = kind - 1;
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
}

// Row half of an inverse transform with the horizontal shift fused: rows
// [ry0, ry1) of the raw plane (Hermitian row transforms of A), filtered
// with L (and dL into Hb when Hb is given) and stored for the window-box
// columns as Ha[(r - ry0) * SW + c - sx0], SW = sx1 - sx0.
template <int E>
__device__ __forceinline__ void inverse_rows_hshift(const PlaneJob& j, const StampClip& cl,
                                                    const float* L, const float* dL, float* Ha,
                                                    float* Hb, const WaveTw<E>& tw) {
  constexpr int G = Batch<E>::G;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int T = j.T, pad = j.pad, logp = j.logp, H = j.H;
  const int SW = cl.sx1 - cl.sx0;
  const bool both = Hb != nullptr;
  const float norm = 1.0f / (float(pad) * float(pad));
  for (int r0 = cl.ry0 + wid * G; r0 < cl.ry1; r0 += kGalWaves * G) {
    float re[G][E], im[G][E];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int r = r0 + g;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = brev_bits(lane + 64 * e, logp);
        float2 y = make_float2(0.0f, 0.0f);
        if (r < cl.ry1) {
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
      if (r0 + g < cl.ry1) wave_fft<E>(re[g], im[g], tw);
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const int r = r0 + g;
      if (r >= cl.ry1) continue;  // uniform across the wave
      float v[E], h[E], hd[E];
#pragma unroll
      for (int e = 0; e < E; ++e) v[e] = re[g][e] * norm;
      lanczos_row<E>(v, L, dL, cl.rx0, cl.rx1, both, h, hd);
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int c = lane + 64 * e;
        if (c < cl.sx0 || c >= cl.sx1) continue;
        const int o = (r - cl.ry0) * SW + c - cl.sx0;
        Ha[o] = h[e];
        if (both) Hb[o] = hd[e];
      }
    }
  }
  __syncthreads();
}

// Vertical 7-tap pass of a horizontally shifted buffer Hs (rows [ry0, ry1),
// SW columns) into the window box of an output plane (side x side, from
// its corner): out = scale * sum_k L[k] Hs(r + k - 3), or out += that sum.
template <int kThreads>
__device__ __forceinline__ void vshift_plane(const float* Hs, const StampClip& cl, const float* L,
                                             float* out, int side, float scale, bool add) {
  const int SW = cl.sx1 - cl.sx0, SH = cl.sy1 - cl.sy0;
  for (int q = threadIdx.x; q < SH * SW; q += kThreads) {
    const int rw = cl.sy0 + q / SW, cc = q % SW;
    float v = 0.0f;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int rr = rw + k - 3;
      if (rr < cl.ry0 || rr >= cl.ry1) continue;
      v = fmaf(L[k], Hs[(rr - cl.ry0) * SW + cc], v);
    }
    const int e = (q / SW) * side + cc;
    out[e] = add ? out[e] + v : v * scale;
  }
}

// All planes of an extended member for one band.  The raw planes are R0 =
// irfft2(S_P (F + core)) (the model), R1, R2 = irfft2(S_dP/dx (F + core)),
// irfft2(S_dP/dy (F + core)) (position), R3..R5 = irfft2(S_P F1 d expo /
// d(logre, ee1, ee2)) and R6 = irfft2(S_P (Fn + dcore)) (Sersic n), S_X the
// spectrum of the PSF plane X (P3 holds P, dP/dx, dP/dy).  The shifted
// planes: 0 = Ly (Lx R0); 1 = Ly (dLx R0) + Ly (Lx R1); 2 = dLy (Lx R0) +
// Ly (Lx R2); 3..5 = Ly (Lx R); 6 = dn/du Ly (Lx R6), each over the window
// box only; `out` holds the member's 7 planes of side x side floats for the
// band.  Ha, Hb hold horizontally shifted rows (in LDS when kLds, else in
// global scratch); each (E, kLds) form has its own call site so that it is
// inlined with its pointers' address spaces known.
template <int E, bool kLds>
__device__ __forceinline__ void render_planes_fused(const PlaneJob& j, const StampClip& cl,
                                                    const float* taps, float dndu, float* Ha,
                                                    float* Hb, float* out, int side) {
  WaveTw<E> tw;
  wave_tw<E>(tw);
  const int TT = j.T * j.T, PP = side * side, need = j.need;
  const float* Lx = taps;
  const float* dLx = taps + 7;
  const float* Ly = taps + 14;
  const float* dLy = taps + 21;
  forward_plane<E>(j, j.P3, tw);
  inverse_columns<E>(j, 0, tw);
  inverse_rows_hshift<E>(j, cl, Lx, dLx, Ha, (need & 1) ? Hb : nullptr, tw);
  vshift_plane<kPreBlock>(Ha, cl, Ly, out, side, 1.0f, false);
  if (need & 2) vshift_plane<kPreBlock>(Ha, cl, dLy, out + (size_t)2 * PP, side, 1.0f, false);
  if (need & 1) vshift_plane<kPreBlock>(Hb, cl, Ly, out + (size_t)PP, side, 1.0f, false);
  for (int q = 0; q < 4; ++q) {
    const int bit = q < 3 ? (4 << q) : 32;
    if (!(need & bit)) continue;
    inverse_columns<E>(j, 1 + q, tw);
    inverse_rows_hshift<E>(j, cl, Lx, dLx, Ha, nullptr, tw);
    vshift_plane<kPreBlock>(Ha, cl, Ly, out + (size_t)(3 + q) * PP, side, q < 3 ? 1.0f : dndu,
                            false);
  }
  for (int q = 0; q < 2; ++q) {
    if (!(need & (1 << q))) continue;
    forward_plane<E>(j, j.P3 + (size_t)(1 + q) * TT, tw);
    inverse_columns<E>(j, 0, tw);
    inverse_rows_hshift<E>(j, cl, Lx, dLx, Ha, nullptr, tw);
    vshift_plane<kPreBlock>(Ha, cl, Ly, out + (size_t)(1 + q) * PP, side, 1.0f, true);
  }
  __syncthreads();
}

// Separable Lanczos-3 shift of the raw planes R (plane 0 the stamp, 1 and 2
// its x and y raw tangents, 3..6 the shape and n tangents) into out:
// horizontal pass into Hh, vertical pass into out, over the window box of
// the stamp (raw inputs outside [0, T) are zero; only the box c.r* is read).
// Plane 1 of out gets the x tap derivative applied to plane 0, plane 2 the y
// tap derivative; plane 6 is scaled by dn/du.  out holds planes of side x
// side floats, the window box stored from its corner (c.sx0, c.sy0).  Planes
// of tangents not in `need` are not written.
template <int kThreads>
__device__ void shift_planes(const float* R, float* Hh, float* out, int T, int side, int need,
                             const float* taps, float dndu, const StampClip& c) {
  const int tid = threadIdx.x, TT = T * T;
  float Lx[7], dLx[7], Ly[7], dLy[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    Lx[k] = taps[k];
    dLx[k] = taps[7 + k];
    Ly[k] = taps[14 + k];
    dLy[k] = taps[21 + k];
  }
  const int SW = c.sx1 - c.sx0, RH = c.ry1 - c.ry0, SH = c.sy1 - c.sy0;
  for (int q = tid; q < RH * SW; q += kThreads) {
    const int rr = c.ry0 + q / SW, cx = c.sx0 + q % SW, e = rr * T + cx;
    float h[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int cc = cx + k - 3;
      if (cc < c.rx0 || cc >= c.rx1) continue;
      const int o = rr * T + cc;
      h[0] = fmaf(Lx[k], R[o], h[0]);
      if (need & 1) h[1] = fmaf(Lx[k], R[TT + o], fmaf(dLx[k], R[o], h[1]));
#pragma unroll
      for (int j = 1; j < 6; ++j)
        if (need & (1 << j)) h[1 + j] = fmaf(Lx[k], R[(size_t)(1 + j) * TT + o], h[1 + j]);
    }
    Hh[e] = h[0];
#pragma unroll
    for (int j = 0; j < 6; ++j)
      if (need & (1 << j)) Hh[(size_t)(1 + j) * TT + e] = h[1 + j];
  }
  __syncthreads();
  const int PP = side * side;
  for (int q = tid; q < SH * SW; q += kThreads) {
    const int rw = c.sy0 + q / SW, cc = c.sx0 + q % SW, e = (q / SW) * side + q % SW;
    float v[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int rr = rw + k - 3;
      if (rr < c.ry0 || rr >= c.ry1) continue;
      const int o = rr * T + cc;
      const float h0 = Hh[o];
      v[0] = fmaf(Ly[k], h0, v[0]);
      if (need & 2) v[2] = fmaf(Ly[k], Hh[(size_t)2 * TT + o], fmaf(dLy[k], h0, v[2]));
#pragma unroll
      for (int j = 0; j < 6; ++j)
        if (j != 1 && (need & (1 << j))) v[1 + j] = fmaf(Ly[k], Hh[(size_t)(1 + j) * TT + o], v[1 + j]);
    }
    out[e] = v[0];
#pragma unroll
    for (int j = 0; j < 6; ++j)
      if (need & (1 << j)) out[(size_t)(1 + j) * PP + e] = j == 5 ? v[6] * dndu : v[1 + j];
  }
  __syncthreads();
}

// Setup shared by the planes kernels, from any one thread: the item's
// member, stamp tier, clip box and validity.  Returns false when the item
// has nothing to render (converged island, stamp outside the window, or a
// buffer overflow, which also raises the error flag).
struct PreItem {
  int g, b, isl, T, side, need;
  size_t slot;
  StampClip clip;
};

__device__ __forceinline__ bool pre_item(const IslandArgsPre& a, int item, bool star,
                                         const float* psrc, PreItem& it) {
  const int nb = a.nb;
  it.g = item / nb;
  it.b = item % nb;
  it.isl = a.pre_isl[it.g];
  const int mode = a.mode;
  if ((mode == kSplitEvalGrad || mode == kSplitEvalTry) && a.conv_out[it.isl]) return false;
  const int32_t* sp = a.shape + 4 * (size_t)it.isl;
  const int T = a.tier_T[sp[2]], B = sp[1];
  it.T = T;
  it.side = pre_side(T, B);
  it.slot = (size_t)a.mem_off[it.isl] + a.pre_slot[it.g];
  int need = 0;
  if (split_grads(mode))
    for (int q = 0; q < 6; ++q)
      if (a.keep[it.slot * 6 + q]) need |= 1 << q;
  it.need = star ? need & 3 : need;
  const float* p6 = psrc + it.slot * 6;
  const float xg = p6[0] - a.x0b[it.isl], yg = p6[1] - a.y0b[it.isl];
  const int ix = int(rintf(xg)) - T / 2, iy = int(rintf(yg)) - T / 2;
  StampClip& c = it.clip;
  c.sx0 = max(0, ix) - ix;
  c.sx1 = min(B, ix + T) - ix;
  c.sy0 = max(0, iy) - iy;
  c.sy1 = min(B, iy + T) - iy;
  c.rx0 = max(0, c.sx0 - 3);
  c.rx1 = min(T, c.sx1 + 3);
  c.ry0 = max(0, c.sy0 - 3);
  c.ry1 = min(T, c.sy1 + 3);
  // a stamp entirely outside the window is never placed
  return c.sx1 > c.sx0 && c.sy1 > c.sy0;
}

__device__ __forceinline__ bool pre_item_fits(const IslandArgsPre& a, const PreItem& it,
                                              bool star) {
  const int T = it.T;
  // the launch's member kind must be the member's own (its spectra offset
  // and plane count follow from it)
  return (a.stype[it.slot] == 0) == star && (a.pre_spec[it.g] < 0) == star &&
         (star || galaxy_fits(T, fft_pad_for(T))) &&
         (star ? star_scratch_floats(T) * sizeof(float) <= a.star_stride
               : pre_scratch(T).total * sizeof(float) <= a.scratch_per_block) &&
         (size_t)a.pre_off[it.g] + (size_t)a.nb * pre_planes(star) * it.side * it.side <=
             a.pbuf_elems &&
         (star || (size_t)a.pre_spec[it.g] + spectra_floats(T) <= a.fspec_elems);
}

// PSF stamp of band b at the member position, and d/dx, d/dy when `gxy`,
// over the box [x0, x0 + W) x [y0, y0 + H) of the T x T stamp into
// dst[0], dst[T^2], dst[2 T^2] (row stride T).
template <int kThreads>
__device__ __forceinline__ void psf_box(const float* Cf, int nt, int T, const float* bas, float xs,
                                        float ys, bool gxy, int x0, int y0, int W, int H,
                                        float* dst) {
  const int TT = T * T;
  for (int q = threadIdx.x; q < W * H; q += kThreads) {
    const int e = (y0 + q / W) * T + x0 + q % W;
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
    dst[e] = v;
    if (gxy) {
      dst[TT + e] = vx / xs;
      dst[2 * TT + e] = vy / ys;
    }
  }
}

// Extended members: (member, band) items [0, n_pre_ext * nb) from a work
// queue; the whole stamp goes through the Fourier mixture in LDS.
__global__ __launch_bounds__(kPreBlock) void prerender_planes_kernel(IslandArgsPre a) {
  __shared__ float2 lds[kPoolFloats / 2];
  __shared__ float bas[63];
  __shared__ float taps[28];
  __shared__ int s_item, s_ok;
  __shared__ PreItem s_it;
  __shared__ float s_dndu, s_xs, s_ys;
  const int tid = threadIdx.x, mode = a.mode, nb = a.nb, nt = a.nt;
  float* base = reinterpret_cast<float*>(a.scratch + (size_t)blockIdx.x * a.scratch_per_block);
  const int n_items = a.n_pre_ext * nb;
  const float* psrc = mode == kSplitEvalInit ? a.p0 : (mode == kSplitEvalTry ? a.p_try : a.p_out);
  for (;;) {
    if (tid == 0) {
      const int item = atomicAdd(a.pre_counter, 1);
      s_item = item;
      if (item < n_items) {
        PreItem it;
        int ok = pre_item(a, item, false, psrc, it);
        if (ok && !pre_item_fits(a, it, false)) {
          atomicOr(a.err, 1);
          ok = 0;
        }
        s_it = it;
        s_ok = ok;
        if (ok) {
          const float* geo = a.psf_geo + 4 * it.b;
          const float* p6 = psrc + it.slot * 6;
          const float X = (p6[0] - geo[0]) / geo[1], Y = (p6[1] - geo[2]) / geo[3];
          psf_basis(X, Y, nt, bas, bas + 21, bas + 42);
          s_xs = geo[1];
          s_ys = geo[3];
          float dndu = 0.0f;
          if (a.free_n[it.slot]) {
            const double u = double(p6[5]);
            const double n = kNMin + (kNMax - kNMin) / (1.0 + exp(-u));
            dndu = float((n - kNMin) * (kNMax - n) / (kNMax - kNMin));
          }
          s_dndu = dndu;
        }
      }
    }
    __syncthreads();
    const int item = s_item;
    if (item >= n_items) break;
    if (!s_ok) {
      __syncthreads();
      continue;
    }
    const int g = s_it.g, b = s_it.b, T = s_it.T, need = s_it.need;
    const size_t slot = s_it.slot;
    if (tid < 28) taps[tid] = a.gtap[(size_t)g * 28 + tid];
    const int pad = fft_pad_for(T), H = pad / 2 + 1;
    const float* Cf = a.tier_coefs[a.shape[4 * (size_t)s_it.isl + 2]] + (size_t)b * nt * T * T;
    PlaneJob j;
    j.g = g;
    j.b = b;
    j.T = T;
    j.pad = pad;
    j.logp = ilog2(pad);
    j.H = H;
    j.need = need;
    j.p6 = psrc + slot * 6;
    j.A = lds;
    j.F = a.fspec + a.pre_spec[g];
    j.dt = a.gdt + (size_t)g * 9;
    const int side = s_it.side;
    float* out = a.pbuf + a.pre_off[g] + (size_t)b * 7 * side * side;
    if (planes_in_lds(T, pad)) {
      // LDS pool: line buffer (H, T), spectrum (H, pad), Ha, Hb (T^2 each), P3
      float* Ha = reinterpret_cast<float*>(lds + H * T + H * pad);
      float* Hb = Ha + T * T;
      float* P3 = Hb + T * T;
      psf_box<kPreBlock>(Cf, nt, T, bas, s_xs, s_ys, need & 3, 0, 0, T, T, P3);
      __syncthreads();
      j.P3 = P3;
      j.S = lds + H * T;
      if (pad == 64)
        render_planes_fused<1, true>(j, s_it.clip, taps, s_dndu, Ha, Hb, out, side);
      else
        render_planes_fused<2, true>(j, s_it.clip, taps, s_dndu, Ha, Hb, out, side);
    } else {
      // line buffer (and the spectrum up to pad 128) in LDS, the rest in the
      // workgroup's global scratch
      const PreScratch L = pre_scratch(T);
      float* P3 = base + L.P3;
      float* Ha = base + L.H2;
      float* Hb = Ha + T * T;
      psf_box<kPreBlock>(Cf, nt, T, bas, s_xs, s_ys, need & 3, 0, 0, T, T, P3);
      __syncthreads();
      j.P3 = P3;
      if (pad <= 128) {
        j.S = lds + H * T;
        render_planes_fused<2, false>(j, s_it.clip, taps, s_dndu, Ha, Hb, out, side);
      } else {
        j.S = reinterpret_cast<float2*>(base + L.S);
        render_planes_fused<4, false>(j, s_it.clip, taps, s_dndu, Ha, Hb, out, side);
      }
    }
  }
}

// Stars: (member, band) items [n_pre_ext * nb, n_pre * nb), strided over
// the grid.  A star needs only the PSF pixels the shift reads for the
// window box, and no LDS beyond its basis and taps, so several workgroups
// share a CU.
constexpr int kStarBlock = 256;
constexpr long long kStarMaxBlocks = 8192;

__global__ __launch_bounds__(kStarBlock) void prerender_stars_kernel(IslandArgsPre a) {
  __shared__ float bas[63];
  __shared__ float taps[28];
  const int tid = threadIdx.x, mode = a.mode, nb = a.nb, nt = a.nt;
  float* base = reinterpret_cast<float*>(a.scratch + (size_t)blockIdx.x * a.star_stride);
  const int i0 = a.n_pre_ext * nb, i1 = a.n_pre * nb;
  const float* psrc = mode == kSplitEvalInit ? a.p0 : (mode == kSplitEvalTry ? a.p_try : a.p_out);
  for (int item = i0 + blockIdx.x; item < i1; item += gridDim.x) {
    // every thread derives the (uniform) item setup itself
    PreItem it;
    if (!pre_item(a, item, true, psrc, it)) continue;
    if (!pre_item_fits(a, it, true)) {
      if (tid == 0) atomicOr(a.err, 1);
      continue;
    }
    const int T = it.T, b = it.b;
    const float* geo = a.psf_geo + 4 * b;
    const float* p6 = psrc + it.slot * 6;
    if (tid < nt) {
      // basis term t = tid of the spatial polynomial, as psf_basis forms it
      const float X = (p6[0] - geo[0]) / geo[1], Y = (p6[1] - geo[2]) / geo[3];
      const int m = kTermM[tid], n = kTermN[tid];
      float px[6], py[6];
      px[0] = 1.0f;
      py[0] = 1.0f;
#pragma unroll
      for (int i = 1; i < 6; ++i) {
        px[i] = px[i - 1] * X;
        py[i] = py[i - 1] * Y;
      }
      bas[tid] = px[m] * py[n];
      bas[21 + tid] = m > 0 ? float(m) * px[m - 1] * py[n] : 0.0f;
      bas[42 + tid] = n > 0 ? float(n) * px[m] * py[n - 1] : 0.0f;
    } else if (tid >= 64 && tid < 64 + 28) {
      taps[tid - 64] = a.gtap[(size_t)it.g * 28 + tid - 64];
    }
    __syncthreads();
    float* R = base;
    const StampClip& c = it.clip;
    psf_box<kStarBlock>(a.tier_coefs[a.shape[4 * (size_t)it.isl + 2]] + (size_t)b * nt * T * T,
                        nt, T, bas, geo[1], geo[3], it.need & 3, c.rx0, c.ry0, c.rx1 - c.rx0,
                        c.ry1 - c.ry0, R);
    __syncthreads();
    shift_planes<kStarBlock>(R, base + round64(3 * (size_t)T * T),
                             a.pbuf + a.pre_off[it.g] + (size_t)b * 3 * it.side * it.side, T,
                             it.side, it.need, taps, 0.0f, c);
  }
}

void prerender_pass(const IslandArgsPre& a, int mode, int* counter, int nblk, hipStream_t s,
                    hipEvent_t ev_spec, hipEvent_t ev_ext) {
  IslandArgsPre p = a;
  p.mode = mode;
  p.pre_counter = counter;
  if (p.n_pre > 0)
    hipLaunchKernelGGL(prerender_spectra_kernel, dim3(p.n_pre), dim3(UPH_BLOCK), 0, s, p);
  if (ev_spec) (void)hipEventRecord(ev_spec, s);
  const long long n_ext = (long long)p.n_pre_ext * p.nb;
  const int g_ext = (int)(n_ext < nblk ? n_ext : nblk);
  if (g_ext > 0) hipLaunchKernelGGL(prerender_planes_kernel, dim3(g_ext), dim3(kPreBlock), 0, s, p);
  if (ev_ext) (void)hipEventRecord(ev_ext, s);
  // the star launch carves the same scratch into smaller per-workgroup
  // blocks (raw planes and horizontal pass of the largest tier only), so
  // more of its workgroups are resident
  int t_max = 0;
  for (int t = 0; t < 3; ++t) t_max = p.tier_T[t] > t_max ? p.tier_T[t] : t_max;
  p.star_stride = align256(star_scratch_floats(t_max) * sizeof(float));
  const size_t total = (size_t)nblk * p.scratch_per_block;
  const long long cap = (long long)(total / p.star_stride) < kStarMaxBlocks
                            ? (long long)(total / p.star_stride)
                            : kStarMaxBlocks;
  const long long n_star = (long long)(p.n_pre - p.n_pre_ext) * p.nb;
  const int g_star = (int)(n_star < cap ? n_star : cap);
  if (g_star > 0)
    hipLaunchKernelGGL(prerender_stars_kernel, dim3(g_star), dim3(kStarBlock), 0, s, p);
}

// ---------------------------------------------------------------------------
// Unit stamps of extended sources for the linear solve (source_model.py:
// 264-328, fft backend; never renormalised): one (source, band) item per
// workgroup visit through the planes pipeline with the model plane only --
// the PSF at the source position, its forward transform, the product with
// the mixture spectrum F + core (evaluated per item on the (H, pad) grid:
// float64 exponents, float32 exponentials), the inverse transform with the
// horizontal Lanczos pass fused into its rows, then the vertical pass into
// the T x T stamp.  Stars are left to the star launch.  kLds: everything in
// one LDS pool (pad <= 128 up to T = 55); otherwise the line buffer in LDS
// and the spectrum, mixture, PSF plane and shifted rows in the workgroup's
// global scratch.
// ---------------------------------------------------------------------------
constexpr int kStampPoolFloats = 39424;

__host__ __device__ inline bool stamps_in_lds(int T, int pad) {
  const int H = pad / 2 + 1;
  return pad <= 128 && 2 * (H * T + H * pad) + H * pad + 2 * T * T <= kStampPoolFloats;
}

__host__ __device__ inline size_t stamp_scratch_floats(int T) {
  const int pad = fft_pad_for(T), H = pad / 2 + 1;
  return round64((size_t)2 * H * pad) + round64((size_t)H * pad) + round64((size_t)T * T) +
         round64((size_t)T * T) + (pad > 256 ? (size_t)2 * H * T : 0);
}

// LDS of a galaxy-stamps form: the pool, or (pad 512) nothing beyond the
// small arrays, its line buffer then being in scratch too.
template <int E>
struct StampLds {
  static constexpr int floats = E == 8 ? 2 : kStampPoolFloats;
};

template <int E, bool kLds>
__global__ __launch_bounds__(kPreBlock) void galaxy_stamps_kernel(
    int n, int T, int nt, const float* __restrict__ coefs, const float* __restrict__ geo_all,
    const double* __restrict__ sersic, const int32_t* __restrict__ src_band,
    const float* __restrict__ xs_, const float* __restrict__ ys_, const float* __restrict__ gx0,
    const float* __restrict__ gy0, const int32_t* __restrict__ stype,
    const float* __restrict__ logre, const float* __restrict__ ee1,
    const float* __restrict__ ee2, const float* __restrict__ sersic_n, float* __restrict__ out,
    char* scratch, size_t per_block) {
  __shared__ float2 lds[StampLds<E>::floats / 2];
  __shared__ float bas[21];
  __shared__ float taps[28];
  __shared__ float s_xs, s_ys;
  __shared__ GalParams<float> gp;
  __shared__ StampClip clip;
  const int tid = threadIdx.x, TT = T * T;
  const int pad = fft_pad_for(T), H = pad / 2 + 1, np = H * pad;
  WaveTw<E> tw;
  wave_tw<E>(tw);
  if (tid == 0) {
    clip.sx0 = clip.sy0 = clip.rx0 = clip.ry0 = 0;
    clip.sx1 = clip.sy1 = clip.rx1 = clip.ry1 = T;
  }
  float2* S;
  float2* A = lds;
  float *F, *P, *Ha;
  if constexpr (kLds) {
    S = lds + H * T;
    F = reinterpret_cast<float*>(S + np);
    P = F + np;
    Ha = P + TT;
  } else {
    float* base = reinterpret_cast<float*>(scratch + (size_t)blockIdx.x * per_block);
    S = reinterpret_cast<float2*>(base);
    F = base + round64((size_t)2 * np);
    P = F + round64((size_t)np);
    Ha = P + round64((size_t)TT);
    if constexpr (E == 8) A = reinterpret_cast<float2*>(Ha + round64((size_t)TT));
  }
  PlaneJob j;
  j.g = 0;
  j.b = 0;
  j.T = T;
  j.pad = pad;
  j.logp = ilog2(pad);
  j.H = H;
  j.need = 0;
  j.p6 = nullptr;
  j.P3 = P;
  j.S = S;
  j.A = A;
  j.F = F;
  j.dt = nullptr;
  const double c2 = -2.0 * 9.869604401089358;  // -2 pi^2
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    if (stype[s] == 0) continue;  // uniform per workgroup
    const int b = src_band[s];
    if (tid == 0) {
      gp.n = double(sersic_n[s]);
      gp.dndu = 0.0;
      gp.K = int(sersic[3]);
      mog_lookup(sersic, gp.n, gp.a, gp.v, gp.core, gp.da, gp.dv, gp.dcore);
      ellipse_terms(double(logre[s]), double(ee1[s]), double(ee2[s]), gp.t, gp.dt);
    }
    if (tid == 64 || tid == 128) {
      const float d = tid == 64 ? xs_[s] - gx0[s] : ys_[s] - gy0[s];
      float* t = taps + (tid == 64 ? 0 : 14);
      lanczos_taps(d - rintf(d), t, t + 7);
    }
    if (tid == 192) {
      const float* geo = geo_all + 4 * b;
      const float X = (xs_[s] - geo[0]) / geo[1], Y = (ys_[s] - geo[2]) / geo[3];
      psf_basis(X, Y, nt, bas, (float*)nullptr, (float*)nullptr);
      s_xs = geo[1];
      s_ys = geo[3];
    }
    __syncthreads();
    // mixture spectrum with the core folded in
    for (int e = tid; e < np; e += kPreBlock) {
      const int k = e / pad, m = e % pad;
      const double v = double(k) / pad;
      const double w = double(m < pad / 2 ? m : m - pad) / pad;
      const double ex = c2 * (gp.t[0] * v * v + gp.t[1] * w * w + 2.0 * gp.t[2] * v * w);
      float f = 0.0f;
      for (int q = 0; q < gp.K; ++q) f = fmaf(float(gp.a[q]), expf(float(gp.v[q] * ex)), f);
      F[e] = f + float(gp.core);
    }
    psf_box<kPreBlock>(coefs + (size_t)b * nt * TT, nt, T, bas, s_xs, s_ys, false, 0, 0, T, T,
                       P);
    __syncthreads();
    forward_plane<E>(j, P, tw);
    inverse_columns<E>(j, 0, tw);
    inverse_rows_hshift<E>(j, clip, taps, taps + 7, Ha, nullptr, tw);
    vshift_plane<kPreBlock>(Ha, clip, taps + 14, out + (size_t)s * TT, T, 1.0f, false);
    __syncthreads();
  }
}

bool launch_galaxy_stamps(int n, int T, int nt, const float* coefs, const float* geo,
                          const double* sersic, const int32_t* src_band, const float* x,
                          const float* y, const float* gx0, const float* gy0,
                          const int32_t* stype, const float* logre, const float* ee1,
                          const float* ee2, const float* sersic_n, float* out, char* scratch,
                          size_t scratch_bytes, int n_cu, hipStream_t s) {
  const int pad = fft_pad_for(T);
  if (!galaxy_fits(T, pad) && pad != 512) return false;
  const bool in_lds = stamps_in_lds(T, pad);
  const size_t per = align256(stamp_scratch_floats(T) * sizeof(float));
  long long nblk = 2LL * n_cu;
  if (!in_lds) {
    if (scratch == nullptr || scratch_bytes < per) return false;
    const long long cap = (long long)(scratch_bytes / per);
    nblk = nblk < cap ? nblk : cap;
  }
  nblk = nblk < n ? nblk : n;
  if (nblk < 1) return true;
#define UPH_GAL_STAMPS(E_, L_)                                                                   \
  hipLaunchKernelGGL((galaxy_stamps_kernel<E_, L_>), dim3((int)nblk), dim3(kPreBlock), 0, s, n, \
                     T, nt, coefs, geo, sersic, src_band, x, y, gx0, gy0, stype, logre, ee1,   \
                     ee2, sersic_n, out, scratch, per)
  if (in_lds && pad == 64)
    UPH_GAL_STAMPS(1, true);
  else if (in_lds)
    UPH_GAL_STAMPS(2, true);
  else if (pad == 128)
    UPH_GAL_STAMPS(2, false);
  else if (pad == 256)
    UPH_GAL_STAMPS(4, false);
  else
    UPH_GAL_STAMPS(8, false);
#undef UPH_GAL_STAMPS
  return true;
}

}  // namespace uph

extern "C" size_t uph_prerender_scratch_bytes(int T) {
  return uph::prerender_supports(T) ? uph::prerender_scratch(T) : 0;
}

extern "C" size_t uph_prerender_spectra_floats(int T) { return uph::spectra_floats(T); }
