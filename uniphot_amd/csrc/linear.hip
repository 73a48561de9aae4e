// Grid-band linear solve (pass 1 / pass 2) and the batched stamp renderer.
//
// render_stamps_kernel: one workgroup per (source, band) stamp; the raw PSF
// and the horizontal Lanczos pass live in LDS.  Follows
// tract7dt/fit/grid_gpu/forced_phot_sparse.py:126-220 (all-star grids,
// unit-sum renormalisation) and source_model.py:264-328.
//
// band_solve_kernel: one workgroup per (grid, band) problem.  Waves reduce
// the per-source self terms (T^T W T diagonal, T^T W d, T^T W B) and the
// neighbour-pair overlaps T_s^T W T_t over the stamp intersection; the block
// reduces the background Gram over the whole window; the (n+M)^2 system is
// assembled in float64, ridge-regularised element-wise and solved by LU,
// with the inverse diagonal for the variances.  Follows
// forced_phot_sparse.py:342-554.
#include <hip/hip_runtime.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"
#include "uph_galaxy.hpp"
#include "uph_island.hpp"

namespace uph {

__global__ __launch_bounds__(UPH_BLOCK) void render_stamps_kernel(
    int n, int T, int nt, const float* __restrict__ coefs, const float* __restrict__ geo_all,
    const int32_t* __restrict__ src_band, const float* __restrict__ xs_, const float* __restrict__ ys_,
    const float* __restrict__ gx0, const float* __restrict__ gy0, const int32_t* __restrict__ stype,
    int renorm, float* __restrict__ out) {
  extern __shared__ float lds[];
  __shared__ float taps[14];
  __shared__ float bas[21];
  __shared__ float red[UPH_WAVES];
  const int s = blockIdx.x, tid = threadIdx.x, TT = T * T;
  if (s >= n) return;
  if (stype && stype[s] != 0) return;  // extended sources: render_galaxy_stamps_kernel
  const int b = src_band[s];
  const float* geo = geo_all + 4 * b;
  const float x = xs_[s], y = ys_[s];
  const float xg = x - gx0[s], yg = y - gy0[s];
  const float dx = xg - rintf(xg), dy = yg - rintf(yg);
  if (tid == 0) lanczos_taps(dx, taps, (float*)nullptr);
  if (tid == 64) lanczos_taps(dy, taps + 7, (float*)nullptr);
  if (tid == 128) {
    const float X = (x - geo[0]) / geo[1], Y = (y - geo[2]) / geo[3];
    psf_basis(X, Y, nt, bas, (float*)nullptr, (float*)nullptr);
  }
  __syncthreads();
  float* raw = lds;
  float* h = lds + TT;
  const float* C = coefs + (size_t)b * nt * TT;
  float part = 0.0f;
  for (int e = tid; e < TT; e += UPH_BLOCK) {
    float v = 0.0f;
    for (int t = 0; t < nt; ++t) v = fmaf(C[(size_t)t * TT + e], bas[t], v);
    raw[e] = v;
    part += v;
  }
  if (renorm) {
    const float tot = block_sum(part, red);
    const float den = fmaxf(tot, 1e-12f);
    for (int e = tid; e < TT; e += UPH_BLOCK) raw[e] = raw[e] / den;
  }
  __syncthreads();
  float Lx[7], Ly[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    Lx[k] = taps[k];
    Ly[k] = taps[7 + k];
  }
  for (int e = tid; e < TT; e += UPH_BLOCK) {
    const int r = e / T, c = e % T;
    float v = 0.0f;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int cc = c + k - 3;
      if (cc >= 0 && cc < T) v = fmaf(Lx[k], raw[r * T + cc], v);
    }
    h[e] = v;
  }
  __syncthreads();
  float* o = out + (size_t)s * TT;
  for (int e = tid; e < TT; e += UPH_BLOCK) {
    const int r = e / T, c = e % T;
    float v = 0.0f;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int rr = r + k - 3;
      if (rr >= 0 && rr < T) v = fmaf(Ly[k], h[rr * T + c], v);
    }
    o[e] = v;
  }
}

// Extended-source stamps (EXP/DEV/SERSIC): raw PSF at (x, y), analytic-FT
// galaxy convolution, separable Lanczos-3 shift; never renormalised
// (source_model.py:264-328 fft backend, forced_phot_sparse.py:624-637).
// Persistent: each workgroup loops over the stamps with its own workspace.
__global__ __launch_bounds__(UPH_BLOCK) void render_galaxy_stamps_kernel(
    int n, int T, int nt, const float* __restrict__ coefs, const float* __restrict__ geo_all,
    const double* __restrict__ sersic, const int32_t* __restrict__ src_band,
    const float* __restrict__ xs_, const float* __restrict__ ys_, const float* __restrict__ gx0,
    const float* __restrict__ gy0, const int32_t* __restrict__ stype,
    const float* __restrict__ logre, const float* __restrict__ ee1,
    const float* __restrict__ ee2, const float* __restrict__ sersic_n, int stars_too, int renorm,
    float* __restrict__ out, char* scratch, size_t per_block) {
  __shared__ GalParams<float> gp;
  __shared__ float taps[14];
  __shared__ float bas[21];
  __shared__ float red[UPH_WAVES];
  const int tid = threadIdx.x, TT = T * T;
  float* ws = reinterpret_cast<float*>(scratch + (size_t)blockIdx.x * per_block);
  const GalLayout L = galaxy_layout(T);
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    const bool star = stype == nullptr || stype[s] == 0;
    if (star && !stars_too) continue;  // uniform per block
    const int b = src_band[s];
    const float* geo = geo_all + 4 * b;
    const float x = xs_[s], y = ys_[s];
    const float xg = x - gx0[s], yg = y - gy0[s];
    const float dx = xg - rintf(xg), dy = yg - rintf(yg);
    if (tid == 0 && !star) {
      gp.n = double(sersic_n[s]);
      gp.dndu = 0.0;
      gp.K = int(sersic[3]);
      mog_lookup(sersic, gp.n, gp.a, gp.v, gp.core, gp.da, gp.dv, gp.dcore);
      ellipse_terms(double(logre[s]), double(ee1[s]), double(ee2[s]), gp.t, gp.dt);
    }
    if (tid == 64) lanczos_taps(dx, taps, (float*)nullptr);
    if (tid == 128) lanczos_taps(dy, taps + 7, (float*)nullptr);
    if (tid == 192) {
      const float X = (x - geo[0]) / geo[1], Y = (y - geo[2]) / geo[3];
      psf_basis(X, Y, nt, bas, (float*)nullptr, (float*)nullptr);
    }
    __syncthreads();
    const float* C = coefs + (size_t)b * nt * TT;
    float part = 0.0f;
    float* rawp = star ? ws + L.G : ws + L.P;
    for (int e = tid; e < TT; e += UPH_BLOCK) {
      float v = 0.0f;
      for (int t = 0; t < nt; ++t) v = fmaf(C[(size_t)t * TT + e], bas[t], v);
      rawp[e] = v;
      part += v;
    }
    if (star && renorm) {
      const float den = fmaxf(block_sum(part, red), 1e-12f);
      for (int e = tid; e < TT; e += UPH_BLOCK) rawp[e] = rawp[e] / den;
    }
    __syncthreads();
    if (!star) galaxy_raw<float, true>(gp, T, 0, ws);
    const float* G = ws + L.G;
    float* h = ws + L.Hh;
    for (int e = tid; e < TT; e += UPH_BLOCK) {
      const int r = e / T, c = e % T;
      float v = 0.0f;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        const int cc = c + k - 3;
        if (cc >= 0 && cc < T) v = fmaf(taps[k], G[r * T + cc], v);
      }
      h[e] = v;
    }
    __syncthreads();
    float* o = out + (size_t)s * TT;
    for (int e = tid; e < TT; e += UPH_BLOCK) {
      const int r = e / T, c = e % T;
      float v = 0.0f;
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        const int rr = r + k - 3;
        if (rr >= 0 && rr < T) v = fmaf(taps[7 + k], h[rr * T + c], v);
      }
      o[e] = v;
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------
struct BandScratch {
  size_t A, X, rhs, piv, per_problem;
};

__host__ __device__ inline BandScratch band_scratch_layout(int max_src, int M) {
  BandScratch s;
  const size_t n = (size_t)max_src + M;
  size_t off = 0;
  s.A = off;
  off = align256(off + n * n * 8);
  s.X = off;
  off = align256(off + n * n * 8);
  s.rhs = off;
  off = align256(off + n * 8);
  s.piv = off;
  off = align256(off + n * 4 + 16);
  s.per_problem = off;
  return s;
}

struct BandArgs {
  int n_prob, T, order, M, k_max, H, W, img_x0, img_y0;
  const int64_t* src_off;
  const int32_t* win;
  const float* data;
  const float* invvar;
  const uint8_t* good;
  double bxo, bxs, byo, bys;
  const float* stamps;
  const int32_t* ix_top;
  const int32_t* iy_top;
  const int32_t* nbr;
  float reg;
  float* fluxes;
  float* flux_var;
  float* bg;
  float* bg_var;
  char* scratch;
  size_t per_problem;
};

// float32 background basis at full-image pixel (px, py): float64
// normalisation rounded to float32 (bg_polynomial.py:125-178).
template <int M>
__device__ __forceinline__ void bg_row(const BandArgs& a, int px, int py, float* out) {
  const double X = (double(px) - a.bxo) / a.bxs;
  const double Y = (double(py) - a.byo) / a.bys;
  constexpr int m_[6] = {0, 1, 0, 2, 1, 0};
  constexpr int n_[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
  for (int t = 0; t < M; ++t) {
    double v = 1.0;
    for (int i = 0; i < m_[t]; ++i) v *= X;
    for (int i = 0; i < n_[t]; ++i) v *= Y;
    out[t] = float(v);
  }
}

template <int M>
__global__ __launch_bounds__(UPH_BLOCK) void band_solve_kernel(BandArgs a) {
  __shared__ float red[UPH_WAVES * 32];
  __shared__ int s_piv;
  const int q = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int band = a.win[q * 5], gx0 = a.win[q * 5 + 1], gy0 = a.win[q * 5 + 2];
  const int gh = a.win[q * 5 + 3], gw = a.win[q * 5 + 4];
  const int64_t s0 = a.src_off[q];
  const int n = int(a.src_off[q + 1] - s0);
  const int NT = n + M, T = a.T, TT = T * T;
  char* base = a.scratch + (size_t)q * a.per_problem;
  if ((size_t)NT * NT * 16 + (size_t)NT * 12 > a.per_problem) {
    // workspace too small for this problem: report NaN rather than overrun
    for (int i = tid; i < NT; i += UPH_BLOCK) {
      const float nanv = __int_as_float(0x7fc00000);
      if (i < n) {
        a.fluxes[s0 + i] = nanv;
        a.flux_var[s0 + i] = nanv;
      } else {
        a.bg[(size_t)q * M + i - n] = nanv;
        a.bg_var[(size_t)q * M + i - n] = nanv;
      }
    }
    return;
  }
  double* A = reinterpret_cast<double*>(base);
  double* X = A + (size_t)NT * NT;
  double* rhs = X + (size_t)NT * NT;
  int* piv = reinterpret_cast<int*>(rhs + NT);
  const size_t plane = (size_t)a.H * a.W;
  const float* D = a.data + (size_t)band * plane;
  const float* IV = a.invvar + (size_t)band * plane;
  const uint8_t* GD = a.good + (size_t)band * plane;
  for (size_t e = tid; e < (size_t)NT * NT; e += UPH_BLOCK) A[e] = 0.0;
  __syncthreads();
  // ---- per-source self terms, one wave per source ----
  for (int s = wid; s < n; s += UPH_WAVES) {
    const float* st = a.stamps + (size_t)(s0 + s) * TT;
    const int ix = a.ix_top[s0 + s], iy = a.iy_top[s0 + s];
    float dg = 0.0f, td = 0.0f, tb[M];
#pragma unroll
    for (int m = 0; m < M; ++m) tb[m] = 0.0f;
    for (int e = lane; e < TT; e += 64) {
      const int py = iy + e / T, px = ix + e % T;
      if (py < 0 || py >= gh || px < 0 || px >= gw) continue;
      const size_t gi = (size_t)(gy0 + py) * a.W + gx0 + px;
      const float wv = GD[gi] ? IV[gi] : 0.0f;
      if (wv == 0.0f) continue;
      const float sv = st[e];
      const float ws = wv * sv;
      dg = fmaf(ws, sv, dg);
      td = fmaf(ws, D[gi], td);
      float br[M];
      bg_row<M>(a, a.img_x0 + gx0 + px, a.img_y0 + gy0 + py, br);
#pragma unroll
      for (int m = 0; m < M; ++m) tb[m] = fmaf(ws, br[m], tb[m]);
    }
    dg = wave_sum(dg);
    td = wave_sum(td);
#pragma unroll
    for (int m = 0; m < M; ++m) tb[m] = wave_sum(tb[m]);
    if (lane == 0) {
      A[(size_t)s * NT + s] = double(dg);
      rhs[s] = double(td);
      for (int m = 0; m < M; ++m) {
        A[(size_t)s * NT + n + m] = double(tb[m]);
        A[(size_t)(n + m) * NT + s] = double(tb[m]);
      }
    }
  }
  // ---- neighbour overlaps, one wave per (source, slot) ----
  for (int u = wid; u < n * a.k_max; u += UPH_WAVES) {
    const int s = u / a.k_max;
    const int t = a.nbr[(s0 + s) * a.k_max + u % a.k_max];
    if (t < 0) continue;  // uniform within the wave
    const float* ss = a.stamps + (size_t)(s0 + s) * TT;
    const float* tt = a.stamps + (size_t)(s0 + t) * TT;
    const int ixs = a.ix_top[s0 + s], iys = a.iy_top[s0 + s];
    const int di = iys - a.iy_top[s0 + t], dj = ixs - a.ix_top[s0 + t];
    float acc = 0.0f;
    for (int e = lane; e < TT; e += 64) {
      const int ky = e / T, kx = e % T;
      const int ty = ky + di, tx = kx + dj;
      if (ty < 0 || ty >= T || tx < 0 || tx >= T) continue;
      const int py = iys + ky, px = ixs + kx;
      if (py < 0 || py >= gh || px < 0 || px >= gw) continue;
      const size_t gi = (size_t)(gy0 + py) * a.W + gx0 + px;
      const float wv = GD[gi] ? IV[gi] : 0.0f;
      acc = fmaf(wv * ss[e], tt[ty * T + tx], acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) A[(size_t)s * NT + t] += double(acc);
  }
  // ---- background Gram and target over the whole window ----
  {
    constexpr int nb2 = M * M + M;
    for (int c0 = 0; c0 < nb2; c0 += 32) {
      float acc[32];
#pragma unroll
      for (int i = 0; i < 32; ++i) acc[i] = 0.0f;
      for (int e = tid; e < gh * gw; e += UPH_BLOCK) {
        const int py = e / gw, px = e % gw;
        const size_t gi = (size_t)(gy0 + py) * a.W + gx0 + px;
        const float wv = GD[gi] ? IV[gi] : 0.0f;
        if (wv == 0.0f) continue;
        float br[M];
        bg_row<M>(a, a.img_x0 + gx0 + px, a.img_y0 + gy0 + py, br);
        const float dv = D[gi];
#pragma unroll
        for (int i = 0; i < 32; ++i) {
          const int k = c0 + i;
          if (k >= nb2) break;
          if (k < M * M) {
            const int r = k / M, c = k % M;
            acc[i] = fmaf(br[r] * wv, br[c], acc[i]);
          } else {
            acc[i] = fmaf(br[k - M * M] * wv, dv, acc[i]);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < 32; ++i) {
        const float v = wave_sum(acc[i]);
        if (lane == 0) red[wid * 32 + i] = v;
      }
      __syncthreads();
      if (tid < 32 && c0 + tid < nb2) {
        float v = red[tid];
        for (int w2 = 1; w2 < UPH_WAVES; ++w2) v += red[w2 * 32 + tid];
        const int k = c0 + tid;
        if (k < M * M)
          A[(size_t)(n + k / M) * NT + n + k % M] = double(v);
        else
          rhs[n + k - M * M] = double(v);
      }
      __syncthreads();
    }
  }
  __syncthreads();
  // element-wise ridge reg * max(|diag|, 1e-30)
  for (int i = tid; i < NT; i += UPH_BLOCK) {
    const double d = A[(size_t)i * NT + i];
    A[(size_t)i * NT + i] = d + double(a.reg) * fmax(fabs(d), 1e-30);
  }
  __syncthreads();
  block_lu(A, NT, piv, &s_piv);
  // inverse columns (one per thread) -> variance diagonal; then the solve
  for (int i = tid; i < NT; i += UPH_BLOCK) {
    double* x = X + (size_t)i * NT;
    for (int j = 0; j < NT; ++j) x[j] = 0.0;
    x[i] = 1.0;
    lu_solve_one(A, piv, NT, x);
  }
  if (tid == 0) lu_solve_one(A, piv, NT, rhs);
  __syncthreads();
  for (int i = tid; i < NT; i += UPH_BLOCK) {
    const double v = X[(size_t)i * NT + i];
    if (i < n) {
      a.fluxes[s0 + i] = float(rhs[i]);
      a.flux_var[s0 + i] = float(v);
    } else {
      a.bg[(size_t)q * M + i - n] = float(rhs[i]);
      a.bg_var[(size_t)q * M + i - n] = float(v);
    }
  }
}

// ---------------------------------------------------------------------------
// Wing images of large extended sources (solver_grid.py:1578-1626): problem
// q's window (band, x0, y0, h, w) gets, per pixel, the sum over its sources
// s in [src_off[q], src_off[q+1]) of flux[s] * (big[s] - small[s]), big the
// Tb x Tb unit stamp at window-local top-left (ix[s], iy[s]) and small the
// Ts x Ts stamp centred on the same pixel.  Sources are summed in order per
// pixel (deterministic); the windows of one launch must be disjoint.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(UPH_BLOCK) void wing_images_kernel(
    int Tb, int Ts, const int64_t* __restrict__ src_off, const int32_t* __restrict__ win,
    const float* __restrict__ big, const float* __restrict__ small,
    const int32_t* __restrict__ ix, const int32_t* __restrict__ iy,
    const float* __restrict__ flux, int H, int W, float* __restrict__ out) {
  const int q = blockIdx.x;
  const int band = win[q * 5], x0 = win[q * 5 + 1], y0 = win[q * 5 + 2];
  const int h = win[q * 5 + 3], w = win[q * 5 + 4];
  const int64_t s0 = src_off[q], s1 = src_off[q + 1];
  const int o = Tb / 2 - Ts / 2;
  float* img = out + (size_t)band * H * W;
  for (int e = blockIdx.y * UPH_BLOCK + threadIdx.x; e < h * w; e += gridDim.y * UPH_BLOCK) {
    const int py = e / w, px = e % w;
    float v = 0.0f;
    for (int64_t s = s0; s < s1; ++s) {
      const int u = px - ix[s], t = py - iy[s];
      if (u < 0 || u >= Tb || t < 0 || t >= Tb) continue;
      float d = big[(size_t)s * Tb * Tb + (size_t)t * Tb + u];
      const int us = u - o, ts = t - o;
      if (us >= 0 && us < Ts && ts >= 0 && ts < Ts) d -= small[(size_t)s * Ts * Ts + ts * Ts + us];
      v = fmaf(flux[s], d, v);
    }
    img[(size_t)(y0 + py) * W + x0 + px] = v;
  }
}

}  // namespace uph

static int n_terms_of_order(int order) { return (order + 1) * (order + 2) / 2; }

extern "C" size_t uph_render_scratch_bytes(int T, int n_blocks) {
  return (size_t)n_blocks * uph::galaxy_scratch_bytes(T, sizeof(float));
}

extern "C" int uph_render_stamps(int n, int T, int n_terms, const float* psf_coefs,
                                 const float* psf_geo, const double* sersic,
                                 const int32_t* src_band, const float* x, const float* y,
                                 const float* gx0, const float* gy0, const int32_t* stype,
                                 const float* logre, const float* ee1, const float* ee2,
                                 const float* sersic_n, int renorm, float* out, void* scratch,
                                 size_t scratch_bytes, void* stream) {
  UPH_CHECK_ARG(n >= 0, "uph_render_stamps: n=%d", n);
  UPH_CHECK_ARG(T >= 7 && T % 2 == 1 && T <= 511, "uph_render_stamps: T=%d must be odd in 7..511",
                T);
  UPH_CHECK_ARG(n_terms == 1 || n_terms == 3 || n_terms == 6 || n_terms == 10 || n_terms == 15 ||
                    n_terms == 21,
                "uph_render_stamps: unsupported n_terms=%d", n_terms);
  UPH_CHECK_ARG(psf_coefs && psf_geo && src_band && x && y && gx0 && gy0 && out,
                "uph_render_stamps: null pointer");
  const bool ext = stype != nullptr;
  UPH_CHECK_ARG(!ext || (logre && ee1 && ee2 && sersic_n && sersic && scratch),
                "uph_render_stamps: extended sources need logre, ee1, ee2, sersic_n, the Sersic "
                "table and scratch");
  if (n == 0) return UPH_OK;
  const size_t lds = (size_t)2 * T * T * sizeof(float);
  int dev = 0, lds_max = 0;
  int st0 = uph::hip_status(hipGetDevice(&dev), "uph_render_stamps hipGetDevice");
  if (st0 != UPH_OK) return st0;
  st0 = uph::hip_status(
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev),
      "uph_render_stamps attribute");
  if (st0 != UPH_OK) return st0;
  // stars whose stamp does not fit LDS go through the persistent workspace path
  const bool stars_lds = lds + 512 <= (size_t)lds_max;
  UPH_CHECK_ARG(stars_lds || scratch != nullptr,
                "uph_render_stamps: T=%d stamps need scratch (uph_render_scratch_bytes)", T);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (stars_lds) {
    hipLaunchKernelGGL(uph::render_stamps_kernel, dim3(n), dim3(UPH_BLOCK), lds, s, n, T,
                       n_terms, psf_coefs, psf_geo, src_band, x, y, gx0, gy0, stype, renorm,
                       out);
    int st = uph::hip_status(hipGetLastError(), "uph_render_stamps launch");
    if (st != UPH_OK || !ext) return st;
  }
  const size_t per = uph::galaxy_scratch_bytes(T, sizeof(float));
  const size_t nb_s = scratch_bytes / per;
  UPH_CHECK_ARG(nb_s >= 1, "uph_render_stamps: scratch of %zu bytes < one workgroup (%zu)",
                scratch_bytes, per);
  const int nblk = (int)(nb_s < (size_t)n ? nb_s : (size_t)n);
  hipLaunchKernelGGL(uph::render_galaxy_stamps_kernel, dim3(nblk), dim3(UPH_BLOCK), 0, s, n, T,
                     n_terms, psf_coefs, psf_geo, sersic, src_band, x, y, gx0, gy0, stype, logre,
                     ee1, ee2, sersic_n, stars_lds ? 0 : 1, renorm, out,
                     static_cast<char*>(scratch), per);
  return uph::hip_status(hipGetLastError(), "uph_render_stamps galaxy launch");
}

extern "C" size_t uph_band_scratch_bytes(int n_prob, int max_src, int bg_order) {
  const int M = n_terms_of_order(bg_order);
  return (size_t)n_prob * uph::band_scratch_layout(max_src, M).per_problem;
}

extern "C" int uph_band_solve(int n_prob, int T, int bg_order, int k_max, const int64_t* src_off,
                              const int32_t* win, int H, int W, int img_x0, int img_y0,
                              const float* data, const float* invvar, const uint8_t* good,
                              double bg_xo,
                              double bg_xs, double bg_yo, double bg_ys, const float* stamps,
                              const int32_t* ix_top, const int32_t* iy_top,
                              const int32_t* nbr_idx, float reg, float* fluxes, float* flux_var,
                              float* bg, float* bg_var, void* scratch, size_t scratch_bytes,
                              void* stream) {
  UPH_CHECK_ARG(n_prob >= 0, "uph_band_solve: n_prob=%d", n_prob);
  UPH_CHECK_ARG(bg_order >= 0 && bg_order <= 2, "uph_band_solve: bg_order=%d outside 0..2",
                bg_order);
  UPH_CHECK_ARG(T >= 1 && k_max >= 0 && H > 0 && W > 0, "uph_band_solve: bad sizes");
  UPH_CHECK_ARG(src_off && win && data && invvar && good && fluxes && flux_var && bg && bg_var &&
                    scratch,
                "uph_band_solve: null pointer");
  UPH_CHECK_ARG(bg_xs != 0.0 && bg_ys != 0.0, "uph_band_solve: zero background scale");
  if (n_prob == 0) return UPH_OK;
  uph::BandArgs a{};
  a.n_prob = n_prob;
  a.T = T;
  a.order = bg_order;
  a.M = n_terms_of_order(bg_order);
  a.k_max = k_max;
  a.H = H;
  a.W = W;
  a.img_x0 = img_x0;
  a.img_y0 = img_y0;
  a.src_off = src_off;
  a.win = win;
  a.data = data;
  a.invvar = invvar;
  a.good = good;
  a.bxo = bg_xo;
  a.bxs = bg_xs;
  a.byo = bg_yo;
  a.bys = bg_ys;
  a.stamps = stamps;
  a.ix_top = ix_top;
  a.iy_top = iy_top;
  a.nbr = nbr_idx;
  a.reg = reg;
  a.fluxes = fluxes;
  a.flux_var = flux_var;
  a.bg = bg;
  a.bg_var = bg_var;
  a.scratch = static_cast<char*>(scratch);
  a.per_problem = (scratch_bytes / (size_t)n_prob) & ~(size_t)255;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (a.M == 1)
    hipLaunchKernelGGL(uph::band_solve_kernel<1>, dim3(n_prob), dim3(UPH_BLOCK), 0, s, a);
  else if (a.M == 3)
    hipLaunchKernelGGL(uph::band_solve_kernel<3>, dim3(n_prob), dim3(UPH_BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL(uph::band_solve_kernel<6>, dim3(n_prob), dim3(UPH_BLOCK), 0, s, a);
  return uph::hip_status(hipGetLastError(), "uph_band_solve launch");
}

extern "C" int uph_wing_images(int n_prob, int Tb, int Ts, const int64_t* src_off,
                               const int32_t* win, const float* big, const float* small,
                               const int32_t* ix, const int32_t* iy, const float* flux, int H,
                               int W, float* out, void* stream) {
  UPH_CHECK_ARG(n_prob >= 0 && Tb >= Ts && Ts >= 1 && (Tb % 2) == 1 && (Ts % 2) == 1,
                "uph_wing_images: n_prob=%d Tb=%d Ts=%d (odd, Tb >= Ts)", n_prob, Tb, Ts);
  if (n_prob == 0) return UPH_OK;
  UPH_CHECK_ARG(src_off && win && big && small && ix && iy && flux && out,
                "uph_wing_images: null argument");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(uph::wing_images_kernel, dim3(n_prob, 16), dim3(UPH_BLOCK), 0, s, Tb, Ts,
                     src_off, win, big, small, ix, iy, flux, H, W, out);
  return uph::hip_status(hipGetLastError(), "uph_wing_images launch");
}
