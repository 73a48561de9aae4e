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

#include <mutex>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"
#include "uph_galaxy.hpp"
#include "uph_island.hpp"
#include "uph_prerender.hpp"

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
// Persistent: each workgroup loops over the stamps with its own workspace;
// stars only with stars_too, extended sources unless skip_gal.
__global__ __launch_bounds__(UPH_BLOCK) void render_galaxy_stamps_kernel(
    int n, int T, int nt, const float* __restrict__ coefs, const float* __restrict__ geo_all,
    const double* __restrict__ sersic, const int32_t* __restrict__ src_band,
    const float* __restrict__ xs_, const float* __restrict__ ys_, const float* __restrict__ gx0,
    const float* __restrict__ gy0, const int32_t* __restrict__ stype,
    const float* __restrict__ logre, const float* __restrict__ ee1,
    const float* __restrict__ ee2, const float* __restrict__ sersic_n, int stars_too,
    int skip_gal, int renorm, float* __restrict__ out, char* scratch, size_t per_block) {
  __shared__ GalParams<float> gp;
  __shared__ float taps[14];
  __shared__ float bas[21];
  __shared__ float red[UPH_WAVES];
  const int tid = threadIdx.x, TT = T * T;
  float* ws = reinterpret_cast<float*>(scratch + (size_t)blockIdx.x * per_block);
  const GalLayout L = galaxy_layout(T);
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    const bool star = stype == nullptr || stype[s] == 0;
    if (star ? !stars_too : skip_gal) continue;  // uniform per block
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
  int stage;  // weighted stamps staged in LDS (UPH_WAVES * T^2 floats of dynamic LDS)
  int lu_apart;  // the LU, inverse diagonal and solve run in band_lu_kernel
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
  // ---- per-source self terms and neighbour overlaps ----
  if (a.stage) {
    // rounds of one source per wave: the wave stages w * stamp of its source
    // (zero off the window and on bad pixels) in LDS with the self terms,
    // then takes each neighbour's dot product over the two stamps' overlap
    // rectangle only, reading the neighbour stamp once
    extern __shared__ float ws_lds[];
    float* wsb = ws_lds + (size_t)wid * TT;
    for (int r0 = 0; r0 < n; r0 += UPH_WAVES) {
      const int s = r0 + wid;
      if (s < n) {
        const float* st = a.stamps + (size_t)(s0 + s) * TT;
        const int ix = a.ix_top[s0 + s], iy = a.iy_top[s0 + s];
        float dg = 0.0f, td = 0.0f, tb[M];
#pragma unroll
        for (int m = 0; m < M; ++m) tb[m] = 0.0f;
        for (int e = lane; e < TT; e += 64) {
          const int py = iy + e / T, px = ix + e % T;
          float ws = 0.0f;
          if (py >= 0 && py < gh && px >= 0 && px < gw) {
            const size_t gi = (size_t)(gy0 + py) * a.W + gx0 + px;
            const float wv = GD[gi] ? IV[gi] : 0.0f;
            if (wv != 0.0f) {
              const float sv = st[e];
              ws = wv * sv;
              dg = fmaf(ws, sv, dg);
              td = fmaf(ws, D[gi], td);
              float br[M];
              bg_row<M>(a, a.img_x0 + gx0 + px, a.img_y0 + gy0 + py, br);
#pragma unroll
              for (int m = 0; m < M; ++m) tb[m] = fmaf(ws, br[m], tb[m]);
            }
          }
          wsb[e] = ws;
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
      __syncthreads();
      if (s < n) {
        const int ixs = a.ix_top[s0 + s], iys = a.iy_top[s0 + s];
        for (int k = 0; k < a.k_max; ++k) {
          const int t = a.nbr[(s0 + s) * a.k_max + k];
          if (t < 0) continue;  // uniform within the wave
          const float* tt = a.stamps + (size_t)(s0 + t) * TT;
          const int di = iys - a.iy_top[s0 + t], dj = ixs - a.ix_top[s0 + t];
          const int ky0 = max(0, -di), ky1 = min(T, T - di);
          const int kx0 = max(0, -dj), kx1 = min(T, T - dj);
          const int ow = kx1 - kx0, oh = ky1 - ky0;
          float acc = 0.0f;
          if (ow > 0 && oh > 0) {
            // four pixels per lane per step, their neighbour-stamp loads
            // issued together
            const int npx = ow * oh;
            for (int e0 = lane; e0 < npx; e0 += 256) {
              float tv[4], wv[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int e = e0 + 64 * u;
                tv[u] = 0.0f;
                wv[u] = 0.0f;
                if (e < npx) {
                  const int ky = ky0 + e / ow, kx = kx0 + e % ow;
                  tv[u] = tt[(ky + di) * T + kx + dj];
                  wv[u] = wsb[ky * T + kx];
                }
              }
#pragma unroll
              for (int u = 0; u < 4; ++u) acc = fmaf(wv[u], tv[u], acc);
            }
          }
          acc = wave_sum(acc);
          if (lane == 0) A[(size_t)s * NT + t] += double(acc);
        }
      }
      __syncthreads();
    }
  } else {
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
  if (a.lu_apart) return;  // band_lu_kernel factors and solves
  __syncthreads();
  block_lu_wpiv(A, NT, piv, &s_piv);
  // inverse columns (one per thread, column i stored as X[j * NT + i]) ->
  // variance diagonal; the solve on a thread without a column
  for (int i = tid; i < NT; i += UPH_BLOCK) {
    for (int j = 0; j < NT; ++j) X[(size_t)j * NT + i] = 0.0;
    X[(size_t)i * NT + i] = 1.0;
    lu_solve_strided(A, piv, NT, X + i, NT);
  }
  if (tid == (NT < UPH_BLOCK ? NT : 0)) lu_solve_one(A, piv, NT, rhs);
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

// ---------------------------------------------------------------------------
// Weighted stamp sums into disjoint image windows: the field-wide STAR
// subtraction at S1 flux (island_stage.py:2040-2117) and an island's member
// add-back (island_solver.py:1795-1818).  Window q = (x0, y0, h, w) of band
// band0 + blockIdx.y of out (bands, H, W) gets scale times the sum over its
// entries k in [off[q], off[q+1]) of flux[s * nbc + bl] * stamp s of band
// chunk bl (T x T at top-left (ix[s], iy[s]) in out coordinates, zero outside),
// s = list[k].  Entries are summed in list order per pixel, so the result does
// not depend on the launch shape; windows of one launch must be disjoint.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(UPH_BLOCK) void stamp_sum_kernel(
    int T, int n_st, int band0, const int64_t* __restrict__ off, const int32_t* __restrict__ win,
    const int32_t* __restrict__ list, const int32_t* __restrict__ ix,
    const int32_t* __restrict__ iy, const float* __restrict__ stamps,
    const float* __restrict__ flux, float scale, int H, int W, float* __restrict__ out) {
  const int q = blockIdx.x, bl = blockIdx.y, nbc = gridDim.y;
  const int x0 = win[q * 4], y0 = win[q * 4 + 1], h = win[q * 4 + 2], w = win[q * 4 + 3];
  const int64_t k0 = off[q], k1 = off[q + 1];
  const float* st = stamps + (size_t)bl * n_st * T * T;
  float* img = out + (size_t)(band0 + bl) * H * W;
  for (int e = blockIdx.z * UPH_BLOCK + threadIdx.x; e < h * w; e += gridDim.z * UPH_BLOCK) {
    const int py = y0 + e / w, px = x0 + e % w;
    float v = 0.0f;
    for (int64_t k = k0; k < k1; ++k) {
      const int s = list[k];
      const int u = px - ix[s], t = py - iy[s];
      if (u < 0 || u >= T || t < 0 || t >= T) continue;
      v = fmaf(flux[(size_t)s * nbc + bl], st[((size_t)s * T + t) * T + u], v);
    }
    float* o = img + (size_t)py * W + px;
    *o = fmaf(scale, v, *o);
  }
}

// Per-grid background planes subtracted from the padded data stack
// (island_stage.py:1731-1751): grid q with core window win[q] = (x0, y0, h,
// w) in image coordinates and coefficients bg[(q * nb + band) * M + m] gets
// out[band][pad + y0 + py][pad + x0 + px] -= sum_m bg_m * basis_m(px, py),
// the basis X^i Y^j of the linear solve (float64 normalisation rounded to
// float32).  Core windows are disjoint.
template <int M>
__global__ __launch_bounds__(UPH_BLOCK) void bg_planes_kernel(
    const int32_t* __restrict__ win, const float* __restrict__ bg, int nb, int pad, double xo,
    double xs, double yo, double ys, int Hp, int Wp, float* __restrict__ out) {
  const int q = blockIdx.x, band = blockIdx.y;
  const int x0 = win[q * 4], y0 = win[q * 4 + 1], h = win[q * 4 + 2], w = win[q * 4 + 3];
  const float* c = bg + ((size_t)q * nb + band) * M;
  float cf[M];
#pragma unroll
  for (int m = 0; m < M; ++m) cf[m] = c[m];
  constexpr int m_[6] = {0, 1, 0, 2, 1, 0};
  constexpr int n_[6] = {0, 0, 1, 0, 1, 2};
  float* img = out + (size_t)band * Hp * Wp;
  for (int e = blockIdx.z * UPH_BLOCK + threadIdx.x; e < h * w; e += gridDim.z * UPH_BLOCK) {
    const int py = y0 + e / w, px = x0 + e % w;
    const double X = (double(px) - xo) / xs, Y = (double(py) - yo) / ys;
    float v = 0.0f;
#pragma unroll
    for (int m = 0; m < M; ++m) {
      double t = 1.0;
      for (int i = 0; i < m_[m]; ++i) t *= X;
      for (int i = 0; i < n_[m]; ++i) t *= Y;
      v = fmaf(float(t), cf[m], v);
    }
    img[(size_t)(pad + py) * Wp + pad + px] -= v;
  }
}

// Island window gather (island_solver.py:1820-1830): window q = (sx, sy, dx,
// dy, side) copies the side x side block at (sx, sy) of every band of src_a
// (and src_b) (bands, Hs, Ws) to (dx, dy) of dst_a (dst_b) (bands, Hd, Wd).
__global__ __launch_bounds__(UPH_BLOCK) void gather_windows_kernel(
    const int32_t* __restrict__ win, int Hs, int Ws, const float* __restrict__ src_a,
    const float* __restrict__ src_b, int Hd, int Wd, float* __restrict__ dst_a,
    float* __restrict__ dst_b) {
  const int q = blockIdx.x, band = blockIdx.y;
  const int sx = win[q * 5], sy = win[q * 5 + 1], dx = win[q * 5 + 2], dy = win[q * 5 + 3];
  const int side = win[q * 5 + 4];
  for (int e = blockIdx.z * UPH_BLOCK + threadIdx.x; e < side * side;
       e += gridDim.z * UPH_BLOCK) {
    const int r = e / side, c = e % side;
    const size_t i = ((size_t)band * Hs + sy + r) * Ws + sx + c;
    const size_t o = ((size_t)band * Hd + dy + r) * Wd + dx + c;
    dst_a[o] = src_a[i];
    if (dst_b) dst_b[o] = src_b[i];
  }
}

// ---------------------------------------------------------------------------
// Blank-response (empty-aperture) median of a grid-band problem
// (solver_grid.py:2867-2943): with the final fluxes and background, the
// weighted residual wr = w (d - wing - model) and w on the window (good
// pixels, w = invvar), then resp = corr(wr, phi) / corr(w, phi^2) ('same'
// correlation with the problem's PSF template phi, T x T, clipped >= 0 and
// unit-sum), NaN where the denominator is 0 or the pixel is not good, and
// b = nanmedian(resp).
// ---------------------------------------------------------------------------
struct BlankArgs {
  int n_prob, T, M, H, W;
  const int64_t* src_off;
  const int32_t* win;
  const float* data;
  const float* invvar;
  const uint8_t* good;
  const float* wing;        // (n_bands, H, W) or null
  const uint8_t* use_wing;  // (n_prob)
  const float* stamps;      // (total_src, T, T) unit stamps
  const int32_t* ix_top;
  const int32_t* iy_top;
  const float* flux;        // (total_src)
  const float* bg;          // (n_prob, M)
  double bxo, bxs, byo, bys;
  const float* phi;         // (n_prob, T, T)
  float* work;              // per problem: wr, w, resp planes of max_pix floats
  size_t max_pix;
  double* b_hat;            // (n_prob)
};

// Factor and solve of the grid-band systems assembled by band_solve_kernel
// (problem q's A (NT x NT, ridge applied) and right-hand side in its
// scratch): A copied into LDS when it fits (NT <= kLuLdsN; larger systems
// keep their trailing kLuLdsN block there), then the same
// LU with partial pivoting, the inverse columns (one per thread) for the
// variance diagonal and the solve, as band_solve_kernel does in place.
constexpr int kLuLdsN = 137;

// block_lu_wpiv for n > kLuLdsN: steps k < k0 = n - kLuLdsN on the matrix
// in global memory, then the trailing block [k0, n) x [k0, n) moves to L
// (LDS, row stride m = n - k0) for the remaining steps, and back at the end.
// The pivots, the operations and their order are those of block_lu_wpiv.
__device__ void block_lu_lds_tail(double* A, double* L, int n, int* piv, int* s_p) {
  const int tid = threadIdx.x, m = kLuLdsN, k0 = n - m;
  for (int k = 0; k < n; ++k) {
    if (k == k0) {
      for (int e = tid; e < m * m; e += UPH_BLOCK)
        L[e] = A[(size_t)(k0 + e / m) * n + k0 + e % m];
      __syncthreads();
    }
    const bool in_l = k >= k0;
    // element (i, j) with i, j >= k: the trailing block once in LDS
    auto at = [&](int i, int j) -> double& {
      return in_l ? L[(i - k0) * m + (j - k0)] : A[(size_t)i * n + j];
    };
    if (tid < 64) {
      double best = tid == 0 ? fabs(at(k, k)) : -1.0;
      int p = tid == 0 ? k : n;
      for (int i = k + tid + (tid == 0 ? 64 : 0); i < n; i += 64) {
        const double v = fabs(at(i, k));
        if (v > best) {
          best = v;
          p = i;
        }
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_down(best, off, 64);
        const int op = __shfl_down(p, off, 64);
        if (ob > best || (ob == best && op < p)) {
          best = ob;
          p = op;
        }
      }
      if (tid == 0) {
        piv[k] = p;
        *s_p = p;
      }
    }
    __syncthreads();
    const int p = *s_p;
    if (p != k) {
      for (int j = tid; j < n; j += UPH_BLOCK) {
        double* x = (in_l && j >= k0) ? &L[(k - k0) * m + j - k0] : &A[(size_t)k * n + j];
        double* y = (in_l && j >= k0) ? &L[(p - k0) * m + j - k0] : &A[(size_t)p * n + j];
        const double t = *x;
        *x = *y;
        *y = t;
      }
      __syncthreads();
    }
    const double akk = at(k, k);
    for (int i = k + 1 + tid; i < n; i += UPH_BLOCK) at(i, k) /= akk;
    __syncthreads();
    const int mm = n - k - 1;
    for (int e = tid; e < mm * mm; e += UPH_BLOCK) {
      const int i = k + 1 + e / mm, j = k + 1 + e % mm;
      at(i, j) -= at(i, k) * at(k, j);
    }
    __syncthreads();
  }
  for (int e = tid; e < m * m; e += UPH_BLOCK) A[(size_t)(k0 + e / m) * n + k0 + e % m] = L[e];
  __syncthreads();
}

template <int M>
__global__ __launch_bounds__(UPH_BLOCK) void band_lu_kernel(BandArgs a) {
  extern __shared__ double lu_lds[];
  __shared__ int s_piv;
  const int q = blockIdx.x, tid = threadIdx.x;
  const int64_t s0 = a.src_off[q];
  const int n = int(a.src_off[q + 1] - s0);
  const int NT = n + M;
  char* base = a.scratch + (size_t)q * a.per_problem;
  if ((size_t)NT * NT * 16 + (size_t)NT * 12 > a.per_problem) return;  // NaN written already
  double* Ag = reinterpret_cast<double*>(base);
  double* X = Ag + (size_t)NT * NT;
  double* rhs = X + (size_t)NT * NT;
  int* piv = reinterpret_cast<int*>(rhs + NT);
  auto finish = [&](double* A) {
    // column i of the inverse as X[j * NT + i] (one column per thread); the
    // solve on a thread without a column
    for (int i = tid; i < NT; i += UPH_BLOCK) {
      for (int j = 0; j < NT; ++j) X[(size_t)j * NT + i] = 0.0;
      X[(size_t)i * NT + i] = 1.0;
      lu_solve_strided(A, piv, NT, X + i, NT);
    }
    if (tid == (NT < UPH_BLOCK ? NT : 0)) lu_solve_one(A, piv, NT, rhs);
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
  };
  if (NT <= kLuLdsN) {
    for (int e = tid; e < NT * NT; e += UPH_BLOCK) lu_lds[e] = Ag[e];
    __syncthreads();
    block_lu_wpiv(lu_lds, NT, piv, &s_piv);
    finish(lu_lds);
  } else {
    block_lu_lds_tail(Ag, lu_lds, NT, piv, &s_piv);
    finish(Ag);
  }
}

template <int M>
__global__ __launch_bounds__(UPH_BLOCK) void blank_residual_kernel(BlankArgs a) {
  const int q = blockIdx.x;
  const int band = a.win[q * 5], x0 = a.win[q * 5 + 1], y0 = a.win[q * 5 + 2];
  const int h = a.win[q * 5 + 3], w = a.win[q * 5 + 4], T = a.T, TT = T * T;
  const int64_t s0 = a.src_off[q], s1 = a.src_off[q + 1];
  const size_t plane = (size_t)a.H * a.W, off = (size_t)band * plane;
  float* WR = a.work + (size_t)q * 3 * a.max_pix;
  float* WW = WR + a.max_pix;
  const float* bgq = a.bg + (size_t)q * M;
  const bool wing = a.wing != nullptr && a.use_wing[q];
  for (int e = blockIdx.y * UPH_BLOCK + threadIdx.x; e < h * w; e += gridDim.y * UPH_BLOCK) {
    const int py = e / w, px = e % w;
    const size_t gi = off + (size_t)(y0 + py) * a.W + x0 + px;
    const float iv = a.invvar[gi];
    const bool gd = a.good[gi] && iv > 0.0f;
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
    const float wv = gd ? iv : 0.0f;
    WR[e] = gd ? wv * (d - model) : 0.0f;
    WW[e] = wv;
  }
}

// Output tile of the correlation launch: kCorrRows x kCorrCols pixels, one
// row of 4 adjacent outputs per thread (kCorrBlock threads).
constexpr int kCorrRows = 32, kCorrCols = 64, kCorrBlock = 512;

__host__ __device__ inline int corr_tp(int T) { return (T + 3) / 4 * 4; }
__host__ __device__ inline int corr_sx(int T) { return kCorrCols + corr_tp(T) + 4; }
__host__ __device__ inline size_t corr_lds_bytes(int T) {
  return ((size_t)2 * T * corr_tp(T) + (size_t)2 * (kCorrRows + T - 1) * corr_sx(T)) *
         sizeof(float);
}

// 'same' correlations of wr with phi and of w with phi^2 over one output
// tile per workgroup visit, inputs (with halo) and template staged in LDS;
// each thread keeps 4 adjacent outputs of both correlations and reads the
// template and the inputs 4 floats at a time, so every input value loaded
// serves 4 outputs.  Template rows are zero-padded to a multiple of 4.
__global__ __launch_bounds__(kCorrBlock) void blank_corr_kernel(BlankArgs a) {
  extern __shared__ float4 lds4[];
  float* lds = reinterpret_cast<float*>(lds4);
  const int q = blockIdx.x;
  const int band = a.win[q * 5], x0 = a.win[q * 5 + 1], y0 = a.win[q * 5 + 2];
  const int h = a.win[q * 5 + 3], w = a.win[q * 5 + 4], T = a.T, c = T / 2;
  const int TP = corr_tp(T), SX = corr_sx(T), SY = kCorrRows + T - 1;
  const int tiles_x = (w + kCorrCols - 1) / kCorrCols, tiles_y = (h + kCorrRows - 1) / kCorrRows;
  float* P1 = lds;                // phi (T rows of TP)
  float* P2 = P1 + T * TP;        // phi^2
  float* A1 = P2 + T * TP;        // wr tile with halo (SY rows of SX)
  float* A2 = A1 + SY * SX;       // w tile with halo
  const float* WR = a.work + (size_t)q * 3 * a.max_pix;
  const float* WW = WR + a.max_pix;
  float* RESP = a.work + (size_t)q * 3 * a.max_pix + 2 * a.max_pix;
  const float* ph = a.phi + (size_t)q * T * T;
  for (int e = threadIdx.x; e < T * TP; e += kCorrBlock) {
    const int ky = e / TP, kx = e % TP;
    const float v = kx < T ? ph[ky * T + kx] : 0.0f;
    P1[e] = v;
    P2[e] = v * v;
  }
  const size_t plane = (size_t)a.H * a.W, off = (size_t)band * plane;
  const int ly = threadIdx.x / (kCorrCols / 4), lx = 4 * (threadIdx.x % (kCorrCols / 4));
  for (int tile = blockIdx.y; tile < tiles_x * tiles_y; tile += gridDim.y) {
    const int ty0 = (tile / tiles_x) * kCorrRows, tx0 = (tile % tiles_x) * kCorrCols;
    __syncthreads();
    for (int e = threadIdx.x; e < SY * SX; e += kCorrBlock) {
      const int yy = ty0 - c + e / SX, xx = tx0 - c + e % SX;
      const bool in = yy >= 0 && yy < h && xx >= 0 && xx < w;
      A1[e] = in ? WR[yy * w + xx] : 0.0f;
      A2[e] = in ? WW[yy * w + xx] : 0.0f;
    }
    __syncthreads();
    float num[4] = {0.0f, 0.0f, 0.0f, 0.0f}, den[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int ky = 0; ky < T; ++ky) {
      const float4* r1 = reinterpret_cast<const float4*>(A1 + (ly + ky) * SX + lx);
      const float4* r2 = reinterpret_cast<const float4*>(A2 + (ly + ky) * SX + lx);
      const float4* k1 = reinterpret_cast<const float4*>(P1 + ky * TP);
      const float4* k2 = reinterpret_cast<const float4*>(P2 + ky * TP);
      float4 u1 = r1[0], u2 = r2[0];
      for (int k4 = 0; k4 < TP / 4; ++k4) {
        const float4 v1 = r1[k4 + 1], v2 = r2[k4 + 1];
        const float4 t1 = k1[k4], t2 = k2[k4];
        const float x1[8] = {u1.x, u1.y, u1.z, u1.w, v1.x, v1.y, v1.z, v1.w};
        const float x2[8] = {u2.x, u2.y, u2.z, u2.w, v2.x, v2.y, v2.z, v2.w};
        const float c1[4] = {t1.x, t1.y, t1.z, t1.w};
        const float c2[4] = {t2.x, t2.y, t2.z, t2.w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            num[j] = fmaf(x1[j + t], c1[t], num[j]);
            den[j] = fmaf(x2[j + t], c2[t], den[j]);
          }
        }
        u1 = v1;
        u2 = v2;
      }
    }
    const int oy = ty0 + ly;
    const float nanv = __int_as_float(0x7fc00000);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int ox = tx0 + lx + j;
      if (oy >= h || ox >= w) continue;
      const size_t gi = off + (size_t)(y0 + oy) * a.W + x0 + ox;
      const bool gd = a.good[gi] && a.invvar[gi] > 0.0f;
      RESP[oy * w + ox] = (gd && den[j] > 0.0f) ? num[j] / fmaxf(den[j], 1e-30f) : nanv;
    }
  }
}

// Total order of floats as unsigned keys (negative values reversed).
__device__ __forceinline__ unsigned float_key(float v) {
  const unsigned u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float key_float(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// The rank-th smallest (0-based) of the non-NaN values of v[0, n): radix
// select, four 8-bit digits from the top, one LDS histogram per digit.
__device__ float select_rank(const float* v, int n, int rank, unsigned* hist, unsigned* s_pref,
                             int* s_rank) {
  const int tid = threadIdx.x;
  unsigned prefix = 0, mask = 0;
  if (tid == 0) *s_rank = rank;
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (int i = tid; i < 256; i += UPH_BLOCK) hist[i] = 0;
    __syncthreads();
    for (int i = tid; i < n; i += UPH_BLOCK) {
      const float x = v[i];
      if (x != x) continue;
      const unsigned k = float_key(x);
      if ((k & mask) == prefix) atomicAdd(&hist[(k >> shift) & 255u], 1u);
    }
    __syncthreads();
    if (tid == 0) {
      int r = *s_rank;
      unsigned d = 0;
      while (d < 255 && r >= int(hist[d])) {
        r -= int(hist[d]);
        ++d;
      }
      *s_rank = r;
      *s_pref = prefix | (d << shift);
    }
    __syncthreads();
    prefix = *s_pref;
    mask |= 255u << shift;
  }
  return key_float(prefix);
}

__global__ __launch_bounds__(UPH_BLOCK) void blank_median_kernel(BlankArgs a) {
  __shared__ unsigned hist[256];
  __shared__ unsigned s_pref;
  __shared__ int s_rank;
  __shared__ int s_cnt;
  const int q = blockIdx.x;
  const int h = a.win[q * 5 + 3], w = a.win[q * 5 + 4], n = h * w;
  const float* RESP = a.work + (size_t)q * 3 * a.max_pix + 2 * a.max_pix;
  if (threadIdx.x == 0) s_cnt = 0;
  __syncthreads();
  int cnt = 0;
  for (int i = threadIdx.x; i < n; i += UPH_BLOCK) cnt += RESP[i] == RESP[i];
  atomicAdd(&s_cnt, cnt);
  __syncthreads();
  const int m = s_cnt;
  if (m == 0) {
    if (threadIdx.x == 0) a.b_hat[q] = __longlong_as_double(0x7ff8000000000000ll);
    return;
  }
  const float lo = select_rank(RESP, n, (m - 1) / 2, hist, &s_pref, &s_rank);
  __syncthreads();
  const float hi = select_rank(RESP, n, m / 2, hist, &s_pref, &s_rank);
  if (threadIdx.x == 0) a.b_hat[q] = 0.5 * (double(lo) + double(hi));
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
  // extended sources through the planes pipeline where it serves T
  int n_cu = 0;
  st0 = uph::hip_status(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev),
                        "uph_render_stamps attribute");
  if (st0 != UPH_OK) return st0;
  const bool gal_done = ext && uph::launch_galaxy_stamps(
                                        n, T, n_terms, psf_coefs, psf_geo, sersic, src_band, x,
                                        y, gx0, gy0, stype, logre, ee1, ee2, sersic_n, out,
                                        static_cast<char*>(scratch), scratch_bytes, n_cu, s);
  st0 = uph::hip_status(hipGetLastError(), "uph_render_stamps galaxy launch");
  if (st0 != UPH_OK || (gal_done && stars_lds)) return st0;
  // the workspace kernel: the extended sources the pipeline did not take,
  // and the stars too large for the LDS kernel
  const size_t per = uph::galaxy_scratch_bytes(T, sizeof(float));
  const size_t nb_s = scratch_bytes / per;
  UPH_CHECK_ARG(nb_s >= 1, "uph_render_stamps: scratch of %zu bytes < one workgroup (%zu)",
                scratch_bytes, per);
  const int nblk = (int)(nb_s < (size_t)n ? nb_s : (size_t)n);
  hipLaunchKernelGGL(uph::render_galaxy_stamps_kernel, dim3(nblk), dim3(UPH_BLOCK), 0, s, n, T,
                     n_terms, psf_coefs, psf_geo, sersic, src_band, x, y, gx0, gy0, stype, logre,
                     ee1, ee2, sersic_n, stars_lds ? 0 : 1, gal_done ? 1 : 0, renorm, out,
                     static_cast<char*>(scratch), per);
  return uph::hip_status(hipGetLastError(), "uph_render_stamps workspace launch");
}

extern "C" size_t uph_band_scratch_bytes(int n_prob, int max_src, int bg_order) {
  const int M = n_terms_of_order(bg_order);
  return (size_t)n_prob * uph::band_scratch_layout(max_src, M).per_problem;
}

static int band_solve_impl(int n_prob, int T, int bg_order, int k_max, const int64_t* src_off,
                           const int32_t* win, int H, int W, int img_x0, int img_y0,
                           const float* data, const float* invvar, const uint8_t* good,
                           double bg_xo, double bg_xs, double bg_yo, double bg_ys,
                           const float* stamps, const int32_t* ix_top, const int32_t* iy_top,
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
  // stage the weighted stamps in LDS up to T = 63 (64 KiB per workgroup)
  a.stage = T <= 63;
  a.lu_apart = 1;
  const size_t lds = a.stage ? (size_t)UPH_WAVES * T * T * sizeof(float) : 0;
  const size_t lds_lu = (size_t)uph::kLuLdsN * uph::kLuLdsN * sizeof(double);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (a.M == 1) {
    hipLaunchKernelGGL(uph::band_solve_kernel<1>, dim3(n_prob), dim3(UPH_BLOCK), lds, s, a);
    hipLaunchKernelGGL(uph::band_lu_kernel<1>, dim3(n_prob), dim3(UPH_BLOCK), lds_lu, s, a);
  } else if (a.M == 3) {
    hipLaunchKernelGGL(uph::band_solve_kernel<3>, dim3(n_prob), dim3(UPH_BLOCK), lds, s, a);
    hipLaunchKernelGGL(uph::band_lu_kernel<3>, dim3(n_prob), dim3(UPH_BLOCK), lds_lu, s, a);
  } else {
    hipLaunchKernelGGL(uph::band_solve_kernel<6>, dim3(n_prob), dim3(UPH_BLOCK), lds, s, a);
    hipLaunchKernelGGL(uph::band_lu_kernel<6>, dim3(n_prob), dim3(UPH_BLOCK), lds_lu, s, a);
  }
  return uph::hip_status(hipGetLastError(), "uph_band_solve launch");
}

extern "C" int uph_band_solve(int n_prob, int T, int bg_order, int k_max, const int64_t* src_off,
                              const int32_t* win, int H, int W, int img_x0, int img_y0,
                              const float* data, const float* invvar, const uint8_t* good,
                              double bg_xo, double bg_xs, double bg_yo, double bg_ys,
                              const float* stamps, const int32_t* ix_top, const int32_t* iy_top,
                              const int32_t* nbr_idx, float reg, float* fluxes, float* flux_var,
                              float* bg, float* bg_var, void* scratch, size_t scratch_bytes,
                              void* stream) {
  return band_solve_impl(n_prob, T, bg_order, k_max, src_off, win, H, W, img_x0, img_y0, data,
                         invvar, good, bg_xo, bg_xs, bg_yo, bg_ys, stamps, ix_top, iy_top,
                         nbr_idx, reg, fluxes, flux_var, bg, bg_var, scratch, scratch_bytes,
                         stream);
}

// ---------------------------------------------------------------------------
// Allocation-free band solve, for callers that cannot allocate (code under
// graph capture, serving loops): the launches of uph_band_solve with their
// workspace in this file-scope device array of UPH_BAND_NOALLOC_BYTES.  Each
// device holds its own copy from module load until the process ends.  In
// host code the array's name is not a device address: the address is taken
// once per device with hipGetSymbolAddress and cached.  The kernels write
// every workspace value they read (the normal matrix is zeroed by the
// assembly launch), so nothing needs clearing between calls; calls that
// overlap on one device share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) char scratch[UPH_BAND_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return uph::hip_status(e, "uph_band_solve_noalloc: hipGetDevice");
  UPH_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "uph_band_solve_noalloc: device %d", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return uph::hip_status(e, "uph_band_solve_noalloc: hipGetSymbolAddress");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return UPH_OK;
}
}  // namespace

extern "C" int uph_band_solve_noalloc(int n_prob, int T, int bg_order, int k_max, int max_src,
                                      const int64_t* src_off, const int32_t* win, int H, int W,
                                      int img_x0, int img_y0, const float* data,
                                      const float* invvar, const uint8_t* good, double bg_xo,
                                      double bg_xs, double bg_yo, double bg_ys,
                                      const float* stamps, const int32_t* ix_top,
                                      const int32_t* iy_top, const int32_t* nbr_idx, float reg,
                                      float* fluxes, float* flux_var, float* bg, float* bg_var,
                                      void* stream) {
  UPH_CHECK_ARG(n_prob >= 0 && max_src >= 0, "uph_band_solve_noalloc: n_prob=%d max_src=%d",
                n_prob, max_src);
  UPH_CHECK_ARG(bg_order >= 0 && bg_order <= 2,
                "uph_band_solve_noalloc: bg_order=%d outside 0..2", bg_order);
  if (n_prob == 0) return UPH_OK;
  const size_t need = uph_band_scratch_bytes(n_prob, max_src, bg_order);
  UPH_CHECK_ARG(need <= (size_t)UPH_BAND_NOALLOC_BYTES,
                "uph_band_solve_noalloc: %d problems of up to %d sources need %zu bytes of "
                "workspace, more than the static %zu; split the batch",
                n_prob, max_src, need, (size_t)UPH_BAND_NOALLOC_BYTES);
  void* ws = nullptr;
  const int st = noalloc_scratch(&ws);
  if (st != UPH_OK) return st;
  return band_solve_impl(n_prob, T, bg_order, k_max, src_off, win, H, W, img_x0, img_y0, data,
                         invvar, good, bg_xo, bg_xs, bg_yo, bg_ys, stamps, ix_top, iy_top,
                         nbr_idx, reg, fluxes, flux_var, bg, bg_var, ws, need, stream);
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

extern "C" int uph_stamp_sum(int n_win, int max_pix, int T, int n_st, int nbc, int band0,
                             int n_bands, const int64_t* off, const int32_t* win,
                             const int32_t* list, const int32_t* ix, const int32_t* iy,
                             const float* stamps, const float* flux, float scale, int H, int W,
                             float* out, void* stream) {
  UPH_CHECK_ARG(n_win >= 0 && max_pix >= 0 && T >= 1 && n_st >= 0 && nbc >= 1 && band0 >= 0 &&
                    band0 + nbc <= n_bands && H >= 1 && W >= 1,
                "uph_stamp_sum: n_win=%d max_pix=%d T=%d n_st=%d bands [%d, %d) of %d, %d x %d",
                n_win, max_pix, T, n_st, band0, band0 + nbc, n_bands, H, W);
  if (n_win == 0 || max_pix == 0) return UPH_OK;
  UPH_CHECK_ARG(off && win && list && ix && iy && stamps && flux && out,
                "uph_stamp_sum: null argument");
  const int per = UPH_BLOCK * 16;
  const int z = std::min(64, (max_pix + per - 1) / per);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(uph::stamp_sum_kernel, dim3(n_win, nbc, z), dim3(UPH_BLOCK), 0, s, T, n_st,
                     band0, off, win, list, ix, iy, stamps, flux, scale, H, W, out);
  return uph::hip_status(hipGetLastError(), "uph_stamp_sum launch");
}

extern "C" int uph_subtract_bg_planes(int n_grid, int max_pix, int n_bands, int bg_order,
                                      const int32_t* win, const float* bg, int pad, double bg_xo,
                                      double bg_xs, double bg_yo, double bg_ys, int Hp, int Wp,
                                      float* out, void* stream) {
  UPH_CHECK_ARG(n_grid >= 0 && max_pix >= 0 && n_bands >= 1 && bg_order >= 0 && bg_order <= 2 &&
                    pad >= 0 && Hp >= 1 && Wp >= 1 && bg_xs != 0.0 && bg_ys != 0.0,
                "uph_subtract_bg_planes: n_grid=%d n_bands=%d bg_order=%d pad=%d %d x %d",
                n_grid, n_bands, bg_order, pad, Hp, Wp);
  if (n_grid == 0 || max_pix == 0) return UPH_OK;
  UPH_CHECK_ARG(win && bg && out, "uph_subtract_bg_planes: null argument");
  const int per = UPH_BLOCK * 16;
  const int z = std::min(64, (max_pix + per - 1) / per);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid(n_grid, n_bands, z);
  if (bg_order == 0)
    hipLaunchKernelGGL(uph::bg_planes_kernel<1>, grid, dim3(UPH_BLOCK), 0, s, win, bg, n_bands,
                       pad, bg_xo, bg_xs, bg_yo, bg_ys, Hp, Wp, out);
  else if (bg_order == 1)
    hipLaunchKernelGGL(uph::bg_planes_kernel<3>, grid, dim3(UPH_BLOCK), 0, s, win, bg, n_bands,
                       pad, bg_xo, bg_xs, bg_yo, bg_ys, Hp, Wp, out);
  else
    hipLaunchKernelGGL(uph::bg_planes_kernel<6>, grid, dim3(UPH_BLOCK), 0, s, win, bg, n_bands,
                       pad, bg_xo, bg_xs, bg_yo, bg_ys, Hp, Wp, out);
  return uph::hip_status(hipGetLastError(), "uph_subtract_bg_planes launch");
}

extern "C" int uph_gather_windows(int n_win, int max_side, int n_bands, const int32_t* win,
                                  int Hs, int Ws, const float* src_a, const float* src_b,
                                  int Hd, int Wd, float* dst_a, float* dst_b, void* stream) {
  UPH_CHECK_ARG(n_win >= 0 && max_side >= 0 && n_bands >= 1 && Hs >= 1 && Ws >= 1 && Hd >= 1 &&
                    Wd >= 1,
                "uph_gather_windows: n_win=%d max_side=%d n_bands=%d src %d x %d dst %d x %d",
                n_win, max_side, n_bands, Hs, Ws, Hd, Wd);
  if (n_win == 0 || max_side == 0) return UPH_OK;
  UPH_CHECK_ARG(win && src_a && dst_a && ((src_b == nullptr) == (dst_b == nullptr)),
                "uph_gather_windows: null argument");
  const int per = UPH_BLOCK * 16;
  const int z = std::min(64, (max_side * max_side + per - 1) / per);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(uph::gather_windows_kernel, dim3(n_win, n_bands, z), dim3(UPH_BLOCK), 0, s,
                     win, Hs, Ws, src_a, src_b, Hd, Wd, dst_a, dst_b);
  return uph::hip_status(hipGetLastError(), "uph_gather_windows launch");
}

extern "C" size_t uph_blank_work_floats(int n_prob, int max_pix) {
  return (size_t)n_prob * 3 * (size_t)(max_pix > 0 ? max_pix : 0);
}

extern "C" int uph_blank_response(int n_prob, int T, int bg_order, const int64_t* src_off,
                                  const int32_t* win, int H, int W, const float* data,
                                  const float* invvar, const uint8_t* good, const float* wing,
                                  const uint8_t* use_wing, const float* stamps,
                                  const int32_t* ix_top, const int32_t* iy_top,
                                  const float* flux, const float* bg, double bg_xo, double bg_xs,
                                  double bg_yo, double bg_ys, const float* phi, float* work,
                                  size_t work_floats, int max_pix, double* b_hat, void* stream) {
  UPH_CHECK_ARG(n_prob >= 0 && T >= 1 && (T % 2) == 1 && bg_order >= 0 && bg_order <= 2,
                "uph_blank_response: n_prob=%d T=%d bg_order=%d", n_prob, T, bg_order);
  if (n_prob == 0) return UPH_OK;
  UPH_CHECK_ARG(src_off && win && data && invvar && good && stamps && ix_top && iy_top && flux &&
                    bg && phi && work && b_hat && (wing == nullptr || use_wing),
                "uph_blank_response: null argument");
  UPH_CHECK_ARG(work_floats >= uph_blank_work_floats(n_prob, max_pix),
                "uph_blank_response: work of %zu floats < %zu", work_floats,
                uph_blank_work_floats(n_prob, max_pix));
  uph::BlankArgs a{};
  a.n_prob = n_prob;
  a.T = T;
  a.M = (bg_order + 1) * (bg_order + 2) / 2;
  a.H = H;
  a.W = W;
  a.src_off = src_off;
  a.win = win;
  a.data = data;
  a.invvar = invvar;
  a.good = good;
  a.wing = wing;
  a.use_wing = use_wing;
  a.stamps = stamps;
  a.ix_top = ix_top;
  a.iy_top = iy_top;
  a.flux = flux;
  a.bg = bg;
  a.bxo = bg_xo;
  a.bxs = bg_xs;
  a.byo = bg_yo;
  a.bys = bg_ys;
  a.phi = phi;
  a.work = work;
  a.max_pix = (size_t)max_pix;
  a.b_hat = b_hat;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (a.M == 1)
    hipLaunchKernelGGL(uph::blank_residual_kernel<1>, dim3(n_prob, 16), dim3(UPH_BLOCK), 0, s, a);
  else if (a.M == 3)
    hipLaunchKernelGGL(uph::blank_residual_kernel<3>, dim3(n_prob, 16), dim3(UPH_BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL(uph::blank_residual_kernel<6>, dim3(n_prob, 16), dim3(UPH_BLOCK), 0, s, a);
  const size_t lds = uph::corr_lds_bytes(T);
  UPH_CHECK_ARG(lds <= 160 * 1024, "uph_blank_response: template T=%d too large", T);
  hipLaunchKernelGGL(uph::blank_corr_kernel, dim3(n_prob, 16), dim3(uph::kCorrBlock), lds, s, a);
  hipLaunchKernelGGL(uph::blank_median_kernel, dim3(n_prob), dim3(UPH_BLOCK), 0, s, a);
  return uph::hip_status(hipGetLastError(), "uph_blank_response launch");
}
