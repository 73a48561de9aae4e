// Persistent per-island VarPro Levenberg-Marquardt kernel (gfx950).
//
// One 256-thread workgroup owns one island at a time and pulls the next
// island from a device work queue, so islands of very different cost share
// one launch and the LM loop never returns to the host.  Per band the
// workgroup gathers the island window, renders every member's unit-flux
// model and its live tangent planes into per-workgroup scratch (L2
// resident), reduces the weighted Gram matrix of [data, models, tangents],
// solves the member fluxes in float64, forms the residual, and accumulates
// the flux-marginalised Gauss-Newton (H, g) from Gram identities:
//   D_j = A^-1 (e_k s_j - f_k U_j)
//   H_ij += D_i.A D_j + f_kj D_i.U_j + f_ki D_j.U_i + f_ki f_kj G_ij
//   g_j  += -(D_j.c + f_kj s_j)
// with A = M^T W M, U = M^T W T, G = T^T W T, c = M^T W R, s = T^T W R.
// Algorithm followed: tract7dt/fit/grid_gpu/island_solver.py models_band
// (:480-562), _fluxes_band/residual_band/chi2_of (:570-684), hg_scan
// (:812-865), one_step (:921-1002), final_fluxes (:1467-1482).
#include <hip/hip_runtime.h>

#include "uph_capi_internal.hpp"
#include "uph_device.hpp"
#include "uph_galaxy.hpp"
#include "uph_prerender.hpp"
#include "uph_island.hpp"

namespace uph {

template <typename Real>
struct IslandSh {
  float p[kMaxLive], ptry[kMaxLive], lo[kMaxLive], hi[kMaxLive];
  int mem[kMaxMembers];  // slot index of real member r
  int st[kMaxMembers];
  float nfz[kMaxMembers];
  uint8_t fn[kMaxMembers];
  int live_r[kMaxLive];  // real-member index of live parameter
  int live_j[kMaxLive];  // parameter 0..5
  int first_live[kMaxMembers + 1];
  Real taps[28];
  Real bas[63];          // PSF basis, d/dX, d/dY (21 each)
  GalParams<Real> gp;    // extended member being rendered
  Real fR[kMaxMembers];  // member fluxes rounded to Real
  double red[UPH_WAVES * 32];
  double redd[UPH_WAVES];
  int Nr, nL, item, pivot;
  int N, B, T, nLc;          // this island's class shape
  long long moff, foff;      // member-slot and flux-block offsets
  const float* coefs;        // PSF coefficients of the island's tier
  // LM state: chi2 and the damping in float64, so the accept test compares
  // float64 sums (the per-band chi2 are float64 reductions of the residual)
  double lam, chi2, chi2_try;
  int consec, nrej, nst, conv, stall;
};

struct Small {
  double *Araw, *LU, *b, *f, *c, *U, *Gt, *s, *D, *AD, *H, *g, *Hd, *delta, *colws;
  int *pivA, *pivH;
};

__device__ inline Small carve_small(char* base, int N, int nL) {
  double* d = reinterpret_cast<double*>(base);
  Small s;
  s.Araw = d; d += N * N;
  s.LU = d; d += N * N;
  s.b = d; d += N;
  s.f = d; d += N;
  s.c = d; d += N;
  s.U = d; d += N * nL;
  s.Gt = d; d += nL * nL;
  s.s = d; d += nL;
  s.D = d; d += nL * N;
  s.AD = d; d += nL * N;
  s.H = d; d += nL * nL;
  s.g = d; d += nL;
  s.Hd = d; d += nL * nL;
  s.delta = d; d += nL;
  s.colws = d; d += 256 * (nL > N ? nL : N);
  int* ip = reinterpret_cast<int*>(d + 8);
  s.pivA = ip; ip += N;
  s.pivH = ip;
  return s;
}

// ---------------------------------------------------------------------------
// Weighted pair reductions: out = sum_p W[p] * Z[a][p] * Z[b][p].
// GRAM mode: all a <= b over planes [0, C) (symmetric fill of out, C x C).
// VEC mode: pairs (q, zr) for q in [0, C) into out[q].
// Chunks of 32 pairs; per-thread split-K accumulation, fixed-order
// wave + block reduction (deterministic).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pair_of(int e, int C, int& a, int& b) {
  a = 0;
  while (e >= C - a) {
    e -= C - a;
    ++a;
  }
  b = a + e;
}

// Q entries (pairs, or columns against plane zr) are accumulated per pass
// over the pixels; each entry sums the same per-pixel terms in the same
// order whatever Q is, so Q only trades registers (2 Q doubles of
// accumulators and Q plane addresses) against passes over W.
template <typename Real, bool GRAM, int Q = 32>
__device__ void weighted_pairs(const Real* __restrict__ W, const Real* __restrict__ Z, int P,
                               int C, int z0, int zr, double* out, double* red) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int E = GRAM ? C * (C + 1) / 2 : C;
  int a0 = 0, b0 = 0;  // first pair of the current chunk (GRAM)
  for (int e0 = 0; e0 < E; e0 += Q) {
    const int ne = min(Q, E - e0);
    int ia[Q], ib[Q];
    {
      int a = a0, b = b0;
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (GRAM) {
          ia[q] = z0 + a;
          ib[q] = z0 + b;
          if (++b >= C) {
            ++a;
            b = a;
          }
        } else {
          ia[q] = z0 + e0 + q;
          ib[q] = zr;
        }
      }
      if (GRAM) {
        a0 = a;
        b0 = b;
      }
    }
    // products in Real, accumulation in float64 (exact-product FMA): the
    // pixel sums carry no float32 summation error into the flux solve.
    double acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = 0.0;
    for (int p = tid; p < P; p += UPH_BLOCK) {
      const Real wv = W[p];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        if (q < ne) {
          const Real za = Z[(size_t)ia[q] * P + p];
          const Real zb = Z[(size_t)ib[q] * P + p];
          acc[q] = fma(double(wv * za), double(zb), acc[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      double v = wave_sum(acc[q]);
      if (lane == 0) red[wid * 32 + q] = v;
    }
    __syncthreads();
    if (tid < ne) {
      double s = red[tid];
#pragma unroll
      for (int w = 1; w < UPH_WAVES; ++w) s += red[w * 32 + tid];
      if (GRAM) {
        int pa, pb;
        pair_of(e0 + tid, C, pa, pb);
        out[pa * C + pb] = s;
        out[pb * C + pa] = s;
      } else {
        out[e0 + tid] = s;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Weighted Gram on the matrix cores: out = Z^T W Z over planes [z0, z0 + C).
// v_mfma_f32_16x16x4f32 tiles of 16 planes; lane l feeds plane (l % 16) at
// four consecutive pixels (float4), lane group l / 16 selects which four, so
// one K-step of four MFMAs consumes 16 pixels.  Products and short sums run
// in float32; every 16 K-steps (256 products per entry) the tile is flushed
// into float64.  Tile pairs (ti <= tj) and, when there are fewer tile pairs
// than waves, pixel splits are spread over the 4 waves; split partials are
// summed in a fixed order (deterministic).
// ---------------------------------------------------------------------------
typedef float uph_f4 __attribute__((ext_vector_type(4)));

__device__ void gram_mfma(const float* __restrict__ W, const float* __restrict__ Z, int P, int C,
                          int z0, double* __restrict__ out, double* __restrict__ part) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nt = (C + 15) >> 4, ntp = nt * (nt + 1) / 2;
  const int S = ntp >= UPH_WAVES ? 1 : UPH_WAVES / ntp;
  const int units = ntp * S;
  const int KS = P >> 4;
  const int col = lane & 15, grp = lane >> 4;
  for (int u = wid; u < units; u += UPH_WAVES) {
    const int tp = u / S, sp = u % S;
    int ti, tj;
    pair_of(tp, nt, ti, tj);
    const int ks0 = (KS * sp) / S, ks1 = (KS * (sp + 1)) / S;
    const int a = ti * 16 + col, b = tj * 16 + col;
    const bool va = a < C, vb = b < C;
    const float* za = Z + (size_t)(z0 + (va ? a : 0)) * P;
    const float* zb = Z + (size_t)(z0 + (vb ? b : 0)) * P;
    double accd[4] = {0.0, 0.0, 0.0, 0.0};
    uph_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    int cnt = 0;
    for (int ks = ks0; ks < ks1; ++ks) {
      const int px = ks * 16 + 4 * grp;
      const uph_f4 wv = *reinterpret_cast<const uph_f4*>(W + px);
      uph_f4 xa = {0.0f, 0.0f, 0.0f, 0.0f}, xb = {0.0f, 0.0f, 0.0f, 0.0f};
      if (va) xa = *reinterpret_cast<const uph_f4*>(za + px);
      if (ti == tj)
        xb = xa;
      else if (vb)
        xb = *reinterpret_cast<const uph_f4*>(zb + px);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.x, wv.x * xb.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.y, wv.y * xb.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.z, wv.z * xb.z, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa.w, wv.w * xb.w, acc, 0, 0, 0);
      if (++cnt == 16) {
        accd[0] += acc.x;
        accd[1] += acc.y;
        accd[2] += acc.z;
        accd[3] += acc.w;
        acc = uph_f4{0.0f, 0.0f, 0.0f, 0.0f};
        cnt = 0;
      }
    }
    accd[0] += acc.x;
    accd[1] += acc.y;
    accd[2] += acc.z;
    accd[3] += acc.w;
    // D[4 * grp + r][col] of the 16 x 16 tile
    double* pt = part + (size_t)u * 256;
#pragma unroll
    for (int r = 0; r < 4; ++r) pt[(4 * grp + r) * 16 + col] = accd[r];
  }
  __syncthreads();
  for (int e = tid; e < ntp * 256; e += UPH_BLOCK) {
    const int tp = e >> 8, i = (e >> 4) & 15, j = e & 15;
    int ti, tj;
    pair_of(tp, nt, ti, tj);
    if (ti == tj && i > j) continue;  // diagonal tiles: upper entries only
    const int a = ti * 16 + i, b = tj * 16 + j;
    if (a >= C || b >= C) continue;
    double v = 0.0;
    for (int sp = 0; sp < S; ++sp) v += part[((size_t)tp * S + sp) * 256 + i * 16 + j];
    out[a * C + b] = v;
    out[b * C + a] = v;
  }
  __syncthreads();
}

// The float64 Gram on the matrix cores (the float64 pre-render evaluation):
// v_mfma_f64_16x16x4f64 tiles of 16 planes with the operand layout of
// gram_mfma (lane l feeds plane l % 16 at four consecutive pixels, two
// double2 loads, lane group l / 16 selects which four); the result tile
// comes back as D[4 r + l / 16][l % 16] in register r of lane l.  Products
// and sums are float64 throughout.  Per 16-pixel step a lane loads 96 bytes
// for 16 products of its tile entries, where the pair loop loads 16 bytes
// per product.
typedef double uph_d2 __attribute__((ext_vector_type(2)));
typedef double uph_d4 __attribute__((ext_vector_type(4)));

__device__ void gram_mfma_f64(const double* __restrict__ W, const double* __restrict__ Z, int P,
                              int C, int z0, double* __restrict__ out,
                              double* __restrict__ part) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nt = (C + 15) >> 4, ntp = nt * (nt + 1) / 2;
  const int S = ntp >= UPH_WAVES ? 1 : UPH_WAVES / ntp;
  const int units = ntp * S;
  const int KS = P >> 4;
  const int col = lane & 15, grp = lane >> 4;
  for (int u = wid; u < units; u += UPH_WAVES) {
    const int tp = u / S, sp = u % S;
    int ti, tj;
    pair_of(tp, nt, ti, tj);
    const int ks0 = (KS * sp) / S, ks1 = (KS * (sp + 1)) / S;
    const int a = ti * 16 + col, b = tj * 16 + col;
    const bool va = a < C, vb = b < C;
    const double* za = Z + (size_t)(z0 + (va ? a : 0)) * P;
    const double* zb = Z + (size_t)(z0 + (vb ? b : 0)) * P;
    uph_d4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int ks = ks0; ks < ks1; ++ks) {
      const int px = ks * 16 + 4 * grp;
      const uph_d2 w0 = *reinterpret_cast<const uph_d2*>(W + px);
      const uph_d2 w1 = *reinterpret_cast<const uph_d2*>(W + px + 2);
      uph_d2 a0 = {0.0, 0.0}, a1 = {0.0, 0.0}, b0 = {0.0, 0.0}, b1 = {0.0, 0.0};
      if (va) {
        a0 = *reinterpret_cast<const uph_d2*>(za + px);
        a1 = *reinterpret_cast<const uph_d2*>(za + px + 2);
      }
      if (ti == tj) {
        b0 = a0;
        b1 = a1;
      } else if (vb) {
        b0 = *reinterpret_cast<const uph_d2*>(zb + px);
        b1 = *reinterpret_cast<const uph_d2*>(zb + px + 2);
      }
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.x, w0.x * b0.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a0.y, w0.y * b0.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.x, w1.x * b1.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a1.y, w1.y * b1.y, acc, 0, 0, 0);
    }
    // register r of lane l holds D[4 r + l / 16][l % 16] (the float64
    // layout differs from the float32 one)
    double* pt = part + (size_t)u * 256;
    pt[(0 + grp) * 16 + col] = acc.x;
    pt[(4 + grp) * 16 + col] = acc.y;
    pt[(8 + grp) * 16 + col] = acc.z;
    pt[(12 + grp) * 16 + col] = acc.w;
  }
  __syncthreads();
  for (int e = tid; e < ntp * 256; e += UPH_BLOCK) {
    const int tp = e >> 8, i = (e >> 4) & 15, j = e & 15;
    int ti, tj;
    pair_of(tp, nt, ti, tj);
    if (ti == tj && i > j) continue;  // diagonal tiles: upper entries only
    const int a = ti * 16 + i, b = tj * 16 + j;
    if (a >= C || b >= C) continue;
    double v = 0.0;
    for (int sp = 0; sp < S; ++sp) v += part[((size_t)tp * S + sp) * 256 + i * 16 + j];
    out[a * C + b] = v;
    out[b * C + a] = v;
  }
  __syncthreads();
}

// The Gram out (C x C) from its 16 x 16 tile pairs (ti <= tj, pair_of
// order), each computed over all pixels as one unit of gram_mfma.
__device__ void gram_from_pairs(const double* __restrict__ pairs, int C,
                                double* __restrict__ out) {
  const int nt = (C + 15) >> 4, ntp = nt * (nt + 1) / 2;
  for (int e = threadIdx.x; e < ntp * 256; e += UPH_BLOCK) {
    const int tp = e >> 8, i = (e >> 4) & 15, j = e & 15;
    int ti, tj;
    pair_of(tp, nt, ti, tj);
    if (ti == tj && i > j) continue;
    const int a = ti * 16 + i, b = tj * 16 + j;
    if (a >= C || b >= C) continue;
    const double v = pairs[e];
    out[a * C + b] = v;
    out[b * C + a] = v;
  }
  __syncthreads();
}

// The same Gram with bfloat16 inputs: one v_mfma_f32_16x16x16_bf16 takes
// each lane's four consecutive pixels (lane l: plane l % 16, pixels
// 4 (l / 16) .. + 3 of the 16-pixel step), so a K-step is one instruction
// instead of four.  Operands are rounded to bfloat16 (nearest even): plane a
// and W * plane b; products accumulate in float32 and are flushed into
// float64 every 16 K-steps as in gram_mfma.  The flux solve then takes three
// refinement steps on float64 residual projections instead of one, which
// restores the fluxes' accuracy.
typedef short uph_s4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rne(float v) {
  unsigned u = __float_as_uint(v);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return short(u >> 16);
}

__device__ void gram_mfma_bf16(const float* __restrict__ W, const float* __restrict__ Z, int P,
                               int C, int z0, double* __restrict__ out,
                               double* __restrict__ part) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int nt = (C + 15) >> 4, ntp = nt * (nt + 1) / 2;
  const int S = ntp >= UPH_WAVES ? 1 : UPH_WAVES / ntp;
  const int units = ntp * S;
  const int KS = P >> 4;
  const int col = lane & 15, grp = lane >> 4;
  for (int u = wid; u < units; u += UPH_WAVES) {
    const int tp = u / S, sp = u % S;
    int ti, tj;
    pair_of(tp, nt, ti, tj);
    const int ks0 = (KS * sp) / S, ks1 = (KS * (sp + 1)) / S;
    const int a = ti * 16 + col, b = tj * 16 + col;
    const bool va = a < C, vb = b < C;
    const float* za = Z + (size_t)(z0 + (va ? a : 0)) * P;
    const float* zb = Z + (size_t)(z0 + (vb ? b : 0)) * P;
    double accd[4] = {0.0, 0.0, 0.0, 0.0};
    uph_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    int cnt = 0;
    for (int ks = ks0; ks < ks1; ++ks) {
      const int px = ks * 16 + 4 * grp;
      const uph_f4 wv = *reinterpret_cast<const uph_f4*>(W + px);
      uph_f4 xa = {0.0f, 0.0f, 0.0f, 0.0f}, xb = {0.0f, 0.0f, 0.0f, 0.0f};
      if (va) xa = *reinterpret_cast<const uph_f4*>(za + px);
      if (ti == tj)
        xb = xa;
      else if (vb)
        xb = *reinterpret_cast<const uph_f4*>(zb + px);
      const uph_s4 av = {bf16_rne(xa.x), bf16_rne(xa.y), bf16_rne(xa.z), bf16_rne(xa.w)};
      const uph_s4 bv = {bf16_rne(wv.x * xb.x), bf16_rne(wv.y * xb.y), bf16_rne(wv.z * xb.z),
                         bf16_rne(wv.w * xb.w)};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
      if (++cnt == 16) {
        accd[0] += acc.x;
        accd[1] += acc.y;
        accd[2] += acc.z;
        accd[3] += acc.w;
        acc = uph_f4{0.0f, 0.0f, 0.0f, 0.0f};
        cnt = 0;
      }
    }
    accd[0] += acc.x;
    accd[1] += acc.y;
    accd[2] += acc.z;
    accd[3] += acc.w;
    double* pt = part + (size_t)u * 256;
#pragma unroll
    for (int r = 0; r < 4; ++r) pt[(4 * grp + r) * 16 + col] = accd[r];
  }
  __syncthreads();
  for (int e = tid; e < ntp * 256; e += UPH_BLOCK) {
    const int tp = e >> 8, i = (e >> 4) & 15, j = e & 15;
    int ti, tj;
    pair_of(tp, nt, ti, tj);
    if (ti == tj && i > j) continue;
    const int a = ti * 16 + i, b = tj * 16 + j;
    if (a >= C || b >= C) continue;
    double v = 0.0;
    for (int sp = 0; sp < S; ++sp) v += part[((size_t)tp * S + sp) * 256 + i * 16 + j];
    out[a * C + b] = v;
    out[b * C + a] = v;
  }
  __syncthreads();
}

template <typename Real>
__device__ __forceinline__ void gram(const Real* W, const Real* Z, int P, int C, int z0,
                                     double* out, double* part, double* red) {
  if constexpr (sizeof(Real) == 4) {
    (void)red;
    gram_mfma(W, Z, P, C, z0, out, part);
  } else {
    (void)part;
    weighted_pairs<Real, true>(W, Z, P, C, z0, 0, out, red);
  }
}

// ---------------------------------------------------------------------------
// Star member: PSF polynomial at (x, y), separable Lanczos-3 shift by the
// remainder of (x - x0b, y - y0b), placed on the B x B window at
// round(pos - origin) - T//2 (source_model.py:322-328, island_solver.py:512).
// Only the part of the stamp that lands in the window is evaluated.  With
// tangents, also d/dx and d/dy (polynomial derivative + tap derivative).
// ---------------------------------------------------------------------------
template <typename Real>
__device__ void render_star(const IslandArgs& a, int T, int B, const float* coefs, int b, float xf,
                            float yf, float x0f, float y0f, Real* __restrict__ m_out,
                            Real* __restrict__ tx_out, Real* __restrict__ ty_out,
                            Real* __restrict__ ws, Real* taps, Real* bas) {
  const int nt = a.nt, tid = threadIdx.x;
  const bool grad = tx_out != nullptr;
  const float* geo = a.psf_geo + 4 * b;
  const Real x = Real(xf), y = Real(yf);
  const Real xg = Real(xf - x0f), yg = Real(yf - y0f);
  const Real rxg = rint(xg), ryg = rint(yg);
  const Real dx = xg - rxg, dy = yg - ryg;
  const int ix = int(rxg) - T / 2, iy = int(ryg) - T / 2;
  const Real xs = Real(geo[1]), ys = Real(geo[3]);
  if (tid == 0) lanczos_taps(dx, taps, grad ? taps + 7 : (Real*)nullptr);
  if (tid == 64) lanczos_taps(dy, taps + 14, grad ? taps + 21 : (Real*)nullptr);
  if (tid == 128) {
    const Real X = (x - Real(geo[0])) / xs, Y = (y - Real(geo[2])) / ys;
    psf_basis(X, Y, nt, bas, bas + 21, bas + 42);
  }
  __syncthreads();
  const int wx0 = max(0, ix), wx1 = min(B, ix + T), wy0 = max(0, iy), wy1 = min(B, iy + T);
  if (wx1 <= wx0 || wy1 <= wy0) return;  // uniform across the block
  const int sx0 = wx0 - ix, sx1 = wx1 - ix, sy0 = wy0 - iy, sy1 = wy1 - iy;
  const int rx0 = max(0, sx0 - 3), rx1 = min(T, sx1 + 3);
  const int ry0 = max(0, sy0 - 3), ry1 = min(T, sy1 + 3);
  const int RW = rx1 - rx0, RH = ry1 - ry0, SW = sx1 - sx0, SH = sy1 - sy0;
  const float* Cf = coefs + (size_t)b * nt * T * T;
  const int TT = T * T;
  Real* P = ws;
  Real* Px = ws + TT;
  Real* Py = ws + 2 * TT;
  Real* H0 = ws + 3 * TT;
  Real* Hx = ws + 4 * TT;
  Real* Hy = ws + 5 * TT;
  for (int e = tid; e < RH * RW; e += UPH_BLOCK) {
    const int r = ry0 + e / RW, c = rx0 + e % RW;
    Real s, sx, sy;
    psf_poly_at(Cf + r * T + c, (size_t)TT, nt, bas, grad, s, sx, sy);
    P[e] = s;
    if (grad) {
      Px[e] = sx / xs;
      Py[e] = sy / ys;
    }
  }
  __syncthreads();
  Real Lx[7], dLx[7], Ly[7], dLy[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    Lx[k] = taps[k];
    Ly[k] = taps[14 + k];
    dLx[k] = grad ? taps[7 + k] : Real(0);
    dLy[k] = grad ? taps[21 + k] : Real(0);
  }
  for (int e = tid; e < RH * SW; e += UPH_BLOCK) {
    const int rr = e / SW, c = sx0 + e % SW;
    Real h = 0, hx = 0, hy = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int cc = c + k - 3;
      if (cc < rx0 || cc >= rx1) continue;
      const int o = rr * RW + cc - rx0;
      const Real v = P[o];
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
    const int r = sy0 + e / SW, cc = e % SW;
    Real m = 0, tx = 0, ty = 0;
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int rr = r + k - 3;
      if (rr < ry0 || rr >= ry1) continue;
      const int o = (rr - ry0) * SW + cc;
      const Real h = H0[o];
      m = fma(Ly[k], h, m);
      if (grad) {
        tx = fma(Ly[k], Hx[o], tx);
        ty = fma(Ly[k], Hy[o], fma(dLy[k], h, ty));
      }
    }
    const int wi = (r + iy) * B + (cc + sx0 + ix);
    m_out[wi] = m;
    if (grad) {
      tx_out[wi] = tx;
      ty_out[wi] = ty;
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Extended member: PSF polynomial over the whole stamp (and its x/y
// derivatives), the analytic-FT galaxy convolution with the raw tangents of
// the member's live parameters, then the same separable Lanczos-3 shift and
// placement as a star (source_model.py:264-328 with the fft backend,
// island_solver.py:441-562).  Tangent planes of live parameter l go to
// tan_base + l * B^2.
// ---------------------------------------------------------------------------
template <typename Real>
__device__ void render_galaxy(const IslandArgs& a, IslandSh<Real>& sh, int b, const float* p6,
                              int r, float x0f, float y0f, Real* __restrict__ m_out,
                              Real* __restrict__ tan_base, Real* __restrict__ gws) {
  const int T = sh.T, B = sh.B, nt = a.nt, tid = threadIdx.x, TT = T * T, P = B * B;
  const bool grads = tan_base != nullptr;
  int need = 0;
  if (grads)
    for (int l = sh.first_live[r]; l < sh.first_live[r + 1]; ++l) need |= 1 << sh.live_j[l];
  const bool gxy = need & 3;
  const float* geo = a.psf_geo + 4 * b;
  const float xf = p6[0], yf = p6[1];
  const Real x = Real(xf), y = Real(yf);
  const Real xg = Real(xf - x0f), yg = Real(yf - y0f);
  const Real rxg = rint(xg), ryg = rint(yg);
  const Real dx = xg - rxg, dy = yg - ryg;
  const int ix = int(rxg) - T / 2, iy = int(ryg) - T / 2;
  const Real xs = Real(geo[1]), ys = Real(geo[3]);
  {
    if (tid == 0) {
      GalParams<Real>& gp = sh.gp;
      double n = double(sh.nfz[r]);
      gp.dndu = 0.0;
      if (sh.fn[r]) {
        const double u = double(p6[5]);
        n = kNMin + (kNMax - kNMin) / (1.0 + exp(-u));
        gp.dndu = (n - kNMin) * (kNMax - n) / (kNMax - kNMin);
      }
      gp.n = n;
      gp.K = int(a.sersic[3]);
      mog_lookup(a.sersic, n, gp.a, gp.v, gp.core, gp.da, gp.dv, gp.dcore);
      ellipse_terms(double(p6[2]), double(p6[3]), double(p6[4]), gp.t, gp.dt);
    }
    if (tid == 64) lanczos_taps(dx, sh.taps, grads ? sh.taps + 7 : (Real*)nullptr);
    if (tid == 128) lanczos_taps(dy, sh.taps + 14, grads ? sh.taps + 21 : (Real*)nullptr);
    if (tid == 192) {
      const Real X = (x - Real(geo[0])) / xs, Y = (y - Real(geo[2])) / ys;
      psf_basis(X, Y, nt, sh.bas, sh.bas + 21, sh.bas + 42);
    }
    __syncthreads();
    const GalLayout L = galaxy_layout(T);
    const float* Cf = sh.coefs + (size_t)b * nt * TT;
    for (int e = tid; e < TT; e += UPH_BLOCK) {
      Real v, vx, vy;
      psf_poly_at(Cf + e, (size_t)TT, nt, (const Real*)sh.bas, gxy, v, vx, vy);
      gws[L.P + e] = v;
      if (gxy) {
        gws[L.Px + e] = vx / xs;
        gws[L.Py + e] = vy / ys;
      }
    }
    __syncthreads();
    galaxy_raw<Real, false>(sh.gp, T, need, gws);
  }
  const GalLayout L = galaxy_layout(T);
  // ---- separable shift + placement of G and the raw tangents ----
  const int wx0 = max(0, ix), wx1 = min(B, ix + T), wy0 = max(0, iy), wy1 = min(B, iy + T);
  if (wx1 <= wx0 || wy1 <= wy0) return;  // uniform
  const int sx0 = wx0 - ix, sx1 = wx1 - ix, sy0 = wy0 - iy, sy1 = wy1 - iy;
  const int ry0 = max(0, sy0 - 3), ry1 = min(T, sy1 + 3);
  const int RH = ry1 - ry0, SW = sx1 - sx0, SH = sy1 - sy0;
  Real Lx[7], dLx[7], Ly[7], dLy[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    Lx[k] = sh.taps[k];
    Ly[k] = sh.taps[14 + k];
    dLx[k] = grads ? sh.taps[7 + k] : Real(0);
    dLy[k] = grads ? sh.taps[21 + k] : Real(0);
  }
  const Real* G = gws + L.G;
  Real* Hh = gws + L.Hh;  // plane q of the horizontal pass at Hh + q * T^2
  const int RS = RH * SW;
  for (int e = tid; e < RS; e += UPH_BLOCK) {
    const int rr = ry0 + e / SW, c = sx0 + e % SW;
    Real h[7] = {0, 0, 0, 0, 0, 0, 0};
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
  const Real dndu = Real(sh.gp.dndu);
  for (int e = tid; e < SH * SW; e += UPH_BLOCK) {
    const int rw = sy0 + e / SW, cc = e % SW;
    Real v[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 7; ++k) {
      const int rr = rw + k - 3;
      if (rr < ry0 || rr >= ry1) continue;
      const int o = (rr - ry0) * SW + cc;
      const Real h0 = Hh[o];
      v[0] = fma(Ly[k], h0, v[0]);
      if (need & 2) v[2] = fma(Ly[k], Hh[(size_t)2 * TT + o], fma(dLy[k], h0, v[2]));
#pragma unroll
      for (int j = 0; j < 6; ++j)
        if (j != 1 && (need & (1 << j))) v[1 + j] = fma(Ly[k], Hh[(size_t)(1 + j) * TT + o], v[1 + j]);
    }
    const int wi = (rw + iy) * B + (cc + sx0 + ix);
    m_out[wi] = v[0];
    if (grads) {
      for (int l = sh.first_live[r]; l < sh.first_live[r + 1]; ++l) {
        const int j = sh.live_j[l];
        tan_base[(size_t)l * P + wi] = j == 5 ? v[6] * dndu : v[1 + j];
      }
    }
  }
  __syncthreads();
}

// Pre-render variant of the member render: copy the window part of the
// member's shifted planes (model, then the live tangents; 7 planes per band
// for an extended member, 3 for a star) into the island planes.  The stamp
// sits at rint(x - x0) - T / 2 as in the in-place render.
// The pre-rendered planes of the evaluation's precision.
template <typename Real>
__device__ __forceinline__ const Real* pre_planes_buf(const IslandArgsPre& ap) {
  if constexpr (sizeof(Real) == 4)
    return ap.pbuf;
  else
    return ap.pbuf64;
}

template <typename Real>
__device__ void place_member(const IslandArgs& a, IslandSh<Real>& sh, int b, const float* p6,
                             int r, float x0f, float y0f, Real* __restrict__ m_out,
                             Real* __restrict__ tan_base) {
  const IslandArgsPre& ap = static_cast<const IslandArgsPre&>(a);
  const int T = sh.T, B = sh.B, P = B * B, tid = threadIdx.x;
  const Real xg = Real(p6[0] - x0f), yg = Real(p6[1] - y0f);
  const int ix = int(rint(xg)) - T / 2, iy = int(rint(yg)) - T / 2;
  const int wx0 = max(0, ix), wx1 = min(B, ix + T), wy0 = max(0, iy), wy1 = min(B, iy + T);
  if (wx1 <= wx0 || wy1 <= wy0) return;  // uniform
  const int sx0 = wx0 - ix, sy0 = wy0 - iy, SW = wx1 - wx0, SH = wy1 - wy0;
  const int g = ap.pre_of_slot[(size_t)sh.moff + sh.mem[r]];
  const int side = pre_side(T, B), PP = side * side;
  const Real* blk = pre_planes_buf<Real>(ap) + ap.pre_off[g] + (size_t)b * pre_planes(sh.st[r] == 0) * PP;
  const int l0 = sh.first_live[r], l1 = sh.first_live[r + 1];
  for (int e = tid; e < SH * SW; e += UPH_BLOCK) {
    const int rs = sy0 + e / SW, cs = sx0 + e % SW;
    const int wi = (rs + iy) * B + (cs + ix), so = (e / SW) * side + e % SW;
    m_out[wi] = Real(blk[so]);
    if (tan_base)
      for (int l = l0; l < l1; ++l)
        tan_base[(size_t)l * P + wi] = Real(blk[(size_t)(1 + sh.live_j[l]) * PP + so]);
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// One band of one island: window copy, member renders, Gram, VarPro fluxes,
// residual, chi2 and (with grads) accumulation of H, g.
// ---------------------------------------------------------------------------
// Residual projections of the pre-render evaluation run 8 columns per pixel
// pass: 112 VGPRs, four waves per SIMD (32 columns need 280 with the
// accumulators and addresses, one wave per SIMD).
constexpr int kPreQ = 8;

template <typename Real, bool kPre>
__device__ double eval_band(const IslandArgs& a, IslandSh<Real>& sh, int isl, int b,
                          const float* pv, bool grads, bool accumulate, Real* planes, Real* ws,
                          Real* gws, double* G, double* part, const Small& sm, float* flux_out,
                          float* ferr_out, int* ngood) {
  const int B = sh.B, P = B * B, tid = threadIdx.x;
  const int Nr = sh.Nr, nL = grads ? sh.nL : 0;
  Real* Wp = planes;
  Real* Rp = planes + P;
  Real* Dp = planes + 2 * P;
  Real* Mp = planes + 3 * P;
  Real* Tp = Mp + (size_t)Nr * P;
  const int ox = a.ox[isl], oy = a.oy[isl];
  int cnt = 0;
  for (int e = tid; e < P; e += UPH_BLOCK) {
    const int r = e / B, c = e % B;
    const size_t gi = ((size_t)b * a.Hp + oy + r) * a.Wp + ox + c;
    const float wv = a.wstack[gi];
    Wp[e] = Real(wv);
    Dp[e] = Real(a.dstack[gi]);
    cnt += wv > 0.0f;
  }
  if (ngood) {
    double tot = block_sum(double(cnt), sh.redd);
    if (tid == 0) *ngood += int(tot);
  }
  for (size_t e = tid; e < (size_t)(Nr + nL) * P; e += UPH_BLOCK) Mp[e] = Real(0);
  __syncthreads();
  const float x0f = a.x0b[isl], y0f = a.y0b[isl];
  for (int r = 0; r < Nr; ++r) {
    const int k = sh.mem[r];
    const float xk = pv[6 * k], yk = pv[6 * k + 1];
    Real* tx = nullptr;
    Real* ty = nullptr;
    if (grads) {
      // star members: live params are x, y (in that order) when kept
      for (int l = sh.first_live[r]; l < sh.first_live[r + 1]; ++l) {
        if (sh.live_j[l] == 0) tx = Tp + (size_t)l * P;
        if (sh.live_j[l] == 1) ty = Tp + (size_t)l * P;
      }
    }
    if constexpr (kPre) {
      place_member<Real>(a, sh, b, pv + 6 * k, r, x0f, y0f, Mp + (size_t)r * P,
                         grads ? Tp : nullptr);
    } else if (sh.st[r] == 0) {
      Real* wsx = ws;
      if (grads && (tx == nullptr || ty == nullptr)) {
        // a frozen coordinate: render both, discard into the R plane
        if (tx == nullptr) tx = Rp;
        if (ty == nullptr) ty = Rp;
      }
      render_star<Real>(a, sh.T, B, sh.coefs, b, xk, yk, x0f, y0f, Mp + (size_t)r * P,
                        grads ? tx : nullptr, grads ? ty : nullptr, wsx, sh.taps, sh.bas);
    } else {
      render_galaxy<Real>(a, sh, b, pv + 6 * k, r, x0f, y0f, Mp + (size_t)r * P,
                          grads ? Tp : nullptr, gws);
    }
  }
  // Gram of [d, models, tangents] under W.
  const int C = 1 + Nr + nL;
  int bf16 = 0;  // bfloat16 Gram operands (float32 pre-render variant, on request)
  if constexpr (kPre && sizeof(Real) == 4) bf16 = static_cast<const IslandArgsPre&>(a).gram_bf16;
  if constexpr (kPre && sizeof(Real) == 4) {
    const IslandArgsPre& ap = static_cast<const IslandArgsPre&>(a);
    const int slot = (ap.big_n > 0 && !bf16) ? ap.big_slot[isl] : -1;
    if (bf16)
      gram_mfma_bf16(Wp, planes, P, C, 2, G, part);
    else if (slot >= 0)
      gram_from_pairs(ap.big_part + ((size_t)slot * a.nb + b) * ap.big_ntp * 256, C, G);
    else
      gram<Real>(Wp, planes, P, C, 2, G, part, sh.red);
  } else if constexpr (kPre) {
    const IslandArgsPre& ap = static_cast<const IslandArgsPre&>(a);
    const int slot = ap.big_n > 0 ? ap.big_slot[isl] : -1;
    if (slot >= 0)
      gram_from_pairs(ap.big_part + ((size_t)slot * a.nb + b) * ap.big_ntp * 256, C, G);
    else
      gram_mfma_f64(Wp, planes, P, C, 2, G, part);
  } else {
    gram<Real>(Wp, planes, P, C, 2, G, part, sh.red);
  }
  // VarPro fluxes (float64 LU), A64 = A + reg I on real members.
  const double reg = double(a.reg);
  for (int e = tid; e < Nr * Nr; e += UPH_BLOCK) {
    const int i = e / Nr, j = e % Nr;
    const double v = G[(1 + i) * C + 1 + j];
    sm.Araw[e] = v;
    sm.LU[e] = v + (i == j ? reg : 0.0);
  }
  for (int i = tid; i < Nr; i += UPH_BLOCK) sm.b[i] = G[1 + i];
  __syncthreads();
  block_lu(sm.LU, Nr, sm.pivA, &sh.pivot);
  if (tid == 0) {
    for (int i = 0; i < Nr; ++i) sm.f[i] = sm.b[i];
    lu_solve_one(sm.LU, sm.pivA, Nr, sm.f);
    for (int i = 0; i < Nr; ++i) sh.fR[i] = Real(sm.f[i]);
  }
  __syncthreads();
  // One refinement step: c0 = M^T W (d - M f0) accumulated in float64 from
  // the per-pixel residual, f = f0 + A64^-1 (c0 - reg f0).  The flux then
  // carries the accuracy of the residual sums, not of the Gram.
  for (int p = tid; p < P; p += UPH_BLOCK) {
    Real model = 0;
    for (int r = 0; r < Nr; ++r) model = fma(sh.fR[r], Mp[(size_t)r * P + p], model);
    Rp[p] = Dp[p] - model;
  }
  __syncthreads();
  weighted_pairs<Real, false, kPre ? kPreQ : 32>(Wp, planes, P, Nr, 3, 1, sm.colws, sh.red);
  if (tid == 0) {
    double* dx = sm.c;
    for (int i = 0; i < Nr; ++i) dx[i] = sm.colws[i] - reg * double(sh.fR[i]);
    lu_solve_one(sm.LU, sm.pivA, Nr, dx);
    for (int i = 0; i < Nr; ++i) {
      sh.fR[i] = Real(double(sh.fR[i]) + dx[i]);
      sm.f[i] = double(sh.fR[i]);
    }
  }
  __syncthreads();
  if constexpr (kPre && sizeof(Real) == 4) {
    // with the bfloat16 Gram two more refinement steps: each one multiplies
    // the flux error by about (Gram error) x (condition number)
    if (bf16) {
      auto refine = [&]() {
        for (int p = tid; p < P; p += UPH_BLOCK) {
          Real model = 0;
          for (int r = 0; r < Nr; ++r) model = fma(sh.fR[r], Mp[(size_t)r * P + p], model);
          Rp[p] = Dp[p] - model;
        }
        __syncthreads();
        weighted_pairs<Real, false, kPreQ>(Wp, planes, P, Nr, 3, 1, sm.colws, sh.red);
        if (tid == 0) {
          double* dx = sm.c;
          double moved = 0.0;  // largest correction relative to its flux
          for (int i = 0; i < Nr; ++i) dx[i] = sm.colws[i] - reg * double(sh.fR[i]);
          lu_solve_one(sm.LU, sm.pivA, Nr, dx);
          for (int i = 0; i < Nr; ++i) {
            moved = fmax(moved, fabs(dx[i]) / fmax(fabs(double(sh.fR[i])), 1.0));
            sh.fR[i] = Real(double(sh.fR[i]) + dx[i]);
            sm.f[i] = double(sh.fR[i]);
          }
          sh.redd[0] = moved;
        }
        __syncthreads();
      };
      refine();
      refine();
      // the refinement has not settled (Gram error times the condition
      // number near 1): the island is refitted on float32 operands
      if (tid == 0 && sh.redd[0] > 1e-6)
        atomicOr(static_cast<const IslandArgsPre&>(a).gram_unsettled + isl, 1);
    }
  }
  if (flux_out) {
    // final fluxes and sqrt(diag(A^-1))
    lu_inverse_diag(sm.LU, sm.pivA, Nr, sm.s, sm.colws);
    __syncthreads();
    for (int r = tid; r < Nr; r += UPH_BLOCK) {
      flux_out[sh.mem[r]] = float(sh.fR[r]);
      ferr_out[sh.mem[r]] = float(sqrt(fmax(sm.s[r], 1e-30)));
    }
    __syncthreads();
    return 0.0;
  }
  // residual R = d - sum_k f_k m_k ; chi2 = sum (sqrt(w) R)^2
  double acc = 0.0;
  for (int p = tid; p < P; p += UPH_BLOCK) {
    Real model = 0;
    for (int r = 0; r < Nr; ++r) model = fma(sh.fR[r], Mp[(size_t)r * P + p], model);
    const Real R = Dp[p] - model;
    const Real rw = sqrt(Wp[p]) * R;
    acc = fma(double(rw), double(rw), acc);
    Rp[p] = R;
  }
  const double chi2b = block_sum(acc, sh.red);
  if (!grads) return chi2b;
  // c = M^T W R, s = T^T W R (one pass over the Nr + nL planes)
  weighted_pairs<Real, false, kPre ? kPreQ : 32>(Wp, planes, P, Nr + nL, 3, 1, sm.colws, sh.red);
  for (int i = tid; i < Nr + nL; i += UPH_BLOCK) {
    if (i < Nr)
      sm.c[i] = sm.colws[i];
    else
      sm.s[i - Nr] = sm.colws[i];
  }
  for (int e = tid; e < Nr * nL; e += UPH_BLOCK) {
    const int r = e / nL, j = e % nL;
    sm.U[e] = G[(1 + r) * C + 1 + Nr + j];
  }
  for (int e = tid; e < nL * nL; e += UPH_BLOCK) {
    const int i = e / nL, j = e % nL;
    sm.Gt[e] = G[(1 + Nr + i) * C + 1 + Nr + j];
  }
  __syncthreads();
  // D_j = A64^-1 (e_k s_j - f_k U_j), one right-hand side per thread
  for (int j = tid; j < nL; j += UPH_BLOCK) {
    double* x = sm.D + (size_t)j * Nr;
    const int kr = sh.live_r[j];
    for (int r = 0; r < Nr; ++r) x[r] = -sm.f[kr] * sm.U[r * nL + j];
    x[kr] += sm.s[j];
    lu_solve_one(sm.LU, sm.pivA, Nr, x);
  }
  __syncthreads();
  for (int e = tid; e < nL * Nr; e += UPH_BLOCK) {
    const int j = e / Nr, r = e % Nr;
    double s = 0.0;
    for (int q = 0; q < Nr; ++q) s += sm.Araw[r * Nr + q] * sm.D[j * Nr + q];
    sm.AD[e] = s;
  }
  __syncthreads();
  for (int e = tid; e < nL * nL; e += UPH_BLOCK) {
    const int i = e / nL, j = e % nL;
    const double* Di = sm.D + (size_t)i * Nr;
    const double* Dj = sm.D + (size_t)j * Nr;
    const double* ADj = sm.AD + (size_t)j * Nr;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int r = 0; r < Nr; ++r) {
      t0 += Di[r] * ADj[r];
      t1 += Di[r] * sm.U[r * nL + j];
      t2 += Dj[r] * sm.U[r * nL + i];
    }
    const double fi = sm.f[sh.live_r[i]], fj = sm.f[sh.live_r[j]];
    const double h = t0 + fj * t1 + fi * t2 + fi * fj * sm.Gt[e];
    sm.H[e] = accumulate ? sm.H[e] + h : h;
  }
  for (int j = tid; j < nL; j += UPH_BLOCK) {
    const double* Dj = sm.D + (size_t)j * Nr;
    double t = 0.0;
    for (int r = 0; r < Nr; ++r) t += Dj[r] * sm.c[r];
    const double gj = -(t + sm.f[sh.live_r[j]] * sm.s[j]);
    sm.g[j] = accumulate ? sm.g[j] + gj : gj;
  }
  __syncthreads();
  return chi2b;
}

template <typename Real>
__device__ double chi2_all(const IslandArgs& a, IslandSh<Real>& sh, int isl, const float* pv,
                          Real* planes, Real* ws, Real* gws, double* G, double* part,
                          const Small& sm) {
  double c = 0.0;
  for (int b = 0; b < a.nb; ++b)
    c += eval_band<Real, false>(a, sh, isl, b, pv, false, false, planes, ws, gws, G, part, sm, nullptr,
                         nullptr, nullptr);
  return c;
}

// Class shape and offsets of island `isl` into sh (called by one thread).
template <typename Real>
__device__ void read_shape(const IslandArgs& a, IslandSh<Real>& sh, int isl) {
  if (a.shape) {
    const int32_t* sp = a.shape + 4 * (size_t)isl;
    sh.N = sp[0];
    sh.B = sp[1];
    sh.T = a.tier_T[sp[2]];
    sh.coefs = a.tier_coefs[sp[2]];
    sh.nLc = sp[3];
    sh.moff = a.mem_off[isl];
    sh.foff = a.flux_off[isl];
  } else {
    sh.N = a.N;
    sh.B = a.B;
    sh.T = a.T;
    sh.coefs = a.psf_coefs;
    sh.nLc = a.nL_max;
    sh.moff = (long long)isl * a.N;
    sh.foff = (long long)isl * a.nb * a.N;
  }
}

// Parameters (from psrc, laid out like p0), box bounds and the real-member
// and live-parameter lists of the island whose shape is in sh.
// Block-cooperative; ends with a barrier.
template <typename Real>
__device__ void describe_island(const IslandArgs& a, IslandSh<Real>& sh, const float* psrc) {
  const int tid = threadIdx.x, N = sh.N, nP = 6 * N;
  const size_t m0 = (size_t)sh.moff;
  for (int q = tid; q < nP; q += UPH_BLOCK) {
    const size_t g = m0 * 6 + q;
    sh.p[q] = psrc[g];
    sh.lo[q] = a.lower[g];
    sh.hi[q] = a.upper[g];
  }
  if (tid == 0) {
    int nr = 0, nl = 0;
    for (int k = 0; k < N; ++k) {
      const size_t g = m0 + k;
      if (!a.mmask[g]) continue;
      sh.mem[nr] = k;
      sh.st[nr] = a.stype[g];
      sh.nfz[nr] = a.n_frozen[g];
      sh.fn[nr] = a.free_n[g];
      sh.first_live[nr] = nl;
      for (int j = 0; j < 6; ++j) {
        if (a.keep[g * 6 + j]) {
          sh.live_r[nl] = nr;
          sh.live_j[nl] = j;
          ++nl;
        }
      }
      ++nr;
    }
    sh.first_live[nr] = nl;
    sh.Nr = nr;
    sh.nL = nl;
  }
  __syncthreads();
}

// Damped diagonal entry of the LM system: H_ii + lam H_ii + 1e-8.
__device__ __forceinline__ double damp_diag(double h, double lam) { return fma(lam, h, h) + 1e-8; }

// LM proposal: LU of the damped system Hd (in place), delta = Hd^-1 rhs
// (rhs overwritten), p_try = clip(p - delta) within the box bounds.
template <typename Real>
__device__ void lm_propose(const IslandArgs& a, IslandSh<Real>& sh, double* Hd, double* rhs,
                           int* piv) {
  const int tid = threadIdx.x, nL = sh.nL, nP = 6 * sh.N;
  block_lu(Hd, nL, piv, &sh.pivot);
  if (tid == 0) lu_solve_one(Hd, piv, nL, rhs);
  __syncthreads();
  for (int q = tid; q < nP; q += UPH_BLOCK) sh.ptry[q] = sh.p[q];
  __syncthreads();
  if (tid == 0) {
    const float clipv[6] = {a.pos_step_max, a.pos_step_max, a.shape_step_max,
                            a.shape_step_max, a.shape_step_max, a.un_step_max};
    for (int l = 0; l < nL; ++l) {
      const int q = 6 * sh.mem[sh.live_r[l]] + sh.live_j[l];
      float d = float(-rhs[l]);
      const float cl = clipv[sh.live_j[l]];
      d = fminf(fmaxf(d, -cl), cl);
      sh.ptry[q] = sh.p[q] + d;
    }
    for (int q = 0; q < nP; ++q) sh.ptry[q] = fminf(fmaxf(sh.ptry[q], sh.lo[q]), sh.hi[q]);
  }
  __syncthreads();
}

// Accept / reject the trial with chi2 `ct` and update the convergence
// prongs (one thread).
template <typename Real>
__device__ void lm_accept(const IslandArgs& a, IslandSh<Real>& sh, double ct) {
  const double cur = sh.chi2;
  const bool accept = ct < cur;
  float pos_step = 0.0f, shape_step = 0.0f;
  if (accept) {
    for (int r = 0; r < sh.Nr; ++r) {
      const int k = sh.mem[r];
      const float sx = sh.ptry[6 * k] - sh.p[6 * k];
      const float sy = sh.ptry[6 * k + 1] - sh.p[6 * k + 1];
      pos_step = fmaxf(pos_step, hypotf(sx, sy));
      for (int j = 2; j < 6; ++j)
        shape_step = fmaxf(shape_step, fabsf(sh.ptry[6 * k + j] - sh.p[6 * k + j]));
    }
    for (int q = 0; q < 6 * sh.N; ++q) sh.p[q] = sh.ptry[q];
    sh.lam = fmax(sh.lam * 0.1, 1e-9);
    sh.chi2 = ct;
  } else {
    sh.lam = fmin(sh.lam * 10.0, 1e6);
  }
  const double dchi2 = cur - ct;
  const bool small = accept && (dchi2 < double(a.ftol_rel) * fmax(cur, 1.0));
  sh.consec = small ? sh.consec + 1 : (accept ? 0 : sh.consec);
  const bool step_small = accept && pos_step < a.step_tol && shape_step < a.shape_step_tol;
  sh.nrej = accept ? 0 : sh.nrej + 1;
  const bool stall_now = sh.nrej >= 4;
  sh.stall = stall_now;
  sh.conv = (sh.consec >= 2) || step_small || stall_now;
  sh.nst += 1;
}

template <typename Real>
__global__ __launch_bounds__(UPH_BLOCK) void island_lm_kernel(IslandArgs a) {
  __shared__ IslandSh<Real> sh;
  const int tid = threadIdx.x;
  char* base = a.scratch + (size_t)blockIdx.x * a.scratch_per_block;
  for (;;) {
    if (tid == 0) {
      const int item = atomicAdd(a.counter, 1);
      sh.item = item;
      if (item < a.n_isl) read_shape(a, sh, a.order ? a.order[item] : item);
    }
    __syncthreads();
    const int item = sh.item;
    if (item >= a.n_isl) break;
    const int isl = a.order ? a.order[item] : item;
    const int N = sh.N, nP = 6 * N;
    const IslandScratch lay = island_scratch_layout(N, sh.B, sh.T, sh.nLc, sizeof(Real));
    Real* planes = reinterpret_cast<Real*>(base + lay.planes);
    Real* ws = reinterpret_cast<Real*>(base + lay.render);
    double* G = reinterpret_cast<double*>(base + lay.gram);
    double* part = reinterpret_cast<double*>(base + lay.part);
    const Small sm = carve_small(base + lay.small, N, sh.nLc);
    Real* gws = reinterpret_cast<Real*>(base + lay.total);  // galaxy workspace (mixed)
    const size_t m0 = (size_t)sh.moff;
    // ---- island description into LDS ----
    describe_island(a, sh, a.p0);
    if (tid == 0) {
      if (a.resume) {
        sh.lam = a.lam_out[isl];
        sh.chi2 = a.chi2_out[isl];
        sh.consec = a.consec_out[isl];
        sh.nrej = a.nrej_out[isl];
        sh.nst = a.nst_out[isl];
      } else {
        sh.lam = 1e-3;
        sh.consec = sh.nrej = sh.nst = 0;
      }
      sh.conv = sh.stall = 0;
      if (sh.Nr == 0) sh.conv = 1;  // nothing to fit (all slots padded)
    }
    __syncthreads();
    if (a.resume) {
      if (sh.Nr == 0 && tid == 0) sh.chi2 = 0.0;
    } else if (sh.Nr > 0) {
      const double c0 = chi2_all<Real>(a, sh, isl, sh.p, planes, ws, gws, G, part, sm);
      if (tid == 0) sh.chi2 = c0;
      __syncthreads();
    } else if (tid == 0) {
      sh.chi2 = 0.0;
    }
    __syncthreads();
    for (int it = 0; it < a.max_steps; ++it) {
      if (sh.conv) break;
      const int nL = sh.nL;
      // H, g at p (un-damped, un-masked over the live set)
      for (int b = 0; b < a.nb; ++b)
        eval_band<Real, false>(a, sh, isl, b, sh.p, true, b > 0, planes, ws, gws, G, part, sm, nullptr,
                        nullptr, nullptr);
      // damped float64 solve: (H + lam diag(H) + 1e-8 I) delta = -g
      const double lam = sh.lam;
      for (int e = tid; e < nL * nL; e += UPH_BLOCK) {
        const int i = e / nL, j = e % nL;
        const double v = sm.H[e];
        sm.Hd[e] = i == j ? damp_diag(v, lam) : v;
      }
      for (int j = tid; j < nL; j += UPH_BLOCK) sm.delta[j] = sm.g[j];
      __syncthreads();
      lm_propose(a, sh, sm.Hd, sm.delta, sm.pivH);
      const double ct = chi2_all<Real>(a, sh, isl, sh.ptry, planes, ws, gws, G, part, sm);
      if (tid == 0) lm_accept(a, sh, ct);
      __syncthreads();
    }
    // ---- final fluxes + outputs ----
    float* fo = a.flux_out + (size_t)sh.foff;
    float* eo = a.ferr_out + (size_t)sh.foff;
    for (int e = tid; e < a.nb * N; e += UPH_BLOCK) {
      fo[e] = 0.0f;
      eo[e] = 1.0f;
    }
    __syncthreads();
    int ngood = 0;
    if (sh.Nr > 0) {
      for (int b = 0; b < a.nb; ++b)
        eval_band<Real, false>(a, sh, isl, b, sh.p, false, false, planes, ws, gws, G, part, sm, fo + b * N,
                        eo + b * N, &ngood);
    }
    for (int q = tid; q < nP; q += UPH_BLOCK) a.p_out[m0 * 6 + q] = sh.p[q];
    if (tid == 0) {
      a.chi2_out[isl] = sh.chi2;
      a.lam_out[isl] = sh.lam;
      a.conv_out[isl] = uint8_t(sh.conv);
      a.stall_out[isl] = uint8_t(sh.stall);
      a.nst_out[isl] = sh.nst;
      a.consec_out[isl] = sh.consec;
      a.nrej_out[isl] = sh.nrej;
      a.ngood_out[isl] = ngood;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Band-split LM.  The same step as island_lm_kernel, with each evaluation
// pass spread over (island, band) items so that one island's bands run on
// many compute units at once; a per-island step launch reduces the band
// blocks in band order (the order of the persistent kernel's accumulation,
// so both paths produce the same bits) and runs the damped solve or the
// accept / reject.  Used for islands whose single-workgroup latency would
// dominate a batch.
// ---------------------------------------------------------------------------
// Argument block of the evaluation kernel: the pre-render variant also
// carries the galaxy plane buffers.
template <bool kPre>
struct EvalArgs {
  using type = IslandArgs;
};
template <>
struct EvalArgs<true> {
  using type = IslandArgsPre;
};

// Post-fit pass, one band of one island at the converged parameters
// (island_solver.py:1549-1612 final_flux_marg, :1643-1715 source_metrics):
// the flux sensitivities dF/dtheta_j = D_j = A^-1 (e_k s_j - f_k U_j) of the
// live parameters (padded slots 0) into the island's post block, and per
// real member the band's metric sums over the window: sum m (the unit
// model), sum (f m)^2, sum (model - f m) f m, sum of m on zero-weight
// pixels, sum m w (d - model)^2, and f.
template <typename Real>
__device__ void post_band(const IslandArgsPre& a, IslandSh<Real>& sh, int isl, int b,
                          const Real* planes, const Small& sm) {
  const int tid = threadIdx.x, P = sh.B * sh.B, Nr = sh.Nr, nL = sh.nL, N = sh.N;
  const Real* Wp = planes;
  const Real* Rp = planes + P;
  const Real* Dp = planes + 2 * P;
  const Real* Mp = planes + 3 * P;
  double* D = a.post_buf + a.post_off[isl] + (size_t)b * sh.nLc * N;
  for (int e = tid; e < nL * N; e += UPH_BLOCK) D[e] = 0.0;
  __syncthreads();
  for (int e = tid; e < nL * Nr; e += UPH_BLOCK) {
    const int j = e / Nr, r = e % Nr;
    D[(size_t)j * N + sh.mem[r]] = sm.D[(size_t)j * Nr + r];
  }
  for (int r = 0; r < Nr; ++r) {
    const double f = double(sh.fR[r]);
    const Real* M = Mp + (size_t)r * P;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int p = tid; p < P; p += UPH_BLOCK) {
      const double m = double(M[p]), w = double(Wp[p]), R = double(Rp[p]);
      const double own = f * m, model = double(Dp[p]) - R;
      v[0] += m;
      v[1] += own * own;
      v[2] += (model - own) * own;
      v[3] += w == 0.0 ? m : 0.0;
      v[4] += m * w * R * R;
    }
#pragma unroll
    for (int q = 0; q < 5; ++q) v[q] = block_sum(v[q], sh.redd);
    if (tid == 0) {
      float* o = a.post_metrics + ((size_t)(sh.moff + sh.mem[r]) * a.nb + b) * 6;
      for (int q = 0; q < 5; ++q) o[q] = float(v[q]);
      o[5] = float(f);
    }
  }
  __syncthreads();
}

template <typename Real, bool kPre>
__global__ __launch_bounds__(UPH_BLOCK) void island_eval_kernel(typename EvalArgs<kPre>::type a) {
  __shared__ IslandSh<Real> sh;
  const int tid = threadIdx.x;
  char* base = a.scratch + (size_t)blockIdx.x * a.scratch_per_block;
  const int mode = a.mode;
  for (;;) {
    if (tid == 0) {
      const int item = atomicAdd(a.counter, 1);
      sh.item = item;
      if (item < a.n_items) {
        const int isl = a.item_isl[item];
        read_shape(a, sh, isl);
        sh.conv = (mode == kSplitEvalGrad || mode == kSplitEvalTry) ? int(a.conv_out[isl]) : 0;
      }
    }
    __syncthreads();
    const int item = sh.item;
    if (item >= a.n_items) break;
    if (sh.conv) {
      __syncthreads();
      continue;
    }
    const int isl = a.item_isl[item], b = a.item_band[item];
    const float* psrc = mode == kSplitEvalInit ? a.p0 : (mode == kSplitEvalTry ? a.p_try : a.p_out);
    describe_island(a, sh, psrc);
    const int N = sh.N, nL = sh.nL;
    const size_t blk = split_band_block(sh.nLc);
    const size_t off = (size_t)a.hg_off[isl] + (size_t)b * blk;
    if (off + blk > a.hg_doubles || nL > sh.nLc) {
      if (tid == 0) atomicOr(a.err, 1);
      __syncthreads();
      continue;
    }
    double* slot = a.hg + off;
    if (sh.Nr == 0) {
      if (tid == 0) slot[(size_t)nL * nL + nL] = 0.0;
      __syncthreads();
      continue;
    }
    const IslandScratch lay = island_scratch_layout(N, sh.B, sh.T, sh.nLc, sizeof(Real));
    Real* planes = reinterpret_cast<Real*>(base + lay.planes);
    Real* ws = reinterpret_cast<Real*>(base + lay.render);
    double* G = reinterpret_cast<double*>(base + lay.gram);
    double* part = reinterpret_cast<double*>(base + lay.part);
    const Small sm = carve_small(base + lay.small, N, sh.nLc);
    Real* gws = reinterpret_cast<Real*>(base + lay.total);
    if (mode == kSplitEvalFinal) {
      float* fo = a.flux_out + (size_t)sh.foff + (size_t)b * N;
      float* eo = a.ferr_out + (size_t)sh.foff + (size_t)b * N;
      int ngood = 0;
      eval_band<Real, kPre>(a, sh, isl, b, sh.p, false, false, planes, ws, gws, G, part, sm, fo, eo,
                      &ngood);
      if (tid == 0) atomicAdd(a.ngood_out + isl, ngood);
    } else {
      const bool grads = split_grads(mode);
      const double c = eval_band<Real, kPre>(a, sh, isl, b, sh.p, grads, false, planes, ws, gws, G, part,
                                     sm, nullptr, nullptr, nullptr);
      if (grads) {
        for (int e = tid; e < nL * nL; e += UPH_BLOCK) slot[e] = sm.H[e];
        for (int j = tid; j < nL; j += UPH_BLOCK) slot[(size_t)nL * nL + j] = sm.g[j];
      }
      if (tid == 0) slot[(size_t)nL * nL + nL] = c;
      if constexpr (kPre && sizeof(Real) == 4) {
        if (mode == kSplitEvalPost) post_band<Real>(a, sh, isl, b, planes, sm);
      }
    }
    __syncthreads();
  }
}

// Gram tile pairs of the large islands for one evaluation pass: workgroup
// (q, g) takes island big_isl[q / nb] in band q % nb and its tile pairs
// 4 g + wave over the columns [d, member models, live tangents] (C = 1 + Nr
// + nL, no tangents on passes without them), each over all window pixels
// as one unit of gram_mfma (float32 MFMA partials flushed to float64 every
// 16 K-steps), with the columns read straight from the data stack and the
// member planes (zero outside a member's stamp box) instead of placed
// window planes.  Converged islands are skipped as the evaluation skips
// them.
constexpr int kMaxCols = 1 + kMaxMembers + kMaxLive;

__global__ __launch_bounds__(UPH_BLOCK) void island_gram_kernel(IslandArgsPre a) {
  __shared__ IslandSh<float> sh;
  __shared__ long long m_blk[kMaxMembers];  // member's band block in pbuf
  __shared__ int m_box[kMaxMembers][5];     // window box wx0, wx1, wy0, wy1; plane side
  __shared__ int c_src[kMaxCols];           // -1: data; else member * 8 + plane
  __shared__ int s_C;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int q = blockIdx.x, b = q % a.nb, isl = a.big_isl[q / a.nb];
  const int mode = a.mode;
  if ((mode == kSplitEvalGrad || mode == kSplitEvalTry) && a.conv_out[isl]) return;  // uniform
  if (tid == 0) read_shape(a, sh, isl);
  __syncthreads();
  const float* psrc = mode == kSplitEvalInit ? a.p0 : (mode == kSplitEvalTry ? a.p_try : a.p_out);
  describe_island(a, sh, psrc);
  const bool grads = split_grads(mode);
  const int Nr = sh.Nr, nL = grads ? sh.nL : 0, C = 1 + Nr + nL;
  const int nt = (C + 15) >> 4, ntp = nt * (nt + 1) / 2;
  if (4 * (int)blockIdx.y >= ntp || Nr == 0) return;  // uniform
  const int B = sh.B, T = sh.T, P = B * B;
  const float x0f = a.x0b[isl], y0f = a.y0b[isl];
  for (int r = tid; r < Nr; r += UPH_BLOCK) {
    const float* p6 = sh.p + 6 * sh.mem[r];
    const int ix = int(rint(p6[0] - x0f)) - T / 2, iy = int(rint(p6[1] - y0f)) - T / 2;
    const int side = pre_side(T, B);
    m_box[r][0] = max(0, ix);
    m_box[r][1] = min(B, ix + T);
    m_box[r][2] = max(0, iy);
    m_box[r][3] = min(B, iy + T);
    m_box[r][4] = side;
    const int g = a.pre_of_slot[(size_t)sh.moff + sh.mem[r]];
    m_blk[r] = (long long)a.pre_off[g] +
               (long long)b * pre_planes(sh.st[r] == 0) * side * side;
  }
  for (int c = tid; c < C; c += UPH_BLOCK) {
    if (c == 0) {
      c_src[c] = -1;
    } else if (c <= Nr) {
      c_src[c] = (c - 1) * 8;
    } else {
      const int l = c - 1 - Nr;
      c_src[c] = sh.live_r[l] * 8 + 1 + sh.live_j[l];
    }
  }
  __syncthreads();
  const int tp = 4 * blockIdx.y + wid;
  if (tp >= ntp) return;  // no barrier follows
  int ti, tj;
  pair_of(tp, nt, ti, tj);
  const int col = lane & 15, grp = lane >> 4;
  const int ca = ti * 16 + col, cb = tj * 16 + col;
  const int sa = ca < C ? c_src[ca] : -2, sb = cb < C ? c_src[cb] : -2;
  const int ox = a.ox[isl], oy = a.oy[isl];
  const float* drow = a.dstack + ((size_t)b * a.Hp + oy) * a.Wp + ox;
  const float* wrow = a.wstack + ((size_t)b * a.Hp + oy) * a.Wp + ox;
  // the value of column source s at window pixels px .. px + 3 (one row)
  auto column4 = [&](int s, int wr, int wc, float (&v)[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) v[t] = 0.0f;
    if (s == -2) return;
    if (s == -1) {
      const float* d = drow + (size_t)wr * a.Wp + wc;
#pragma unroll
      for (int t = 0; t < 4; ++t) v[t] = d[t];
      return;
    }
    const int r = s >> 3, pl = s & 7;
    const int* bx = m_box[r];
    if (wr < bx[2] || wr >= bx[3]) return;
    const int side = bx[4];
    const float* src = a.pbuf + m_blk[r] + (size_t)pl * side * side +
                       (size_t)(wr - bx[2]) * side - bx[0];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int c = wc + t;
      if (c >= bx[0] && c < bx[1]) v[t] = src[c];
    }
  };
  const int KS = P >> 4;
  double accd[4] = {0.0, 0.0, 0.0, 0.0};
  uph_f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
  int cnt = 0;
  for (int ks = 0; ks < KS; ++ks) {
    const int px = ks * 16 + 4 * grp;
    const int wr = px / B, wc = px % B;
    float xa[4], xb[4], wv[4];
    const float* w = wrow + (size_t)wr * a.Wp + wc;
#pragma unroll
    for (int t = 0; t < 4; ++t) wv[t] = w[t];
    column4(sa, wr, wc, xa);
    if (ti == tj) {
#pragma unroll
      for (int t = 0; t < 4; ++t) xb[t] = xa[t];
    } else {
      column4(sb, wr, wc, xb);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[0], wv[0] * xb[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[1], wv[1] * xb[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[2], wv[2] * xb[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[3], wv[3] * xb[3], acc, 0, 0, 0);
    if (++cnt == 16) {
      accd[0] += acc.x;
      accd[1] += acc.y;
      accd[2] += acc.z;
      accd[3] += acc.w;
      acc = uph_f4{0.0f, 0.0f, 0.0f, 0.0f};
      cnt = 0;
    }
  }
  accd[0] += acc.x;
  accd[1] += acc.y;
  accd[2] += acc.z;
  accd[3] += acc.w;
  double* pt = a.big_part + (((size_t)(q / a.nb) * a.nb + b) * a.big_ntp + tp) * 256;
#pragma unroll
  for (int r = 0; r < 4; ++r) pt[(4 * grp + r) * 16 + col] = accd[r];
}

// The float64 form of island_gram_kernel for the float64 band-split
// sequence: columns read as doubles from the float64 member planes (data and
// weights widened from the float32 stacks), products on
// v_mfma_f64_16x16x4f64 with the accumulation chain of gram_mfma_f64 (one
// unit per tile pair over all pixels, the pixels of a K-step in the same
// order), so the Gram equals the evaluation's own.  The float64 result tile
// comes back as D[4 r + l / 16][l % 16] in register r of lane l.
__global__ __launch_bounds__(UPH_BLOCK) void island_gram64_kernel(IslandArgsPre a) {
  __shared__ IslandSh<float> sh;
  __shared__ long long m_blk[kMaxMembers];
  __shared__ int m_box[kMaxMembers][5];
  __shared__ int c_src[kMaxCols];
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int q = blockIdx.x, b = q % a.nb, isl = a.big_isl[q / a.nb];
  const int mode = a.mode;
  if ((mode == kSplitEvalGrad || mode == kSplitEvalTry) && a.conv_out[isl]) return;  // uniform
  if (tid == 0) read_shape(a, sh, isl);
  __syncthreads();
  const float* psrc = mode == kSplitEvalInit ? a.p0 : (mode == kSplitEvalTry ? a.p_try : a.p_out);
  describe_island(a, sh, psrc);
  const bool grads = split_grads(mode);
  const int Nr = sh.Nr, nL = grads ? sh.nL : 0, C = 1 + Nr + nL;
  const int nt = (C + 15) >> 4, ntp = nt * (nt + 1) / 2;
  if (4 * (int)blockIdx.y >= ntp || Nr == 0) return;  // uniform
  const int B = sh.B, T = sh.T, P = B * B;
  const float x0f = a.x0b[isl], y0f = a.y0b[isl];
  const double* pb = a.pbuf64;
  for (int r = tid; r < Nr; r += UPH_BLOCK) {
    const float* p6 = sh.p + 6 * sh.mem[r];
    const int ix = int(rint(double(p6[0] - x0f))) - T / 2;
    const int iy = int(rint(double(p6[1] - y0f))) - T / 2;
    const int side = pre_side(T, B);
    m_box[r][0] = max(0, ix);
    m_box[r][1] = min(B, ix + T);
    m_box[r][2] = max(0, iy);
    m_box[r][3] = min(B, iy + T);
    m_box[r][4] = side;
    const int g = a.pre_of_slot[(size_t)sh.moff + sh.mem[r]];
    m_blk[r] = (long long)a.pre_off[g] +
               (long long)b * pre_planes(sh.st[r] == 0) * side * side;
  }
  for (int c = tid; c < C; c += UPH_BLOCK) {
    if (c == 0) {
      c_src[c] = -1;
    } else if (c <= Nr) {
      c_src[c] = (c - 1) * 8;
    } else {
      const int l = c - 1 - Nr;
      c_src[c] = sh.live_r[l] * 8 + 1 + sh.live_j[l];
    }
  }
  __syncthreads();
  const int tp = 4 * blockIdx.y + wid;
  if (tp >= ntp) return;  // no barrier follows
  int ti, tj;
  pair_of(tp, nt, ti, tj);
  const int col = lane & 15, grp = lane >> 4;
  const int ca = ti * 16 + col, cb = tj * 16 + col;
  const int sa = ca < C ? c_src[ca] : -2, sb = cb < C ? c_src[cb] : -2;
  const int ox = a.ox[isl], oy = a.oy[isl];
  const float* drow = a.dstack + ((size_t)b * a.Hp + oy) * a.Wp + ox;
  const float* wrow = a.wstack + ((size_t)b * a.Hp + oy) * a.Wp + ox;
  auto column4 = [&](int s, int wr, int wc, double (&v)[4]) {
#pragma unroll
    for (int t = 0; t < 4; ++t) v[t] = 0.0;
    if (s == -2) return;
    if (s == -1) {
      const float* d = drow + (size_t)wr * a.Wp + wc;
#pragma unroll
      for (int t = 0; t < 4; ++t) v[t] = double(d[t]);
      return;
    }
    const int r = s >> 3, pl = s & 7;
    const int* bx = m_box[r];
    if (wr < bx[2] || wr >= bx[3]) return;
    const int side = bx[4];
    const double* src = pb + m_blk[r] + (size_t)pl * side * side +
                        (size_t)(wr - bx[2]) * side - bx[0];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int c = wc + t;
      if (c >= bx[0] && c < bx[1]) v[t] = src[c];
    }
  };
  const int KS = P >> 4;
  uph_d4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int ks = 0; ks < KS; ++ks) {
    const int px = ks * 16 + 4 * grp;
    const int wr = px / B, wc = px % B;
    double xa[4], xb[4], wv[4];
    const float* w = wrow + (size_t)wr * a.Wp + wc;
#pragma unroll
    for (int t = 0; t < 4; ++t) wv[t] = double(w[t]);
    column4(sa, wr, wc, xa);
    if (ti == tj) {
#pragma unroll
      for (int t = 0; t < 4; ++t) xb[t] = xa[t];
    } else {
      column4(sb, wr, wc, xb);
    }
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[0], wv[0] * xb[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[1], wv[1] * xb[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[2], wv[2] * xb[2], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[3], wv[3] * xb[3], acc, 0, 0, 0);
  }
  double* pt = a.big_part + (((size_t)(q / a.nb) * a.nb + b) * a.big_ntp + tp) * 256;
  pt[(0 + grp) * 16 + col] = acc.x;
  pt[(4 + grp) * 16 + col] = acc.y;
  pt[(8 + grp) * 16 + col] = acc.z;
  pt[(12 + grp) * 16 + col] = acc.w;
}

template <typename Real>
__global__ __launch_bounds__(UPH_BLOCK) void island_step_kernel(IslandArgs a) {
  __shared__ IslandSh<Real> sh;
  __shared__ int piv[kMaxLive];
  __shared__ double rhs[kMaxLive];
  const int tid = threadIdx.x, isl = blockIdx.x, mode = a.mode;
  if (mode != kSplitStepInit && a.conv_out[isl]) return;  // uniform over the block
  if (tid == 0) read_shape(a, sh, isl);
  __syncthreads();
  describe_island(a, sh, mode == kSplitStepInit ? a.p0 : a.p_out);
  const int N = sh.N, nP = 6 * N, nL = sh.nL, nb = a.nb;
  const size_t m0 = (size_t)sh.moff;
  const size_t blk = split_band_block(sh.nLc);
  const size_t off0 = (size_t)a.hg_off[isl];
  if (off0 + (size_t)nb * blk > a.hg_doubles || nL > sh.nLc) {
    if (tid == 0) {
      atomicOr(a.err, 1);
      a.conv_out[isl] = 1;
    }
    return;
  }
  double* h0 = a.hg + off0;
  const size_t c_at = (size_t)nL * nL + nL;  // chi2 entry of a band block
  if (mode == kSplitStepInit) {
    if (tid == 0 && !a.resume) {
      double c = 0.0;
      if (sh.Nr > 0)
        for (int b = 0; b < nb; ++b) c += h0[(size_t)b * blk + c_at];
      a.chi2_out[isl] = c;
      a.lam_out[isl] = 1e-3;
      a.consec_out[isl] = 0;
      a.nrej_out[isl] = 0;
      a.nst_out[isl] = 0;
    } else if (tid == 0 && sh.Nr == 0) {
      a.chi2_out[isl] = 0.0;
    }
    if (tid == 0) {
      a.conv_out[isl] = uint8_t(sh.Nr == 0);
      a.stall_out[isl] = 0;
      a.ngood_out[isl] = 0;
    }
    for (int q = tid; q < nP; q += UPH_BLOCK) a.p_out[m0 * 6 + q] = sh.p[q];
    float* fo = a.flux_out + (size_t)sh.foff;
    float* eo = a.ferr_out + (size_t)sh.foff;
    for (int e = tid; e < nb * N; e += UPH_BLOCK) {
      fo[e] = 0.0f;
      eo[e] = 1.0f;
    }
    return;
  }
  if (tid == 0) {
    sh.chi2 = a.chi2_out[isl];
    sh.lam = a.lam_out[isl];
    sh.consec = a.consec_out[isl];
    sh.nrej = a.nrej_out[isl];
    sh.nst = a.nst_out[isl];
    sh.conv = a.conv_out[isl];
    sh.stall = a.stall_out[isl];
  }
  __syncthreads();
  if (mode == kSplitStepSolve) {
    // H = sum_b H_b and g = sum_b g_b in band order; damped in place in block 0
    const double lam = sh.lam;
    for (int e = tid; e < nL * nL; e += UPH_BLOCK) {
      double v = h0[e];
      for (int b = 1; b < nb; ++b) v = v + h0[(size_t)b * blk + e];
      const int i = e / nL, j = e % nL;
      h0[e] = i == j ? damp_diag(v, lam) : v;
    }
    for (int j = tid; j < nL; j += UPH_BLOCK) {
      const size_t o = (size_t)nL * nL + j;
      double v = h0[o];
      for (int b = 1; b < nb; ++b) v = v + h0[(size_t)b * blk + o];
      rhs[j] = v;
    }
    __syncthreads();
    lm_propose(a, sh, h0, rhs, piv);
    for (int q = tid; q < nP; q += UPH_BLOCK) a.p_try[m0 * 6 + q] = sh.ptry[q];
    return;
  }
  // kSplitStepAccept
  for (int q = tid; q < nP; q += UPH_BLOCK) sh.ptry[q] = a.p_try[m0 * 6 + q];
  __syncthreads();
  if (tid == 0) {
    double c = 0.0;
    for (int b = 0; b < nb; ++b) c += h0[(size_t)b * blk + c_at];
    lm_accept(a, sh, c);
    a.chi2_out[isl] = sh.chi2;
    a.lam_out[isl] = sh.lam;
    a.consec_out[isl] = sh.consec;
    a.nrej_out[isl] = sh.nrej;
    a.nst_out[isl] = sh.nst;
    a.conv_out[isl] = uint8_t(sh.conv);
    a.stall_out[isl] = uint8_t(sh.stall);
  }
  __syncthreads();
  for (int q = tid; q < nP; q += UPH_BLOCK) a.p_out[m0 * 6 + q] = sh.p[q];
}

// Post-fit covariance, one workgroup per island (island_solver.py:1484-1547
// final_errors without priors, :1549-1612 final_flux_marg): Sigma = inverse
// of the un-damped flux-marginalised GN Hessian summed over bands (f64 LU),
// the marginal variances of the live parameters and cov(ee1, ee2) per
// member slot (1 and 0 where a parameter is not live), and per band the
// flux variance increment diag(D Sigma D^T) of each extended member over
// the extended members' live parameters (0 for stars).
__global__ __launch_bounds__(UPH_BLOCK) void island_post_kernel(IslandArgsPre a) {
  __shared__ IslandSh<float> sh;
  __shared__ int piv[kMaxLive];
  __shared__ int s_p;
  __shared__ double scl[kMaxLive];   // 1 / sqrt(H_ll)
  __shared__ double cmax[UPH_BLOCK];
  const int tid = threadIdx.x, isl = blockIdx.x, nb = a.nb;
  if (tid == 0) read_shape(a, sh, isl);
  __syncthreads();
  describe_island(a, sh, a.p_out);
  const int N = sh.N, nL = sh.nL;
  const size_t blk = split_band_block(sh.nLc);
  const size_t off0 = (size_t)a.hg_off[isl];
  const size_t poff = (size_t)a.post_off[isl];
  if (off0 + (size_t)nb * blk > a.hg_doubles || nL > sh.nLc ||
      poff + post_block(nb, sh.nLc, N) > a.post_doubles) {
    if (tid == 0) atomicOr(a.err, 1);
    return;  // uniform
  }
  for (int e = tid; e < N * 7; e += UPH_BLOCK)
    a.post_var[(size_t)sh.moff * 7 + e] = e % 7 == 6 ? 0.0f : 1.0f;
  for (int e = tid; e < nb * N; e += UPH_BLOCK) a.post_dvar[(size_t)sh.foff + e] = 0.0f;
  if (nL == 0 || sh.Nr == 0) {
    if (tid == 0) a.post_kappa[isl] = 1.0f;
    return;  // uniform
  }
  double* h0 = a.hg + off0;
  const double* D = a.post_buf + poff;
  double* S = a.post_buf + poff + (size_t)nb * sh.nLc * N;
  __syncthreads();
  for (int e = tid; e < nL * nL; e += UPH_BLOCK) {
    double v = h0[e];
    for (int b = 1; b < nb; ++b) v = v + h0[(size_t)b * blk + e];
    h0[e] = v;
  }
  if (a.post_hg) {
    // the band sums of H and g in band order (the step launch's sums), for
    // the host's post-hoc stall probes
    double* o = a.post_hg + a.post_hg_off[isl];
    for (int j = tid; j < nL; j += UPH_BLOCK) {
      const size_t at = (size_t)nL * nL + j;
      double v = h0[at];
      for (int b = 1; b < nb; ++b) v = v + h0[(size_t)b * blk + at];
      o[(size_t)nL * nL + j] = v;
    }
  }
  __syncthreads();
  if (a.post_hg) {
    double* o = a.post_hg + a.post_hg_off[isl];
    for (int e = tid; e < nL * nL; e += UPH_BLOCK) o[e] = h0[e];
  }
  // conditioning of the Gauss-Newton system in parameter units scaled to a
  // unit diagonal: kappa_1 = |Hs|_1 |Hs^-1|_1, Hs = diag(s) H diag(s),
  // s_l = H_ll^-1/2 (how much float32 model error the LM step amplifies)
  for (int l = tid; l < nL; l += UPH_BLOCK) {
    const double d = h0[(size_t)l * nL + l];
    scl[l] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
  }
  __syncthreads();
  double cm = 0.0;
  for (int j = tid; j < nL; j += UPH_BLOCK) {
    double c = 0.0;
    for (int i = 0; i < nL; ++i) c += fabs(h0[(size_t)i * nL + j]) * scl[i];
    cm = fmax(cm, c * scl[j]);
  }
  cmax[tid] = cm;
  __syncthreads();
  block_lu(h0, nL, piv, &s_p);
  for (int j = tid; j < nL; j += UPH_BLOCK) {
    double* x = S + (size_t)j * nL;  // column j of the (symmetric) inverse
    for (int i = 0; i < nL; ++i) x[i] = i == j ? 1.0 : 0.0;
    lu_solve_one(h0, piv, nL, x);
  }
  __syncthreads();
  if (tid == 0) {
    double n1 = 0.0, n2 = 0.0;
    bool ok = true;
    for (int t = 0; t < UPH_BLOCK; ++t) n1 = fmax(n1, cmax[t]);
    for (int j = 0; j < nL; ++j) {
      if (!(scl[j] > 0.0)) ok = false;
      double c = 0.0;
      for (int i = 0; i < nL; ++i) c += fabs(S[(size_t)j * nL + i]) / (scl[i] * scl[j]);
      n2 = fmax(n2, c);
    }
    const double k = n1 * n2;
    a.post_kappa[isl] = ok && k == k ? float(fmin(k, 3.0e38)) : 3.0e38f;
  }
  for (int l = tid; l < nL; l += UPH_BLOCK) {
    const int r = sh.live_r[l], q = sh.live_j[l];
    float* o = a.post_var + ((size_t)sh.moff + sh.mem[r]) * 7;
    o[q] = float(S[(size_t)l * nL + l]);
    if (q == 3)
      for (int m = sh.first_live[r]; m < sh.first_live[r + 1]; ++m)
        if (sh.live_j[m] == 4) o[6] = float(S[(size_t)l * nL + m]);
  }
  for (int e = tid; e < nb * sh.Nr; e += UPH_BLOCK) {
    const int b = e / sh.Nr, r = e % sh.Nr;
    if (sh.st[r] == 0) continue;  // star rows carry no shape marginalisation
    const double* Db = D + (size_t)b * sh.nLc * N;
    const int k = sh.mem[r];
    double dv = 0.0;
    for (int i = 0; i < nL; ++i) {
      if (sh.st[sh.live_r[i]] == 0) continue;
      const double di = Db[(size_t)i * N + k];
      double t = 0.0;
      for (int c = 0; c < nL; ++c)
        if (sh.st[sh.live_r[c]] != 0) t += S[(size_t)i * nL + c] * Db[(size_t)c * N + k];
      dv += di * t;
    }
    a.post_dvar[(size_t)sh.foff + (size_t)b * N + k] = float(fmax(dv, 0.0));
  }
}

}  // namespace uph

extern "C" size_t uph_island_scratch_per_block(int N, int B, int T, int n_live_max,
                                               int precision) {
  const size_t rb = precision == UPH_PRECISION_FP64 ? 8 : 4;
  size_t base = uph::island_scratch_layout(N, B, T, n_live_max, rb).total;
  // n_live_max > 2N marks a class with extended members
  return base + (n_live_max > 2 * N ? uph::galaxy_scratch_bytes(T, rb) : 0);
}

// Validate and copy the inputs, outputs and options shared by every island
// entry point into `a`.
static int fill_island_args(uph::IslandArgs& a, const uph_island_inputs* in,
                            const uph_island_outputs* out, const uph_lm_options* opt,
                            int precision, void* scratch, size_t per) {
  UPH_CHECK_ARG(in->dstack && in->wstack && in->psf_geo && in->ox && in->oy && in->x0b &&
                    in->y0b && in->stype && in->n_frozen && in->free_n && in->mmask &&
                    in->keep && in->p0 && in->lower && in->upper,
                "uph_island_fit: null input pointer");
  UPH_CHECK_ARG(out->p && out->fluxes && out->ferr && out->chi2 && out->lam && out->conv &&
                    out->stall && out->nst && out->consec && out->nrej && out->n_good,
                "uph_island_fit: null output pointer");
  UPH_CHECK_ARG(opt->max_steps >= 0, "uph_island_fit: max_steps=%d", opt->max_steps);
  UPH_CHECK_ARG(precision == UPH_PRECISION_FP32 || precision == UPH_PRECISION_FP64,
                "uph_island_fit: precision=%d", precision);
  UPH_CHECK_ARG(scratch != nullptr, "uph_island_fit: null scratch");
  a.max_steps = opt->max_steps;
  a.resume = opt->resume;
  a.reg = opt->reg;
  a.pos_step_max = opt->pos_step_max;
  a.shape_step_max = opt->shape_step_max;
  a.un_step_max = opt->un_step_max;
  a.ftol_rel = opt->ftol_rel;
  a.step_tol = opt->step_tol;
  a.shape_step_tol = opt->shape_step_tol;
  a.dstack = in->dstack;
  a.wstack = in->wstack;
  a.psf_geo = in->psf_geo;
  a.sersic = in->sersic;
  a.ox = in->ox;
  a.oy = in->oy;
  a.x0b = in->x0b;
  a.y0b = in->y0b;
  a.stype = in->stype;
  a.n_frozen = in->n_frozen;
  a.free_n = in->free_n;
  a.mmask = in->mmask;
  a.keep = in->keep;
  a.p0 = in->p0;
  a.lower = in->lower;
  a.upper = in->upper;
  a.order = in->order;
  a.p_out = out->p;
  a.flux_out = out->fluxes;
  a.ferr_out = out->ferr;
  a.chi2_out = out->chi2;
  a.lam_out = out->lam;
  a.conv_out = out->conv;
  a.stall_out = out->stall;
  a.nst_out = out->nst;
  a.consec_out = out->consec;
  a.nrej_out = out->nrej;
  a.ngood_out = out->n_good;
  a.scratch = static_cast<char*>(scratch);
  a.scratch_per_block = per;
  return UPH_OK;
}

static int island_launch(uph::IslandArgs& a, const uph_island_inputs* in,
                         const uph_island_outputs* out, const uph_lm_options* opt, int precision,
                         void* scratch, size_t scratch_bytes, size_t per, int32_t* counter,
                         void* stream) {
  int st0 = fill_island_args(a, in, out, opt, precision, scratch, per);
  if (st0 != UPH_OK) return st0;
  UPH_CHECK_ARG(counter != nullptr, "uph_island_fit: null counter");
  if (a.n_isl <= 0) return UPH_OK;
  const size_t nblk_s = scratch_bytes / per;
  UPH_CHECK_ARG(nblk_s >= 1, "uph_island_fit: scratch of %zu bytes < one workgroup (%zu)",
                scratch_bytes, per);
  const int nblk = (int)(nblk_s < (size_t)a.n_isl ? nblk_s : (size_t)a.n_isl);
  a.counter = counter;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = uph::hip_status(hipMemsetAsync(counter, 0, sizeof(int32_t), s), "uph_island_fit memset");
  if (st != UPH_OK) return st;
  if (precision == UPH_PRECISION_FP32)
    hipLaunchKernelGGL(uph::island_lm_kernel<float>, dim3(nblk), dim3(UPH_BLOCK), 0, s, a);
  else
    hipLaunchKernelGGL(uph::island_lm_kernel<double>, dim3(nblk), dim3(UPH_BLOCK), 0, s, a);
  return uph::hip_status(hipGetLastError(), "uph_island_fit launch");
}

static bool valid_terms(int nt) {
  return nt == 1 || nt == 3 || nt == 6 || nt == 10 || nt == 15 || nt == 21;
}

extern "C" int uph_island_fit(const uph_island_class* cls, const uph_island_inputs* in,
                              const uph_island_outputs* out, const uph_lm_options* opt,
                              int precision, void* scratch, size_t scratch_bytes,
                              int32_t* counter, void* stream) {
  UPH_CHECK_ARG(cls && in && out && opt, "uph_island_fit: null argument");
  UPH_CHECK_ARG(cls->N >= 1 && cls->N <= uph::kMaxMembers, "uph_island_fit: N=%d outside 1..%d",
                cls->N, uph::kMaxMembers);
  UPH_CHECK_ARG(cls->B >= 16 && cls->B % 16 == 0, "uph_island_fit: B=%d must be a multiple of 16",
                cls->B);
  UPH_CHECK_ARG(cls->T >= 7 && cls->T % 2 == 1, "uph_island_fit: T=%d must be odd >= 7", cls->T);
  UPH_CHECK_ARG(valid_terms(cls->n_terms), "uph_island_fit: unsupported n_terms=%d",
                cls->n_terms);
  UPH_CHECK_ARG(cls->n_bands >= 1, "uph_island_fit: n_bands=%d", cls->n_bands);
  UPH_CHECK_ARG(cls->n_live_max >= 1 && cls->n_live_max <= 6 * cls->N,
                "uph_island_fit: n_live_max=%d outside 1..6N", cls->n_live_max);
  UPH_CHECK_ARG(cls->Hp >= cls->B && cls->Wp >= cls->B, "uph_island_fit: stack %dx%d < B=%d",
                cls->Hp, cls->Wp, cls->B);
  UPH_CHECK_ARG(!cls->mixed || (in->sersic != nullptr && cls->n_live_max > 2 * cls->N),
                "uph_island_fit: mixed classes need the Sersic table and n_live_max > 2N");
  UPH_CHECK_ARG(in->psf_coefs != nullptr, "uph_island_fit: null psf_coefs");
  uph::IslandArgs a{};
  a.n_isl = cls->n_isl;
  a.N = cls->N;
  a.B = cls->B;
  a.T = cls->T;
  a.nt = cls->n_terms;
  a.nb = cls->n_bands;
  a.nL_max = cls->n_live_max;
  a.Hp = cls->Hp;
  a.Wp = cls->Wp;
  a.psf_coefs = in->psf_coefs;
  const size_t per = uph_island_scratch_per_block(cls->N, cls->B, cls->T, cls->n_live_max,
                                                  precision);
  return island_launch(a, in, out, opt, precision, scratch, scratch_bytes, per, counter, stream);
}

// Heterogeneous batch description -> `a` (shape arrays, tiers, sizes).
static int fill_batch_args(uph::IslandArgs& a, const uph_island_batch* bat,
                           const uph_island_inputs* in, size_t scratch_per_block) {
  UPH_CHECK_ARG(bat->shape && bat->mem_off && bat->flux_off,
                "uph_island_fit_batch: null shape/offset arrays");
  UPH_CHECK_ARG(bat->n_tiers >= 1 && bat->n_tiers <= 3, "uph_island_fit_batch: n_tiers=%d",
                bat->n_tiers);
  UPH_CHECK_ARG(valid_terms(bat->n_terms), "uph_island_fit_batch: unsupported n_terms=%d",
                bat->n_terms);
  UPH_CHECK_ARG(bat->n_bands >= 1, "uph_island_fit_batch: n_bands=%d", bat->n_bands);
  UPH_CHECK_ARG(bat->max_B >= 16 && bat->Hp >= bat->max_B && bat->Wp >= bat->max_B,
                "uph_island_fit_batch: stack %dx%d smaller than max_B=%d", bat->Hp, bat->Wp,
                bat->max_B);
  UPH_CHECK_ARG(!bat->mixed || in->sersic != nullptr,
                "uph_island_fit_batch: mixed islands need the Sersic table");
  for (int t = 0; t < bat->n_tiers; ++t) {
    UPH_CHECK_ARG(bat->tier_T[t] >= 7 && bat->tier_T[t] % 2 == 1 && bat->psf_coefs[t],
                  "uph_island_fit_batch: tier %d: T=%d / coefs", t, bat->tier_T[t]);
    a.tier_coefs[t] = bat->psf_coefs[t];
    a.tier_T[t] = bat->tier_T[t];
  }
  UPH_CHECK_ARG(scratch_per_block >= 256 && scratch_per_block % 256 == 0,
                "uph_island_fit_batch: scratch_per_block=%zu", scratch_per_block);
  a.n_isl = bat->n_isl;
  a.nt = bat->n_terms;
  a.nb = bat->n_bands;
  a.Hp = bat->Hp;
  a.Wp = bat->Wp;
  a.shape = bat->shape;
  a.mem_off = bat->mem_off;
  a.flux_off = bat->flux_off;
  return UPH_OK;
}

extern "C" int uph_island_fit_batch(const uph_island_batch* bat, const uph_island_inputs* in,
                                    const uph_island_outputs* out, const uph_lm_options* opt,
                                    int precision, void* scratch, size_t scratch_bytes,
                                    size_t scratch_per_block, int32_t* counter, void* stream) {
  UPH_CHECK_ARG(bat && in && out && opt, "uph_island_fit_batch: null argument");
  UPH_CHECK_ARG(bat->shape && bat->mem_off && bat->flux_off,
                "uph_island_fit_batch: null shape/offset arrays");
  uph::IslandArgs a{};
  int st = fill_batch_args(a, bat, in, scratch_per_block);
  if (st != UPH_OK) return st;
  return island_launch(a, in, out, opt, precision, scratch, scratch_bytes, scratch_per_block,
                       counter, stream);
}

namespace uph {
// dst[0, n) = v, one workgroup.
__global__ __launch_bounds__(UPH_BLOCK) void fill_i32_kernel(int32_t* dst, int n, int v) {
  for (int i = threadIdx.x; i < n; i += UPH_BLOCK) dst[i] = v;
}
}  // namespace uph

template <typename Real>
static void split_eval(uph::IslandArgs a, int mode, int* counter, int nblk, hipStream_t s) {
  a.mode = mode;
  a.counter = counter;
  hipLaunchKernelGGL((uph::island_eval_kernel<Real, false>), dim3(nblk), dim3(UPH_BLOCK), 0, s, a);
}

template <typename Real>
static void split_eval_pre(uph::IslandArgsPre a, int mode, int* counter, int nblk,
                           hipStream_t s) {
  a.mode = mode;
  a.counter = counter;
  hipLaunchKernelGGL((uph::island_eval_kernel<Real, true>), dim3(nblk), dim3(UPH_BLOCK), 0, s,
                     a);
}

template <typename Real>
static void split_step(uph::IslandArgs a, int mode, hipStream_t s) {
  a.mode = mode;
  hipLaunchKernelGGL(uph::island_step_kernel<Real>, dim3(a.n_isl), dim3(UPH_BLOCK), 0, s, a);
}

// Evaluation pass k of the sequence: the member pre-render (when `pre` is
// given), the large-island Gram launch (when it has islands), then the
// (island, band) evaluation, on work-queue counters 3k (evaluation), 3k + 1
// (member planes) and 3k + 2 (float64 planes of the larger stamp tiers); the
// caller's timing events 6k (start),
// 6k + 1 (after the taps/spectra launch), 6k + 2 (after the extended-member
// planes), 6k + 3 (after the star planes), 6k + 4 (after the large-island
// Gram, before the evaluation) and 6k + 5 (end) when given.
template <typename Real>
static void split_pass(const uph::IslandArgs& a, const uph::IslandArgsPre* pre, int mode,
                       int* counters, int nblk, hipStream_t s, void* const* ev, int& k) {
  auto evk = [&](int q) { return ev ? static_cast<hipEvent_t>(ev[6 * k + q]) : nullptr; };
  auto rec = [&](int q) {
    if (ev) (void)hipEventRecord(evk(q), s);
  };
  rec(0);
  if (pre && sizeof(Real) == 8) {
    uph::prerender64_pass(*pre, mode, counters + 3 * k + 1, nblk, s, evk(1), evk(2));
  } else if (pre) {
    uph::prerender_pass(*pre, mode, counters + 3 * k + 1, nblk, s, evk(1), evk(2));
  } else {
    rec(1);
    rec(2);
  }
  rec(3);
  if (sizeof(Real) == 4 && pre && pre->big_n > 0 && !pre->gram_bf16) {
    uph::IslandArgsPre g = *pre;
    g.mode = mode;
    hipLaunchKernelGGL(uph::island_gram_kernel, dim3(g.big_n * g.nb, (g.big_ntp + 3) / 4),
                       dim3(UPH_BLOCK), 0, s, g);
  } else if (sizeof(Real) == 8 && pre && pre->big_n > 0) {
    uph::IslandArgsPre g = *pre;
    g.mode = mode;
    hipLaunchKernelGGL(uph::island_gram64_kernel, dim3(g.big_n * g.nb, (g.big_ntp + 3) / 4),
                       dim3(UPH_BLOCK), 0, s, g);
  }
  rec(4);
  if (pre)
    split_eval_pre<Real>(*pre, mode, counters + 3 * k, nblk, s);
  else
    split_eval<Real>(a, mode, counters + 3 * k, nblk, s);
  rec(5);
  ++k;
}

template <typename Real>
static void split_sequence(const uph::IslandArgs& a, const uph::IslandArgsPre* pre,
                           int* counters, int nblk, hipStream_t s, void* const* ev) {
  int k = 0;
  split_pass<Real>(a, pre, uph::kSplitEvalInit, counters, nblk, s, ev, k);
  split_step<Real>(a, uph::kSplitStepInit, s);
  for (int it = 0; it < a.max_steps; ++it) {
    split_pass<Real>(a, pre, uph::kSplitEvalGrad, counters, nblk, s, ev, k);
    split_step<Real>(a, uph::kSplitStepSolve, s);
    split_pass<Real>(a, pre, uph::kSplitEvalTry, counters, nblk, s, ev, k);
    split_step<Real>(a, uph::kSplitStepAccept, s);
  }
  split_pass<Real>(a, pre, uph::kSplitEvalFinal, counters, nblk, s, ev, k);
  if (pre && pre->post) {
    split_pass<Real>(a, pre, uph::kSplitEvalPost, counters, nblk, s, ev, k);
    hipLaunchKernelGGL(uph::island_post_kernel, dim3(a.n_isl), dim3(UPH_BLOCK), 0, s, *pre);
  }
}

extern "C" size_t uph_island_split_counters(int max_steps) {
  // evaluation passes: initial, two per step, final, post-fit
  const size_t passes = (size_t)(2 * (max_steps > 0 ? max_steps : 0) + 3);
  // per pass: evaluation queue, planes queue, large-tier planes queue; then
  // the error flag
  return 3 * passes + 1;
}

extern "C" int uph_island_fit_split(const uph_island_batch* bat, const uph_island_inputs* in,
                                    const uph_island_outputs* out, const uph_lm_options* opt,
                                    int precision, const uph_island_split* sp, void* scratch,
                                    size_t scratch_bytes, size_t scratch_per_block,
                                    void* stream) {
  UPH_CHECK_ARG(bat && in && out && opt && sp, "uph_island_fit_split: null argument");
  UPH_CHECK_ARG(bat->shape && bat->mem_off && bat->flux_off,
                "uph_island_fit_split: null shape/offset arrays");
  UPH_CHECK_ARG(sp->n_items >= 0 && (sp->n_items == 0 || (sp->item_isl && sp->item_band)),
                "uph_island_fit_split: bad item arrays");
  UPH_CHECK_ARG(sp->hg_off && sp->hg && sp->p_try && sp->counters,
                "uph_island_fit_split: null work arrays");
  uph::IslandArgs a{};
  int st = fill_batch_args(a, bat, in, scratch_per_block);
  if (st != UPH_OK) return st;
  st = fill_island_args(a, in, out, opt, precision, scratch, scratch_per_block);
  if (st != UPH_OK) return st;
  if (a.n_isl <= 0) return UPH_OK;
  UPH_CHECK_ARG(sp->n_items == a.n_isl * a.nb,
                "uph_island_fit_split: n_items=%d must be n_isl * n_bands = %d", sp->n_items,
                a.n_isl * a.nb);
  const size_t nblk_s = scratch_bytes / scratch_per_block;
  UPH_CHECK_ARG(nblk_s >= 1, "uph_island_fit_split: scratch of %zu bytes < one workgroup (%zu)",
                scratch_bytes, scratch_per_block);
  const int nblk = (int)(nblk_s < (size_t)sp->n_items ? nblk_s : (size_t)sp->n_items);
  a.n_items = sp->n_items;
  a.item_isl = sp->item_isl;
  a.item_band = sp->item_band;
  a.hg_off = sp->hg_off;
  a.hg = sp->hg;
  a.hg_doubles = sp->hg_doubles;
  a.p_try = sp->p_try;
  const size_t nc = uph_island_split_counters(a.max_steps);
  a.err = sp->counters + nc - 1;
  uph::IslandArgsPre pre{};
  const bool use_pre = sp->n_pre > 0;
  UPH_CHECK_ARG(!sp->post || use_pre,
                "uph_island_fit_split: the post-fit pass needs the member pre-render");
  UPH_CHECK_ARG(!sp->gram_bf16 || use_pre,
                "uph_island_fit_split: the bfloat16 Gram needs the member pre-render");
  const bool fp64 = precision == UPH_PRECISION_FP64;
  if (use_pre) {
    UPH_CHECK_ARG(sp->pre_isl && sp->pre_slot && sp->pre_off && sp->pre_spec &&
                      sp->pre_of_slot && sp->pbuf && sp->fspec && (fp64 || (sp->gdt && sp->gtap)),
                  "uph_island_fit_split: null pre-render arrays");
    UPH_CHECK_ARG(sp->n_pre_ext >= 0 && sp->n_pre_ext <= sp->n_pre,
                  "uph_island_fit_split: n_pre_ext=%d outside [0, n_pre=%d]", sp->n_pre_ext,
                  sp->n_pre);
    UPH_CHECK_ARG(!fp64 || (!sp->post && !sp->gram_bf16),
                  "uph_island_fit_split: the float64 pre-render takes no post-fit pass or "
                  "bfloat16 Gram");
    for (int t = 0; t < bat->n_tiers; ++t) {
      if (fp64)
        UPH_CHECK_ARG(uph::prerender64_scratch(bat->tier_T[t]) <= scratch_per_block,
                      "uph_island_fit_split: stamp tier %d needs %zu bytes of float64 "
                      "pre-render scratch per workgroup (%zu given)",
                      bat->tier_T[t], uph::prerender64_scratch(bat->tier_T[t]),
                      scratch_per_block);
      else
        UPH_CHECK_ARG(uph::prerender_supports(bat->tier_T[t]) &&
                          uph::prerender_scratch(bat->tier_T[t]) <= scratch_per_block,
                      "uph_island_fit_split: stamp tier %d not supported by the member "
                      "pre-render",
                      bat->tier_T[t]);
    }
    static_cast<uph::IslandArgs&>(pre) = a;
    pre.n_pre = sp->n_pre;
    pre.n_pre_ext = sp->n_pre_ext;
    pre.pre_isl = sp->pre_isl;
    pre.pre_slot = sp->pre_slot;
    pre.pre_off = sp->pre_off;
    pre.pre_spec = sp->pre_spec;
    pre.pre_of_slot = sp->pre_of_slot;
    if (fp64) {
      pre.pbuf64 = reinterpret_cast<double*>(sp->pbuf);
      pre.fspec64 = reinterpret_cast<double*>(sp->fspec);
    } else {
      pre.pbuf = sp->pbuf;
      pre.fspec = sp->fspec;
    }
    pre.pbuf_elems = sp->pbuf_elems;
    pre.fspec_elems = sp->fspec_elems;
    pre.gdt = sp->gdt;
    pre.gtap = sp->gtap;
    pre.post = sp->post;
    pre.gram_bf16 = sp->gram_bf16;
    pre.gram_unsettled = sp->gram_unsettled;
    UPH_CHECK_ARG(sp->big_n >= 0 && (sp->big_n == 0 || (sp->big_slot && sp->big_isl &&
                                                          sp->big_part && sp->big_ntp >= 1)),
                  "uph_island_fit_split: big_n=%d needs big_slot, big_isl, big_part and "
                  "big_ntp >= 1",
                  sp->big_n);
    pre.big_n = sp->big_n;
    pre.big_ntp = sp->big_ntp;
    pre.big_slot = sp->big_slot;
    pre.big_isl = sp->big_isl;
    pre.big_part = sp->big_part;
    UPH_CHECK_ARG(!sp->gram_bf16 || sp->gram_unsettled,
                  "uph_island_fit_split: the bfloat16 Gram needs gram_unsettled");
    if (sp->post) {
      UPH_CHECK_ARG(sp->post_off && sp->post_buf && sp->post_var && sp->post_dvar &&
                        sp->post_metrics && sp->post_kappa,
                    "uph_island_fit_split: null post-fit arrays");
      pre.post_off = sp->post_off;
      pre.post_buf = sp->post_buf;
      pre.post_doubles = sp->post_doubles;
      pre.post_var = sp->post_var;
      pre.post_dvar = sp->post_dvar;
      pre.post_metrics = sp->post_metrics;
      pre.post_kappa = sp->post_kappa;
      pre.post_hg = sp->post_hg;
      pre.post_hg_off = sp->post_hg_off;
    }
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  // counters cleared by a kernel, not a memset node, so that the whole
  // sequence can be captured into a graph and replayed
  hipLaunchKernelGGL(uph::fill_i32_kernel, dim3(1), dim3(UPH_BLOCK), 0, s, sp->counters,
                     (int)nc, 0);
  st = uph::hip_status(hipGetLastError(), "uph_island_fit_split counter reset");
  if (st != UPH_OK) return st;
  if (precision == UPH_PRECISION_FP32)
    split_sequence<float>(a, use_pre ? &pre : nullptr, sp->counters, nblk, s, sp->eval_events);
  else
    split_sequence<double>(a, use_pre ? &pre : nullptr, sp->counters, nblk, s, sp->eval_events);
  return uph::hip_status(hipGetLastError(), "uph_island_fit_split launch");
}
