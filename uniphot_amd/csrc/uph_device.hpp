// Device-side building blocks shared by the island and grid kernels:
// polynomial-PSF basis, Lanczos-3 taps, block reductions and a small
// block-cooperative float64 LU.  gfx950 only (64-wide wavefronts).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define UPH_BLOCK 256
#define UPH_WAVES (UPH_BLOCK / 64)

namespace uph {

// Polynomial term exponents (m, n) of X^m Y^n for orders 0..5, low total
// degree first and, within a degree, decreasing power of X.  Same term
// order as tract7dt/fit/grid_gpu/bg_polynomial.py:40-100.
__constant__ const int8_t kTermM[21] = {0, 1, 0, 2, 1, 0, 3, 2, 1, 0, 4, 3, 2, 1, 0, 5, 4, 3, 2, 1, 0};
__constant__ const int8_t kTermN[21] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0, 1, 2, 3, 4, 0, 1, 2, 3, 4, 5};

template <typename Real>
__device__ __forceinline__ Real ipow(Real v, int e) {
  Real r = Real(1);
  for (int i = 0; i < e; ++i) r *= v;
  return r;
}

// Basis values and d/dX, d/dY of the nt-term spatial polynomial.
template <typename Real>
__device__ __forceinline__ void psf_basis(Real X, Real Y, int nt, Real* b, Real* bx, Real* by) {
  Real px[6], py[6];
  px[0] = Real(1);
  py[0] = Real(1);
#pragma unroll
  for (int i = 1; i < 6; ++i) {
    px[i] = px[i - 1] * X;
    py[i] = py[i - 1] * Y;
  }
  for (int t = 0; t < nt; ++t) {
    int m = kTermM[t], n = kTermN[t];
    b[t] = px[m] * py[n];
    if (bx) bx[t] = m > 0 ? Real(m) * px[m - 1] * py[n] : Real(0);
    if (by) by[t] = n > 0 ? Real(n) * px[m] * py[n - 1] : Real(0);
  }
}

// The nt-term polynomial PSF at one stamp pixel (coefficients cp[t * TT],
// t < nt, float) and, with `grad`, its X / Y derivative sums over the basis
// b[t], b[21 + t], b[42 + t].  The coefficient loads go out in groups of 8
// before their FMAs, which run in term order.
template <typename Real>
__device__ __forceinline__ void psf_poly_at(const float* cp, size_t TT, int nt, const Real* b,
                                            bool grad, Real& s, Real& sx, Real& sy) {
  s = Real(0);
  sx = Real(0);
  sy = Real(0);
  for (int t0 = 0; t0 < nt; t0 += 8) {
    float cv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) cv[u] = t0 + u < nt ? cp[(size_t)(t0 + u) * TT] : 0.0f;
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (t0 + u >= nt) break;
      const Real c = Real(cv[u]);
      s = fma(c, b[t0 + u], s);
      if (grad) {
        sx = fma(c, b[21 + t0 + u], sx);
        sy = fma(c, b[42 + t0 + u], sy);
      }
    }
  }
}

// Lanczos-3 window and its derivative at x, float64
// (tract7dt/fit/optimize/gpu/kernels/lanczos.py:50-99).
__device__ __forceinline__ void l3_val_der(double x, double& v, double& d) {
  double ax = fabs(x);
  if (ax >= 3.0) {
    v = 0.0;
    d = 0.0;
    return;
  }
  if (ax < 1e-12) {
    v = 1.0;
    d = 0.0;
    return;
  }
  const double pi = 3.141592653589793;
  double px = pi * x;
  double s1, c1, s3, c3;
  sincos(px, &s1, &c1);
  sincos(px / 3.0, &s3, &c3);
  double inv = 1.0 / (px * px);
  v = 3.0 * s1 * s3 * inv;
  d = (3.0 * pi * c1 * s3 + pi * s1 * c3) * inv - 2.0 * v / x;
}

// Normalised 7-tap filter at offsets k-3+s, k = 0..6, and its derivative in
// s.  Computed in float64 and rounded to Real, like the reference (taps in
// float64, cast to float32).
template <typename Real>
__device__ __forceinline__ void lanczos_taps(Real s, Real* taps, Real* dtaps) {
  double v[7], dv[7], sv = 0.0, sd = 0.0;
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    l3_val_der(double(k - 3) + double(s), v[k], dv[k]);
    sv += v[k];
    sd += dv[k];
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    taps[k] = Real(v[k] / sv);
    if (dtaps) dtaps[k] = Real((dv[k] * sv - v[k] * sd) / (sv * sv));
  }
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum of one value per thread across the block; every thread gets the
// result.  `red` must hold UPH_WAVES entries.  Fixed reduction order.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  T s = red[0];
#pragma unroll
  for (int i = 1; i < UPH_WAVES; ++i) s += red[i];
  return s;
}

// In-place LU with partial pivoting (first maximal |pivot|, as LAPACK getrf)
// of a row-major n x n float64 matrix, block-cooperative.
__device__ inline void block_lu(double* A, int n, int* piv, int* s_p) {
  const int tid = threadIdx.x;
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      int p = k;
      double best = fabs(A[k * n + k]);
      for (int i = k + 1; i < n; ++i) {
        double a = fabs(A[i * n + k]);
        if (a > best) {
          best = a;
          p = i;
        }
      }
      piv[k] = p;
      *s_p = p;
    }
    __syncthreads();
    int p = *s_p;
    if (p != k) {
      for (int j = tid; j < n; j += UPH_BLOCK) {
        double t = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
      __syncthreads();
    }
    double akk = A[k * n + k];
    for (int i = k + 1 + tid; i < n; i += UPH_BLOCK) A[i * n + k] /= akk;
    __syncthreads();
    int m = n - k - 1;
    for (int e = tid; e < m * m; e += UPH_BLOCK) {
      int i = k + 1 + e / m, j = k + 1 + e % m;
      A[i * n + j] -= A[i * n + k] * A[k * n + j];
    }
    __syncthreads();
  }
}

// block_lu with the pivot search spread over wave 0 (each lane scans rows
// k + lane, k + lane + 64, ..., then a shuffle reduction): the same pivots
// (largest |A[i][k]|, the smallest i among ties, the diagonal when it is
// NaN) and the same arithmetic, without one thread walking the column.
__device__ inline void block_lu_wpiv(double* A, int n, int* piv, int* s_p) {
  const int tid = threadIdx.x;
  for (int k = 0; k < n; ++k) {
    if (tid < 64) {
      // lane 0 starts from the diagonal as the one-thread scan does (a NaN
      // there stays the pivot); the other lanes from "no candidate"
      double best = tid == 0 ? fabs(A[k * n + k]) : -1.0;
      int p = tid == 0 ? k : n;
      for (int i = k + tid + (tid == 0 ? 64 : 0); i < n; i += 64) {
        const double v = fabs(A[i * n + k]);
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
    int p = *s_p;
    if (p != k) {
      for (int j = tid; j < n; j += UPH_BLOCK) {
        double t = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
      __syncthreads();
    }
    double akk = A[k * n + k];
    for (int i = k + 1 + tid; i < n; i += UPH_BLOCK) A[i * n + k] /= akk;
    __syncthreads();
    int m = n - k - 1;
    for (int e = tid; e < m * m; e += UPH_BLOCK) {
      int i = k + 1 + e / m, j = k + 1 + e % m;
      A[i * n + j] -= A[i * n + k] * A[k * n + j];
    }
    __syncthreads();
  }
}

// Solve LU x = P b for one right-hand side (sequential, one thread).
__device__ inline void lu_solve_one(const double* LU, const int* piv, int n, double* x) {
  for (int k = 0; k < n; ++k) {
    int p = piv[k];
    if (p != k) {
      double t = x[k];
      x[k] = x[p];
      x[p] = t;
    }
  }
  for (int i = 0; i < n; ++i) {
    double s = x[i];
    for (int j = 0; j < i; ++j) s -= LU[i * n + j] * x[j];
    x[i] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    for (int j = i + 1; j < n; ++j) s -= LU[i * n + j] * x[j];
    x[i] = s / LU[i * n + i];
  }
}

// lu_solve_one on a vector stored with stride st (x[k] at x[k * st]): the
// same operations in the same order.  Columns of a row-major block solved
// one per thread this way read consecutive addresses across the wave.
__device__ inline void lu_solve_strided(const double* LU, const int* piv, int n, double* x,
                                        int st) {
  for (int k = 0; k < n; ++k) {
    int p = piv[k];
    if (p != k) {
      double t = x[(size_t)k * st];
      x[(size_t)k * st] = x[(size_t)p * st];
      x[(size_t)p * st] = t;
    }
  }
  for (int i = 0; i < n; ++i) {
    double s = x[(size_t)i * st];
    for (int j = 0; j < i; ++j) s -= LU[i * n + j] * x[(size_t)j * st];
    x[(size_t)i * st] = s;
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = x[(size_t)i * st];
    for (int j = i + 1; j < n; ++j) s -= LU[i * n + j] * x[(size_t)j * st];
    x[(size_t)i * st] = s / LU[i * n + i];
  }
}

// Diagonal of A^-1 from its LU factors: column i of the inverse solved with
// e_i, one column per thread; `col` is per-thread workspace (n doubles each).
__device__ inline void lu_inverse_diag(const double* LU, const int* piv, int n, double* diag,
                                       double* col) {
  for (int i = threadIdx.x; i < n; i += UPH_BLOCK) {
    double* x = col + (size_t)threadIdx.x * n;
    for (int j = 0; j < n; ++j) x[j] = 0.0;
    x[i] = 1.0;
    lu_solve_one(LU, piv, n, x);
    diag[i] = x[i];
  }
}

}  // namespace uph
