// CPU restatement of the per-island VarPro-LM and the per grid-band linear
// solve, in float32 pixel arithmetic with float64 solves (TEST
// INFRASTRUCTURE: the CPU baseline timed beside the GPU path; only bench.py's
// cpu_baseline leg and tests/ load it).  OpenMP over islands and over
// grid-band problems.
//
// It follows the float64 NumPy oracle (oracle/island.py, oracle/core.py,
// oracle/linear.py) step for step -- the same member renders, explicit
// whitened-residual Jacobian through the closed-form VarPro flux solve,
// legacy LM step with its clips, bounds and three convergence prongs -- which
// restate tract7dt/fit/grid_gpu/island_solver.py:430-1002 (models_band,
// _fluxes_band, chi2_of, hg_scan, one_step), :1467-1482 (final_fluxes),
// sersic_mog.py:87-207 (table lookup, ellipse terms, FFT render),
// lanczos.py:50-181, psf_polynomial.py:134-197 and forced_phot_sparse.py:
// 342-554 (solver_inner).  Transforms are this file's own radix-2 complex
// FFT (two real rows per complex transform).  The test suite checks it
// against the float64 oracle on small islands and grid problems.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// Pixel arithmetic type: float32 for the baseline; a float64 build of the
// same source (-DUPHC_REAL=double) checks the restatement against the
// float64 oracle.
#ifndef UPHC_REAL
#define UPHC_REAL float
#endif

namespace {

using real = UPHC_REAL;
using cf = std::complex<real>;
constexpr double kPi = 3.141592653589793;
constexpr double kLam0 = 1e-3, kUp = 10.0, kDown = 0.1;
constexpr int kStallK = 4;
const int kTermM[21] = {0, 1, 0, 2, 1, 0, 3, 2, 1, 0, 4, 3, 2, 1, 0, 5, 4, 3, 2, 1, 0};
const int kTermN[21] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0, 1, 2, 3, 4, 0, 1, 2, 3, 4, 5};

struct Opts {
  float reg, pos_max, shape_max, un_max, ftol_rel, step_tol, shape_tol;
  int max_steps;
};

struct Table {
  int nk, K;
  const double *kn, *amps, *lv, *core;
  double nmin, nmax;
};

struct Psf {  // one band at one stamp tier: planar (nt, T, T)
  const float* c;
  int T, nt;
  float xo, xs, yo, ys;
};

// ---------------------------------------------------------------- FFT
struct Fft {
  int n = 0;
  std::vector<cf> tw;
  std::vector<int> rev;
  void init(int nn) {
    n = nn;
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    rev.assign(n, 0);
    for (int i = 0; i < n; ++i) {
      int r = 0;
      for (int b = 0; b < lg; ++b)
        if (i & (1 << b)) r |= 1 << (lg - 1 - b);
      rev[i] = r;
    }
    tw.resize(n / 2);
    for (int k = 0; k < n / 2; ++k)
      tw[k] = cf(real(std::cos(-2.0 * kPi * k / n)), real(std::sin(-2.0 * kPi * k / n)));
  }
  // in place; inverse without the 1/n factor
  void run(cf* a, bool inv) const {
    for (int i = 0; i < n; ++i)
      if (i < rev[i]) std::swap(a[i], a[rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
      const int half = len >> 1, step = n / len;
      for (int i = 0; i < n; i += len)
        for (int j = 0; j < half; ++j) {
          cf w = tw[j * step];
          if (inv) w = std::conj(w);
          const cf u = a[i + j], v = a[i + j + half] * w;
          a[i + j] = u + v;
          a[i + j + half] = u - v;
        }
    }
  }
};

// rfft2 of a T x T image zero-padded to pad x pad: S (pad rows w) x (pad/2+1
// columns v), row-major.
void rfft2(const Fft& f, const real* img, int T, int pad, cf* S, std::vector<cf>& tmp) {
  const int hv = pad / 2 + 1;
  tmp.assign(size_t(pad), cf(0, 0));
  std::fill(S, S + size_t(pad) * hv, cf(0, 0));
  for (int r = 0; r < T; r += 2) {  // two real rows per complex transform
    for (int c = 0; c < pad; ++c) {
      const real a = c < T ? img[r * T + c] : 0.0f;
      const real b = (c < T && r + 1 < T) ? img[(r + 1) * T + c] : 0.0f;
      tmp[c] = cf(a, b);
    }
    f.run(tmp.data(), false);
    for (int k = 0; k < hv; ++k) {
      const cf z = tmp[k], zc = std::conj(tmp[(pad - k) % pad]);
      S[size_t(r) * hv + k] = real(0.5) * (z + zc);
      if (r + 1 < T) S[size_t(r + 1) * hv + k] = cf(0, -0.5f) * (z - zc);
    }
  }
  std::vector<cf> col(pad);
  for (int k = 0; k < hv; ++k) {
    for (int r = 0; r < pad; ++r) col[r] = S[size_t(r) * hv + k];
    f.run(col.data(), false);
    for (int r = 0; r < pad; ++r) S[size_t(r) * hv + k] = col[r];
  }
}

// irfft2 of X (pad x (pad/2+1)), cropped to the top-left T x T, into out
// (scaled by 1 / pad^2); X is overwritten.
void irfft2_crop(const Fft& f, cf* X, int T, int pad, real* out, std::vector<cf>& tmp) {
  const int hv = pad / 2 + 1;
  std::vector<cf> col(pad);
  for (int k = 0; k < hv; ++k) {
    for (int r = 0; r < pad; ++r) col[r] = X[size_t(r) * hv + k];
    f.run(col.data(), true);
    for (int r = 0; r < pad; ++r) X[size_t(r) * hv + k] = col[r];
  }
  tmp.assign(size_t(pad), cf(0, 0));
  const real sc = 1.0f / (real(pad) * real(pad));
  for (int r = 0; r < T; r += 2) {
    const cf* a = X + size_t(r) * hv;
    const cf* b = r + 1 < T ? X + size_t(r + 1) * hv : nullptr;
    for (int k = 0; k < pad; ++k) {
      // a real row's spectrum: the DC and Nyquist bins count by their real
      // parts (as a complex-to-real transform reads them)
      const bool lo = k < hv, edge = k == 0 || k == pad / 2;
      cf xa = lo ? a[k] : std::conj(a[pad - k]);
      cf xb = b ? (lo ? b[k] : std::conj(b[pad - k])) : cf(0, 0);
      if (edge) {
        xa = cf(xa.real(), 0.0f);
        xb = cf(xb.real(), 0.0f);
      }
      tmp[k] = xa + cf(0, 1) * xb;
    }
    f.run(tmp.data(), true);
    for (int c = 0; c < T; ++c) {
      out[r * T + c] = tmp[c].real() * sc;
      if (r + 1 < T) out[(r + 1) * T + c] = tmp[c].imag() * sc;
    }
  }
}

int fft_pad_for(int T) {
  int n = 1;
  while (n < T + T / 2) n <<= 1;
  return n;
}

// ---------------------------------------------------------------- pieces
void lanczos_taps(real s, real* taps, real* dtaps) {
  double v[7], dv[7], sv = 0, sd = 0;
  for (int k = 0; k < 7; ++k) {
    const double x = double(k - 3) + double(s), ax = std::fabs(x);
    if (ax >= 3.0) {
      v[k] = dv[k] = 0.0;
    } else if (ax < 1e-12) {
      v[k] = 1.0;
      dv[k] = 0.0;
    } else {
      const double px = kPi * x, s1 = std::sin(px), c1 = std::cos(px);
      const double s3 = std::sin(px / 3.0), c3 = std::cos(px / 3.0), inv = 1.0 / (px * px);
      v[k] = 3.0 * s1 * s3 * inv;
      dv[k] = (3.0 * kPi * c1 * s3 + kPi * s1 * c3) * inv - 2.0 * v[k] / x;
    }
    sv += v[k];
    sd += dv[k];
  }
  for (int k = 0; k < 7; ++k) {
    taps[k] = real(v[k] / sv);
    if (dtaps) dtaps[k] = real((dv[k] * sv - v[k] * sd) / (sv * sv));
  }
}

void shift_x(const real* in, int T, const real* t, real* out) {
  for (int r = 0; r < T; ++r)
    for (int c = 0; c < T; ++c) {
      real v = 0.0f;
      for (int k = 0; k < 7; ++k) {
        const int cc = c + k - 3;
        if (cc >= 0 && cc < T) v += t[k] * in[r * T + cc];
      }
      out[r * T + c] = v;
    }
}

void shift_y(const real* in, int T, const real* t, real* out) {
  for (int r = 0; r < T; ++r)
    for (int c = 0; c < T; ++c) {
      real v = 0.0f;
      for (int k = 0; k < 7; ++k) {
        const int rr = r + k - 3;
        if (rr >= 0 && rr < T) v += t[k] * in[rr * T + c];
      }
      out[r * T + c] = v;
    }
}

void mog_at(const Table& tb, double n, double* a, double* v, double& core, double* da,
            double* dv, double& dcore) {
  const double nn = std::min(std::max(n, tb.nmin), tb.nmax);
  const bool clipped = nn != n;
  int i = int(std::upper_bound(tb.kn, tb.kn + tb.nk, nn) - tb.kn) - 1;
  i = std::min(std::max(i, 0), tb.nk - 2);
  const double n0 = tb.kn[i], n1 = tb.kn[i + 1];
  const double t = n1 > n0 ? (nn - n0) / (n1 - n0) : 0.0;
  const double dt = (clipped || !(n1 > n0)) ? 0.0 : 1.0 / (n1 - n0);
  for (int k = 0; k < tb.K; ++k) {
    const double a0 = tb.amps[i * tb.K + k], a1 = tb.amps[(i + 1) * tb.K + k];
    const double l0 = tb.lv[i * tb.K + k], l1 = tb.lv[(i + 1) * tb.K + k];
    a[k] = (1 - t) * a0 + t * a1;
    v[k] = std::exp((1 - t) * l0 + t * l1);
    da[k] = dt * (a1 - a0);
    dv[k] = v[k] * dt * (l1 - l0);
  }
  core = (1 - t) * tb.core[i] + t * tb.core[i + 1];
  dcore = dt * (tb.core[i + 1] - tb.core[i]);
}

template <typename R>
void cov_terms_an(R lr, R e1, R e2, R* out) {
  const R re = std::exp(lr), ee = std::sqrt(e1 * e1 + e2 * e2), q = std::exp(-ee);
  const R e = R(1.0) - q, inv_ba = (R(1.0) + e) / q, s2 = R(1.0) / (inv_ba * inv_ba);
  const R cos2 = e1 / ee, sin2 = e2 / ee, half = re * re / R(2.0);
  out[0] = half * ((R(1.0) + s2) - cos2 * (R(1.0) - s2));
  out[1] = half * ((R(1.0) + s2) + cos2 * (R(1.0) - s2));
  out[2] = half * sin2 * (R(1.0) - s2);
}

// (cxx, cyy, cxy) of the ellipse with the reference's clips, and their
// derivatives in (logre, ee1, ee2) (complex step of the analytic form; the
// clipped cases as in the oracle).
void cov_terms(double lr, double e1, double e2, double* t, double d[3][3]) {
  if (e1 == 0.0 && e2 == 0.0) e1 = 1e-4;
  const double re = std::min(std::max(std::exp(lr), 0.1), 1000.0);
  const double ee = std::hypot(e1, e2), e = 1.0 - std::exp(-ee);
  const double inv_ba = std::min((1.0 + e) / std::max(1.0 - e, 1e-4), 1000.0);
  const double th = 0.5 * std::atan2(e2, e1), ct = std::cos(th), st = std::sin(th);
  const double a = re * ct / inv_ba, b = re * st, c = -re * st / inv_ba, dd = re * ct;
  t[0] = a * a + b * b;
  t[1] = c * c + dd * dd;
  t[2] = a * c + b * dd;
  const bool re_clip = !(0.1 < std::exp(lr) && std::exp(lr) < 1000.0);
  const bool ba_clip = !((1.0 - e) > 1e-4 && (1.0 + e) / (1.0 - e) < 1000.0);
  const double h = 1e-30;
  using cd = std::complex<double>;
  for (int k = 0; k < 3; ++k) {
    cd args[3] = {cd(lr), cd(e1), cd(e2)};
    args[k] += cd(0, h);
    cd o[3];
    cov_terms_an<cd>(args[0], args[1], args[2], o);
    for (int j = 0; j < 3; ++j) d[k][j] = o[j].imag() / h;
    if (k == 0 && re_clip) d[k][0] = d[k][1] = d[k][2] = 0.0;
    if (k > 0 && ba_clip) {
      const double s2 = 1.0 / (inv_ba * inv_ba), half = re * re / 2.0;
      double dc, ds;
      if (k == 1) {
        dc = (ee * ee - e1 * e1) / (ee * ee * ee);
        ds = -e1 * e2 / (ee * ee * ee);
      } else {
        dc = -e1 * e2 / (ee * ee * ee);
        ds = (ee * ee - e2 * e2) / (ee * ee * ee);
      }
      d[k][0] = -half * dc * (1 - s2);
      d[k][1] = half * dc * (1 - s2);
      d[k][2] = half * ds * (1 - s2);
    }
  }
}

enum { WX = 1, WY = 2, WLR = 4, WE1 = 8, WE2 = 16, WN = 32 };

struct Work {
  std::vector<real> raw, dpx, dpy, g[6], hx, tmp2;
  std::vector<cf> S, Sx, X, tmp;
  std::vector<real> F, dF[4], expo;
  Fft fft;
  int fft_n = 0;
};

// Unit-flux T x T member stamp at (x, y), shifted by the remainder of
// (pos - origin), with the derivative planes selected by `want` in grads[j]
// (j = x, y, lr, e1, e2, n).
void member_stamp(const Psf& P, const Table& tb, real x, real y, real x0, real y0, int st,
                  real lr, real e1, real e2, real nval, unsigned want, real* stamp,
                  real* const* grads, Work& w) {
  const int T = P.T, TT = T * T;
  w.raw.assign(TT, 0.0f);
  const bool wxy = want & (WX | WY);
  if (wxy) {
    w.dpx.assign(TT, 0.0f);
    w.dpy.assign(TT, 0.0f);
  }
  {
    const real X = (x - P.xo) / P.xs, Y = (y - P.yo) / P.ys;
    real px[6], py[6], b[21], bx[21], by[21];
    px[0] = py[0] = 1.0f;
    for (int i = 1; i < 6; ++i) {
      px[i] = px[i - 1] * X;
      py[i] = py[i - 1] * Y;
    }
    for (int t = 0; t < P.nt; ++t) {
      const int m = kTermM[t], n = kTermN[t];
      b[t] = px[m] * py[n];
      bx[t] = m > 0 ? real(m) * px[m - 1] * py[n] / P.xs : 0.0f;
      by[t] = n > 0 ? real(n) * px[m] * py[n - 1] / P.ys : 0.0f;
    }
    for (int t = 0; t < P.nt; ++t) {
      const float* c = P.c + size_t(t) * TT;
      for (int e = 0; e < TT; ++e) w.raw[e] += c[e] * b[t];
      if (wxy)
        for (int e = 0; e < TT; ++e) {
          w.dpx[e] += c[e] * bx[t];
          w.dpy[e] += c[e] * by[t];
        }
    }
  }
  real* gr[6];
  for (int j = 0; j < 6; ++j) {
    gr[j] = nullptr;
    if (want & (1u << j)) {
      w.g[j].assign(TT, 0.0f);
      gr[j] = w.g[j].data();
    }
  }
  std::vector<real> model;
  const real* rawv = w.raw.data();
  if (st == 0) {
    if (gr[0]) std::copy(w.dpx.begin(), w.dpx.end(), gr[0]);
    if (gr[1]) std::copy(w.dpy.begin(), w.dpy.end(), gr[1]);
  } else {
    if (e1 == 0.0f && e2 == 0.0f) e1 = 1e-4f;
    const int pad = fft_pad_for(T), hv = pad / 2 + 1;
    if (w.fft_n != pad) {
      w.fft.init(pad);
      w.fft_n = pad;
    }
    double a[32], v[32], da[32], dv[32], core, dcore;
    mog_at(tb, nval, a, v, core, da, dv, dcore);
    double t3[3], d3[3][3];
    cov_terms(lr, e1, e2, t3, d3);
    const size_t G = size_t(pad) * hv;
    w.F.assign(G, 0.0f);
    w.expo.assign(G, 0.0f);
    const bool wsh = want & (WLR | WE1 | WE2), wn = want & WN;
    for (int j = 0; j < 4; ++j)
      if ((j < 3 && wsh) || (j == 3 && wn)) w.dF[j].assign(G, 0.0f);
    for (int r = 0; r < pad; ++r) {
      const double fw = (r < (pad + 1) / 2 ? r : r - pad) / double(pad);
      for (int k = 0; k < hv; ++k) {
        const double fv = k / double(pad);
        const double q = -2.0 * kPi * kPi;
        const double ex = q * (t3[0] * fv * fv + t3[1] * fw * fw + 2.0 * t3[2] * fv * fw);
        double Fs = 0.0, dFs[3] = {0, 0, 0}, dFn = 0.0;
        double dex[3];
        for (int j = 0; j < 3; ++j)
          dex[j] = q * (d3[j][0] * fv * fv + d3[j][1] * fw * fw + 2.0 * d3[j][2] * fv * fw);
        for (int c = 0; c < tb.K; ++c) {
          const double ec = std::exp(v[c] * ex);
          Fs += a[c] * ec;
          if (wsh)
            for (int j = 0; j < 3; ++j) dFs[j] += a[c] * v[c] * ec * dex[j];
          if (wn) dFn += da[c] * ec + a[c] * ec * ex * dv[c];
        }
        const size_t o = size_t(r) * hv + k;
        w.F[o] = real(Fs);
        if (wsh)
          for (int j = 0; j < 3; ++j) w.dF[j][o] = real(dFs[j]);
        if (wn) w.dF[3][o] = real(dFn);
      }
    }
    w.S.resize(G);
    w.X.resize(G);
    rfft2(w.fft, w.raw.data(), T, pad, w.S.data(), w.tmp);
    model.assign(TT, 0.0f);
    auto inv_with = [&](const cf* Sp, const real* Fp, real* out) {
      for (size_t o = 0; o < G; ++o) w.X[o] = Sp[o] * Fp[o];
      irfft2_crop(w.fft, w.X.data(), T, pad, out, w.tmp);
    };
    inv_with(w.S.data(), w.F.data(), model.data());
    const real cf_ = real(core);
    for (int e = 0; e < TT; ++e) model[e] += cf_ * w.raw[e];
    if (gr[0] || gr[1]) {
      w.Sx.resize(G);
      for (int j = 0; j < 2; ++j) {
        if (!gr[j]) continue;
        const real* src = j == 0 ? w.dpx.data() : w.dpy.data();
        rfft2(w.fft, src, T, pad, w.Sx.data(), w.tmp);
        inv_with(w.Sx.data(), w.F.data(), gr[j]);
        for (int e = 0; e < TT; ++e) gr[j][e] += cf_ * src[e];
      }
    }
    for (int j = 0; j < 3; ++j)
      if (gr[2 + j]) inv_with(w.S.data(), w.dF[j].data(), gr[2 + j]);
    if (gr[5]) {
      inv_with(w.S.data(), w.dF[3].data(), gr[5]);
      const real dc = real(dcore);
      for (int e = 0; e < TT; ++e) gr[5][e] += dc * w.raw[e];
    }
    rawv = model.data();
  }
  const real xg = x - x0, yg = y - y0;
  const real dxs = xg - std::nearbyint(xg), dys = yg - std::nearbyint(yg);
  real tx[7], dtx[7], ty[7], dty[7];
  lanczos_taps(dxs, tx, dtx);
  lanczos_taps(dys, ty, dty);
  w.hx.resize(TT);
  w.tmp2.resize(TT);
  shift_x(rawv, T, tx, w.hx.data());
  shift_y(w.hx.data(), T, ty, stamp);
  for (int j = 0; j < 6; ++j) {
    if (!gr[j]) continue;
    shift_x(gr[j], T, tx, w.tmp2.data());
    shift_y(w.tmp2.data(), T, ty, grads[j]);
  }
  if (gr[0]) {
    shift_x(rawv, T, dtx, w.tmp2.data());
    std::vector<real> t2(TT);
    shift_y(w.tmp2.data(), T, ty, t2.data());
    for (int e = 0; e < TT; ++e) grads[0][e] += t2[e];
  }
  if (gr[1]) {
    std::vector<real> t2(TT);
    shift_y(w.hx.data(), T, dty, t2.data());
    for (int e = 0; e < TT; ++e) grads[1][e] += t2[e];
  }
}

// float64 LU with partial pivoting of a row-major n x n matrix, in place.
bool lu64(std::vector<double>& A, std::vector<int>& piv, int n) {
  piv.resize(n);
  for (int k = 0; k < n; ++k) {
    int p = k;
    double best = std::fabs(A[size_t(k) * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (std::fabs(A[size_t(i) * n + k]) > best) {
        best = std::fabs(A[size_t(i) * n + k]);
        p = i;
      }
    piv[k] = p;
    if (p != k)
      for (int j = 0; j < n; ++j) std::swap(A[size_t(k) * n + j], A[size_t(p) * n + j]);
    const double d = A[size_t(k) * n + k];
    if (d == 0.0) return false;
    for (int i = k + 1; i < n; ++i) {
      const double l = A[size_t(i) * n + k] / d;
      A[size_t(i) * n + k] = l;
      if (l == 0.0) continue;
      for (int j = k + 1; j < n; ++j) A[size_t(i) * n + j] -= l * A[size_t(k) * n + j];
    }
  }
  return true;
}

void lu_solve(const std::vector<double>& LU, const std::vector<int>& piv, std::vector<double>& b,
              int n) {
  // rows were exchanged whole during the factorisation: permute b first
  for (int k = 0; k < n; ++k) std::swap(b[k], b[piv[k]]);
  for (int k = 0; k < n; ++k)
    for (int i = k + 1; i < n; ++i) b[i] -= LU[size_t(i) * n + k] * b[k];
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= LU[size_t(i) * n + j] * b[j];
    b[i] = s / LU[size_t(i) * n + i];
  }
}

// A x = b in float64 (A destroyed, b becomes x); false if singular.
bool solve64(std::vector<double>& A, std::vector<double>& b, int n) {
  std::vector<int> piv;
  if (!lu64(A, piv, n)) return false;
  lu_solve(A, piv, b, n);
  return true;
}

template <typename A, typename B>
double dotf(const A* a, const B* b, size_t n) {
  double acc = 0.0;
  for (size_t s = 0; s < n; s += 1024) {
    const size_t e = std::min(n, s + 1024);
    real p = 0;
    for (size_t i = s; i < e; ++i) p += real(a[i]) * real(b[i]);
    acc += p;
  }
  return acc;
}

struct IslandC {
  int nb, N, B, T;
  const float *data, *w;  // (nb, B, B)
  float x0, y0;
  const int32_t* stype;
  const float* n_frozen;
  const uint8_t *free_n, *mmask, *keep;
  const float *lower, *upper;
  const Psf* psf;  // per band
  const Table* tb;
};

struct Fit {
  const IslandC& I;
  const Opts& o;
  Work w;
  std::vector<std::pair<int, int>> live;
  std::vector<real> m, tan, st, gbuf;  // models (N, B2), tangents (L, B2)
  explicit Fit(const IslandC& isl, const Opts& op) : I(isl), o(op) {
    for (int k = 0; k < I.N; ++k)
      for (int j = 0; j < 6; ++j)
        if (I.mmask[k] && I.keep[k * 6 + j]) live.emplace_back(k, j);
  }
  // placed member models (and tangents of the live parameters) of band b
  void models(const float* p, int b, bool grad) {
    const int B = I.B, T = I.T, B2 = B * B, TT = T * T;
    m.assign(size_t(I.N) * B2, 0.0f);
    if (grad) tan.assign(live.size() * size_t(B2), 0.0f);
    st.resize(TT);
    gbuf.resize(6 * size_t(TT));
    for (int k = 0; k < I.N; ++k) {
      if (!I.mmask[k]) continue;
      const float* q = p + 6 * k;
      const int s = I.stype[k];
      real lr = 0.0f, e1 = 1.0f, e2 = 0.0f;
      if (s != 0) {
        lr = q[2];
        e1 = q[3];
        e2 = q[4];
        if (e1 == 0.0f && e2 == 0.0f) e1 = 1e-4f;
      }
      const double nmin = I.tb->nmin, nmax = I.tb->nmax;
      const double nval = I.free_n[k] ? nmin + (nmax - nmin) / (1.0 + std::exp(-double(q[5])))
                                      : double(I.n_frozen[k]);
      unsigned want = 0;
      std::vector<int> li;
      if (grad)
        for (size_t i = 0; i < live.size(); ++i)
          if (live[i].first == k) {
            want |= 1u << live[i].second;
            li.push_back(int(i));
          }
      real* g6[6];
      for (int j = 0; j < 6; ++j) g6[j] = gbuf.data() + size_t(j) * TT;
      member_stamp(I.psf[b], *I.tb, q[0], q[1], I.x0, I.y0, s, lr, e1, e2, real(nval), want,
                   st.data(), g6, w);
      const int ix = int(std::nearbyint(q[0] - I.x0)) - T / 2;
      const int iy = int(std::nearbyint(q[1] - I.y0)) - T / 2;
      auto place = [&](const real* src, real* dst, real scale) {
        for (int r = 0; r < T; ++r) {
          const int yy = iy + r;
          if (yy < 0 || yy >= B) continue;
          for (int c = 0; c < T; ++c) {
            const int xx = ix + c;
            if (xx >= 0 && xx < B) dst[yy * B + xx] = src[r * T + c] * scale;
          }
        }
      };
      place(st.data(), m.data() + size_t(k) * B2, 1.0f);
      for (int i : li) {
        const int j = live[i].second;
        real sc = 1.0f;
        if (j == 5) sc = real((nval - nmin) * (nmax - nval) / (nmax - nmin));
        place(g6[j], tan.data() + size_t(i) * B2, sc);
      }
    }
  }
  // VarPro fluxes of band b from the current models: f (N), A64 (N x N)
  void fluxes(int b, std::vector<double>& f, std::vector<double>& A64) {
    const int N = I.N, B2 = I.B * I.B;
    const float* d = I.data + size_t(b) * B2;
    const float* wt = I.w + size_t(b) * B2;
    std::vector<real> wm(size_t(N) * B2);
    for (int k = 0; k < N; ++k)
      for (int e = 0; e < B2; ++e) wm[size_t(k) * B2 + e] = wt[e] * m[size_t(k) * B2 + e];
    A64.assign(size_t(N) * N, 0.0);
    std::vector<double> rhs(N);
    for (int k = 0; k < N; ++k) {
      for (int l = 0; l <= k; ++l) {
        const double v = dotf(wm.data() + size_t(k) * B2, m.data() + size_t(l) * B2, B2);
        A64[size_t(k) * N + l] = A64[size_t(l) * N + k] = v;
      }
      rhs[k] = dotf(wm.data() + size_t(k) * B2, d, B2);
    }
    for (int k = 0; k < N; ++k) A64[size_t(k) * N + k] += I.mmask[k] ? double(o.reg) : 1.0;
    std::vector<double> A = A64;
    f = rhs;
    solve64(A, f, N);
    for (int k = 0; k < N; ++k)
      if (!I.mmask[k]) f[k] = 0.0;
  }
  double chi2(const float* p) {
    double c = 0.0;
    const int B2 = I.B * I.B;
    std::vector<double> f, A;
    for (int b = 0; b < I.nb; ++b) {
      models(p, b, false);
      fluxes(b, f, A);
      const float* d = I.data + size_t(b) * B2;
      const float* wt = I.w + size_t(b) * B2;
      for (int e = 0; e < B2; ++e) {
        real mod = 0.0f;
        for (int k = 0; k < I.N; ++k) mod += real(f[k]) * m[size_t(k) * B2 + e];
        const real r = std::sqrt(wt[e]) * (d[e] - mod);
        c += double(r) * double(r);
      }
    }
    return c;
  }
  // Gauss-Newton H (L x L) and g (L) over the live parameters
  void hg(const float* p, std::vector<double>& H, std::vector<double>& g) {
    const int L = int(live.size()), N = I.N, B2 = I.B * I.B;
    H.assign(size_t(L) * L, 0.0);
    g.assign(L, 0.0);
    std::vector<double> f, A64;
    std::vector<real> J(size_t(L) * B2), r(B2), sw(B2), wm(size_t(N) * B2);
    for (int b = 0; b < I.nb; ++b) {
      models(p, b, true);
      fluxes(b, f, A64);
      const float* d = I.data + size_t(b) * B2;
      const float* wt = I.w + size_t(b) * B2;
      for (int e = 0; e < B2; ++e) {
        real mod = 0.0f;
        for (int k = 0; k < N; ++k) mod += real(f[k]) * m[size_t(k) * B2 + e];
        sw[e] = std::sqrt(wt[e]);
        r[e] = sw[e] * (d[e] - mod);
      }
      for (int k = 0; k < N; ++k)
        for (int e = 0; e < B2; ++e) wm[size_t(k) * B2 + e] = wt[e] * m[size_t(k) * B2 + e];
      std::vector<double> LU = A64;
      std::vector<int> piv;
      lu64(LU, piv, N);
      for (int i = 0; i < L; ++i) {
        const int k = live[i].first;
        const real* t = tan.data() + size_t(i) * B2;
        std::vector<double> u(N), rhs(N, 0.0);
        for (int l = 0; l < N; ++l) u[l] = dotf(wm.data() + size_t(l) * B2, t, B2);
        std::vector<real> wt_t(B2);
        for (int e = 0; e < B2; ++e) wt_t[e] = wt[e] * t[e];
        const double dbk = dotf(wt_t.data(), d, B2);
        // db - dA f with dA = e_k u^T + u e_k^T
        double uf = 0.0;
        for (int l = 0; l < N; ++l) uf += u[l] * f[l];
        for (int l = 0; l < N; ++l) rhs[l] = -u[l] * f[k];
        rhs[k] += dbk - uf;
        lu_solve(LU, piv, rhs, N);
        for (int l = 0; l < N; ++l)
          if (!I.mmask[l]) rhs[l] = 0.0;
        real* Ji = J.data() + size_t(i) * B2;
        const real fk = real(f[k]);
        for (int e = 0; e < B2; ++e) {
          real dm = fk * t[e];
          for (int l = 0; l < N; ++l) dm += real(rhs[l]) * m[size_t(l) * B2 + e];
          Ji[e] = -sw[e] * dm;
        }
      }
      for (int i = 0; i < L; ++i) {
        const real* Ji = J.data() + size_t(i) * B2;
        for (int j = 0; j <= i; ++j) {
          const double v = dotf(Ji, J.data() + size_t(j) * B2, B2);
          H[size_t(i) * L + j] += v;
          if (j != i) H[size_t(j) * L + i] += v;
        }
        g[i] += dotf(Ji, r.data(), B2);
      }
    }
  }
  // legacy LM loop from p (6N floats, updated in place)
  void run(float* p, int& nst, bool& conv, bool& stall) {
    const int N = I.N, nP = 6 * N, L = int(live.size());
    double lam = kLam0, c2 = chi2(p);
    int consec = 0, nrej = 0;
    nst = 0;
    conv = stall = false;
    std::vector<float> pt(nP);
    const float cv[6] = {o.pos_max, o.pos_max, o.shape_max, o.shape_max, o.shape_max, o.un_max};
    for (int it = 0; it < o.max_steps && !conv; ++it) {
      std::vector<double> Hl, gl;
      hg(p, Hl, gl);
      std::vector<double> Hd(size_t(nP) * nP, 0.0), rhs(nP, 0.0);
      std::vector<int> idx(L);
      for (int i = 0; i < L; ++i) idx[i] = 6 * live[i].first + live[i].second;
      std::vector<char> keep(nP, 0);
      for (int i = 0; i < L; ++i) keep[idx[i]] = 1;
      for (int i = 0; i < nP; ++i)
        if (!keep[i]) Hd[size_t(i) * nP + i] = 1.0;
      for (int i = 0; i < L; ++i) {
        for (int j = 0; j < L; ++j) Hd[size_t(idx[i]) * nP + idx[j]] = Hl[size_t(i) * L + j];
        rhs[idx[i]] = -gl[i];
      }
      for (int i = 0; i < nP; ++i) {
        double& dg = Hd[size_t(i) * nP + i];
        dg = dg + lam * dg + 1e-8;
      }
      solve64(Hd, rhs, nP);
      for (int i = 0; i < nP; ++i) {
        float dl = keep[i] ? float(rhs[i]) : 0.0f;
        dl = std::min(std::max(dl, -cv[i % 6]), cv[i % 6]);
        pt[i] = std::min(std::max(p[i] + dl, I.lower[i]), I.upper[i]);
      }
      const double ct = chi2(pt.data());
      const bool acc = ct < c2;
      float pos_step = 0.0f, shape_step = 0.0f;
      if (acc)
        for (int k = 0; k < N; ++k) {
          if (!I.mmask[k]) continue;
          const float sx = pt[6 * k] - p[6 * k], sy = pt[6 * k + 1] - p[6 * k + 1];
          pos_step = std::max(pos_step, std::hypot(sx, sy));
          for (int j = 2; j < 6; ++j)
            shape_step = std::max(shape_step, std::fabs(pt[6 * k + j] - p[6 * k + j]));
        }
      const bool small = acc && (c2 - ct) < double(o.ftol_rel) * std::max(c2, 1.0);
      consec = small ? consec + 1 : (acc ? 0 : consec);
      const bool step_small = acc && pos_step < o.step_tol && shape_step < o.shape_tol;
      nrej = acc ? 0 : nrej + 1;
      stall = nrej >= kStallK;
      conv = consec >= 2 || step_small || stall;
      if (acc) {
        std::copy(pt.begin(), pt.end(), p);
        c2 = ct;
        lam = std::max(lam * kDown, 1e-9);
      } else {
        lam = std::min(lam * kUp, 1e6);
      }
      ++nst;
    }
  }
  // fluxes and sqrt(diag(A64^-1)) per band (final_fluxes)
  void final_fluxes(const float* p, float* out_f, float* out_e /* (nb, N) */) {
    std::vector<double> f, A;
    std::vector<int> piv;
    const int N = I.N;
    for (int b = 0; b < I.nb; ++b) {
      models(p, b, false);
      fluxes(b, f, A);
      lu64(A, piv, N);
      for (int k = 0; k < N; ++k) {
        out_f[size_t(b) * N + k] = float(f[k]);
        std::vector<double> e(N, 0.0);
        e[k] = 1.0;
        lu_solve(A, piv, e, N);
        out_e[size_t(b) * N + k] = float(std::sqrt(std::max(e[k], 1e-30)));
      }
    }
  }
};

}  // namespace

extern "C" {

int uphc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Fit n_isl islands.  Island i: window side B[i], stamp tier T[i], N[i]
// member slots at mem_off[i] (member arrays: stype, n_frozen, free_n,
// mmask, keep (x6), p0 / lower / upper (x6)), pixels at pix_off[i] in data
// and w ((nb, B, B) each); x0b, y0b window origin.  PSF tiers: tier_T[t]
// with planar coefficients (nb, nt, T, T) at tier_off[t] of coefs; geo
// (nb, 4).  Sersic table: nk knots, K components.  opts = reg, pos / shape
// / u clips, ftol_rel, step_tol, shape_step_tol.  Outputs: p (6 per
// member), flux and its error sqrt(diag(A^-1)) (nb per member slot,
// band-major within the island: flux[mem_off[i] * nb + b * N + k]), nst, conv, stall and, when sec_out is
// given, each island's wall time on its thread.
int uphc_fit_islands(int n_isl, int nb, int nt, const int32_t* iB, const int32_t* iT,
                     const int32_t* iN, const int64_t* pix_off, const int32_t* mem_off,
                     const float* data, const float* w, const float* x0b, const float* y0b,
                     const int32_t* stype, const float* n_frozen, const uint8_t* free_n,
                     const uint8_t* mmask, const uint8_t* keep, const float* p0,
                     const float* lower, const float* upper, int n_tiers, const int32_t* tier_T,
                     const float* coefs, const int64_t* tier_off, const float* geo, int nk,
                     int K, const double* knots, const double* amps, const double* log_vars,
                     const double* core, double nmin, double nmax, const float* opts,
                     int max_steps, int n_threads, float* p_out, float* flux_out,
                     float* ferr_out, int32_t* nst_out, uint8_t* conv_out, uint8_t* stall_out,
                     double* sec_out) {
  if (n_isl <= 0) return 0;
  if (K > 32 || nt > 21 || n_tiers < 1) return 1;
  const Table tb{nk, K, knots, amps, log_vars, core, nmin, nmax};
  const Opts o{opts[0], opts[1], opts[2], opts[3], opts[4], opts[5], opts[6], max_steps};
  std::vector<std::vector<Psf>> psf(n_tiers, std::vector<Psf>(nb));
  for (int t = 0; t < n_tiers; ++t)
    for (int b = 0; b < nb; ++b) {
      const int T = tier_T[t];
      psf[t][b] = Psf{coefs + tier_off[t] + size_t(b) * nt * T * T, T, nt, geo[4 * b],
                      geo[4 * b + 1], geo[4 * b + 2], geo[4 * b + 3]};
    }
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
  for (int i = 0; i < n_isl; ++i) {
    int tix = -1;
    for (int t = 0; t < n_tiers; ++t)
      if (tier_T[t] == iT[i]) tix = t;
    if (tix < 0) {
      ++bad;
      continue;
    }
#ifdef _OPENMP
    const double t0 = omp_get_wtime();
#endif
    const int mo = mem_off[i], N = iN[i];
    IslandC I{nb, N, iB[i], iT[i], data + pix_off[i], w + pix_off[i], x0b[i], y0b[i],
              stype + mo, n_frozen + mo, free_n + mo, mmask + mo, keep + 6 * size_t(mo),
              lower + 6 * size_t(mo), upper + 6 * size_t(mo), psf[tix].data(), &tb};
    Fit fit(I, o);
    float* p = p_out + 6 * size_t(mo);
    std::copy(p0 + 6 * size_t(mo), p0 + 6 * size_t(mo + N), p);
    int nst;
    bool conv, stall;
    fit.run(p, nst, conv, stall);
    fit.final_fluxes(p, flux_out + size_t(mo) * nb, ferr_out + size_t(mo) * nb);
    nst_out[i] = nst;
    conv_out[i] = conv;
    stall_out[i] = stall;
#ifdef _OPENMP
    if (sec_out) sec_out[i] = omp_get_wtime() - t0;
#endif
  }
  return bad ? 1 : 0;
}

// Gauss-Newton H (L x L, live parameters in member-major order), g and
// chi2 of island 0 of the same argument layout at the parameters p (6N):
// the check of the Jacobian against the float64 oracle.
int uphc_island_hg(int nb, int nt, const int32_t* iB, const int32_t* iT, const int32_t* iN,
                   const float* data, const float* w, const float* x0b, const float* y0b,
                   const int32_t* stype, const float* n_frozen, const uint8_t* free_n,
                   const uint8_t* mmask, const uint8_t* keep, const float* lower,
                   const float* upper, const float* coefs, const float* geo, int nk, int K,
                   const double* knots, const double* amps, const double* log_vars,
                   const double* core, double nmin, double nmax, const float* opts,
                   const float* p, double* H, double* g, double* chi2) {
  const Table tb{nk, K, knots, amps, log_vars, core, nmin, nmax};
  const Opts o{opts[0], opts[1], opts[2], opts[3], opts[4], opts[5], opts[6], 0};
  std::vector<Psf> psf(nb);
  const int T = iT[0];
  for (int b = 0; b < nb; ++b)
    psf[b] = Psf{coefs + size_t(b) * nt * T * T, T, nt, geo[4 * b], geo[4 * b + 1],
                 geo[4 * b + 2], geo[4 * b + 3]};
  IslandC I{nb, iN[0], iB[0], T, data, w, x0b[0], y0b[0], stype, n_frozen, free_n, mmask,
            keep, lower, upper, psf.data(), &tb};
  Fit fit(I, o);
  std::vector<double> Hv, gv;
  *chi2 = fit.chi2(p);
  fit.hg(p, Hv, gv);
  std::copy(Hv.begin(), Hv.end(), H);
  std::copy(gv.begin(), gv.end(), g);
  return int(gv.size());
}

// Unit-flux stamps (n, T, T) of the linear solve (render_stamps,
// forced_phot_sparse.py:126-220): source s of band band[s] at (x, y),
// shifted by the remainder against (gx0[s], gy0[s]); `renorm` rescales the
// raw PSF of an all-star set to unit sum before the shift.  PSF planar
// (nb, nt, T, T) with geo (nb, 4); Sersic table as uphc_fit_islands.
int uphc_render_stamps(int n, int T, int nt, const float* coefs, const float* geo,
                       const int32_t* band, const float* x, const float* y, const float* gx0,
                       const float* gy0, const int32_t* stype, const float* logre,
                       const float* ee1, const float* ee2, const float* sersic_n, int renorm,
                       int nk, int K, const double* knots, const double* amps,
                       const double* log_vars, const double* core, double nmin, double nmax,
                       int n_threads, float* out) {
  const Table tb{nk, K, knots, amps, log_vars, core, nmin, nmax};
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  const size_t TT = size_t(T) * T;
#pragma omp parallel
  {
    Work w;
    std::vector<real> st(TT);
#pragma omp for schedule(dynamic, 4)
    for (int s = 0; s < n; ++s) {
      const int b = band[s];
      const Psf P{coefs + size_t(b) * nt * TT, T, nt, geo[4 * b], geo[4 * b + 1],
                  geo[4 * b + 2], geo[4 * b + 3]};
      if (renorm && stype[s] == 0) {
        // all-star set: PSF / its sum, then the shift
        std::vector<real> raw(TT, real(0));
        const real X = (x[s] - P.xo) / P.xs, Y = (y[s] - P.yo) / P.ys;
        real px[6], py[6];
        px[0] = py[0] = real(1);
        for (int i = 1; i < 6; ++i) {
          px[i] = px[i - 1] * X;
          py[i] = py[i - 1] * Y;
        }
        for (int t = 0; t < nt; ++t) {
          const real bt = px[kTermM[t]] * py[kTermN[t]];
          for (size_t e = 0; e < TT; ++e) raw[e] += P.c[t * TT + e] * bt;
        }
        real tot = 0;
        for (size_t e = 0; e < TT; ++e) tot += raw[e];
        tot = std::max(tot, real(1e-12));
        for (size_t e = 0; e < TT; ++e) raw[e] /= tot;
        const real xg = x[s] - gx0[s], yg = y[s] - gy0[s];
        real tx[7], ty[7];
        lanczos_taps(xg - std::nearbyint(xg), tx, nullptr);
        lanczos_taps(yg - std::nearbyint(yg), ty, nullptr);
        shift_x(raw.data(), T, tx, st.data());
        shift_y(st.data(), T, ty, raw.data());
        for (size_t e = 0; e < TT; ++e) out[size_t(s) * TT + e] = float(raw[e]);
        continue;
      }
      real* g6[6] = {};
      member_stamp(P, tb, x[s], y[s], gx0[s], gy0[s], stype[s], logre[s], ee1[s], ee2[s],
                   sersic_n[s], 0u, st.data(), g6, w);
      for (size_t e = 0; e < TT; ++e) out[size_t(s) * TT + e] = float(st[e]);
    }
  }
  return 0;
}

// Grid-band linear solves (solver_inner): problem q has sources
// [src_off[q], src_off[q+1]) with stamps (T x T at window-local top-left
// ix, iy), neighbour lists nbr (k_max per source, -1 = none, source indices
// local to the problem), its window's data / invvar / good at pix_off[q]
// (h[q] x w[q]) and the plane background (order 1) in X = (x + gx0 - xo)/xs,
// Y = (y + gy0 - yo)/ys.  Outputs per source: flux, var; per problem: bg
// (3), bg_var (3).
int uphc_band_solve(int n_prob, int T, int k_max, const int64_t* src_off, const int64_t* pix_off,
                    const int32_t* hh, const int32_t* ww, const float* gx0, const float* gy0,
                    const float* data, const float* invvar, const uint8_t* good,
                    const float* stamps, const int32_t* ix, const int32_t* iy,
                    const int32_t* nbr, double xo, double xs, double yo, double ys, double reg,
                    int n_threads, float* flux, float* var, float* bg, float* bg_var) {
#ifdef _OPENMP
  if (n_threads > 0) omp_set_num_threads(n_threads);
#endif
  const int M = 3, TT = T * T;
#pragma omp parallel for schedule(dynamic, 1)
  for (int q = 0; q < n_prob; ++q) {
    const int64_t s0 = src_off[q];
    const int n = int(src_off[q + 1] - s0), h = hh[q], wd = ww[q], D = n + M;
    const float* d = data + pix_off[q];
    const float* iv = invvar + pix_off[q];
    const uint8_t* gd = good + pix_off[q];
    std::vector<float> wg(size_t(h) * wd), bx(size_t(h) * wd), by(size_t(h) * wd);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < wd; ++x) {
        const size_t e = size_t(y) * wd + x;
        wg[e] = gd[e] ? iv[e] : 0.0f;
        bx[e] = float((x + gx0[q] - xo) / xs);
        by[e] = float((y + gy0[q] - yo) / ys);
      }
    std::vector<double> A(size_t(D) * D, 0.0), rhs(D, 0.0);
    std::vector<float> eff(size_t(n) * TT), wu(size_t(n) * TT);
    for (int s = 0; s < n; ++s) {
      const float* st = stamps + (s0 + s) * TT;
      double dg = 0, tdv = 0, tb[3] = {0, 0, 0};
      for (int r = 0; r < T; ++r)
        for (int c = 0; c < T; ++c) {
          const int yy = iy[s0 + s] + r, xx = ix[s0 + s] + c;
          const bool in = yy >= 0 && yy < h && xx >= 0 && xx < wd;
          const size_t e = in ? size_t(yy) * wd + xx : 0;
          const float wv = in ? wg[e] : 0.0f, ev = in ? st[r * T + c] : 0.0f;
          eff[size_t(s) * TT + r * T + c] = ev;
          wu[size_t(s) * TT + r * T + c] = wv;
          const float ws = wv * ev;
          dg += double(ws) * ev;
          if (in) {
            tdv += double(ws) * d[e];
            tb[0] += ws;
            tb[1] += double(ws) * bx[e];
            tb[2] += double(ws) * by[e];
          }
        }
      A[size_t(s) * D + s] = dg;
      rhs[s] = tdv;
      for (int m = 0; m < M; ++m) {
        A[size_t(s) * D + n + m] = tb[m];
        A[size_t(n + m) * D + s] = tb[m];
      }
    }
    for (int s = 0; s < n; ++s)
      for (int k = 0; k < k_max; ++k) {
        const int t = nbr[(s0 + s) * k_max + k];
        if (t < 0) continue;
        const int di = iy[s0 + s] - iy[s0 + t], dj = ix[s0 + s] - ix[s0 + t];
        double acc = 0.0;
        for (int r = 0; r < T; ++r) {
          const int tr = r + di;
          if (tr < 0 || tr >= T) continue;
          for (int c = 0; c < T; ++c) {
            const int tc = c + dj;
            if (tc < 0 || tc >= T) continue;
            acc += double(wu[size_t(s) * TT + r * T + c] * eff[size_t(s) * TT + r * T + c]) *
                   eff[size_t(t) * TT + tr * T + tc];
          }
        }
        A[size_t(s) * D + t] += acc;
      }
    double bb[3][3] = {}, bd[3] = {};
    for (size_t e = 0; e < size_t(h) * wd; ++e) {
      const double v[3] = {1.0, bx[e], by[e]};
      for (int a = 0; a < 3; ++a) {
        bd[a] += double(wg[e]) * v[a] * d[e];
        for (int b = 0; b < 3; ++b) bb[a][b] += double(wg[e]) * v[a] * v[b];
      }
    }
    for (int a = 0; a < 3; ++a) {
      rhs[n + a] = bd[a];
      for (int b = 0; b < 3; ++b) A[size_t(n + a) * D + n + b] = bb[a][b];
    }
    for (int i = 0; i < D; ++i) {
      double& dg = A[size_t(i) * D + i];
      dg += reg * std::max(std::fabs(dg), 1e-30);
    }
    // solve and the inverse diagonal (one LU, D + 1 right-hand sides)
    std::vector<int> piv;
    std::vector<double> x = rhs;
    lu64(A, piv, D);
    lu_solve(A, piv, x, D);
    for (int s = 0; s < n; ++s) flux[s0 + s] = float(x[s]);
    for (int m = 0; m < M; ++m) bg[q * M + m] = float(x[n + m]);
    for (int i = 0; i < D; ++i) {
      std::vector<double> e(D, 0.0);
      e[i] = 1.0;
      lu_solve(A, piv, e, D);
      if (i < n)
        var[s0 + i] = float(e[i]);
      else
        bg_var[q * M + i - n] = float(e[i]);
    }
  }
  return 0;
}

}  // extern "C"
