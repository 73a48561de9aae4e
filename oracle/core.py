"""FP64 restatement of the rendering primitives (TEST INFRASTRUCTURE).

Covers the polynomial PSF, the Lanczos-3 sub-pixel shift, the Sersic
mixture-of-Gaussians table, the analytic-FT galaxy renderer and the
unit-flux member stamps with their parameter derivatives.

Reference (paths under ``tract7dt/fit``):
  * PSF polynomial: ``grid_gpu/psf_polynomial.py:134-197``; term order
    ``grid_gpu/bg_polynomial.py:40-100``.
  * Lanczos-3: ``optimize/gpu/kernels/lanczos.py:50-181``.
  * Sersic table lookup: ``grid_gpu/sersic_mog.py:87-100``; generator
    ``grid_gpu/sersic_mog_gen.py:61-111`` (Shajib 2019 rows).
  * Ellipse terms and FFT galaxy render: ``grid_gpu/sersic_mog.py:108-207``.
  * Stamp assembly: ``grid_gpu/source_model.py:264-390``; island render
    ``grid_gpu/island_solver.py:441-562``.

Parity pins (tests/test_oracle_*.py):
  * PSF eval: n_t=1 constant stamp; closed-form polynomial on random coefs.
  * Lanczos taps: sum to 1; d=0 is the identity; shift matches
    ``scipy.ndimage.correlate1d(mode="constant")``; tap derivative vs
    central finite differences.
  * Sersic table: every row equals the reference's committed table rows
    (golden fixture: n >= 0.8 to 1e-12, the Tractor rows below exactly);
    Tractor rows sum to 1; Shajib amplitudes integrate to the row total.
  * Galaxy FFT render: Gaussian(1 component) x Gaussian PSF matches the
    closed-form convolved Gaussian away from the wrap; flux = total - wings.
  * Member-stamp derivatives: central finite differences in float64.
  * Direct backend (``direct_stamp``): the shift-and-add convolution equals
    ``scipy.signal.fftconvolve(mode="same")``; a one-component round
    profile convolved with a Gaussian PSF matches the sampled Gaussian of
    the summed variance; the stamp sums to 1 before the shift; STAR rows
    equal the PSF render.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

SOURCE_STAR, SOURCE_EXP, SOURCE_DEV, SOURCE_SERSIC = 0, 1, 2, 3
SERSIC_N_MIN = 0.29
SERSIC_N_MAX = 6.3


# ---------------------------------------------------------------------------
# Polynomial PSF  (psf_polynomial.py:134-197, bg_polynomial.py:76-100)
# ---------------------------------------------------------------------------

def terms_for_order(order: int):
    """(m, n) exponent pairs of X^m Y^n, low total degree first and, within
    one degree, decreasing power of X (bg_polynomial.py:40-75)."""
    out = []
    for d in range(order + 1):
        for m in range(d, -1, -1):
            out.append((m, d - m))
    return tuple(out)


def order_for_n_terms(n_terms: int) -> int:
    table = {1: 0, 3: 1, 6: 2, 10: 3, 15: 4, 21: 5}
    if n_terms not in table:
        raise ValueError(f"unsupported n_terms={n_terms}")
    return table[n_terms]


def spatial_basis(X: float, Y: float, n_terms: int):
    """Basis values and their X / Y partial derivatives (float64)."""
    terms = terms_for_order(order_for_n_terms(n_terms))
    b = np.empty(n_terms)
    bx = np.zeros(n_terms)
    by = np.zeros(n_terms)
    for t, (m, n) in enumerate(terms):
        b[t] = X ** m * Y ** n
        if m > 0:
            bx[t] = m * X ** (m - 1) * Y ** n
        if n > 0:
            by[t] = n * X ** m * Y ** (n - 1)
    return b, bx, by


def eval_psf(psf: dict, x: float, y: float, grad: bool = False):
    """PSF stamp at image position (x, y) (psf_polynomial.py:174-197).

    ``psf`` = dict(coefs (T,T,nt), x_origin, x_scale, y_origin, y_scale).
    With ``grad`` also returns d(stamp)/dx and d(stamp)/dy.
    """
    c = np.asarray(psf["coefs"], dtype=np.float64)
    X = (x - psf["x_origin"]) / psf["x_scale"]
    Y = (y - psf["y_origin"]) / psf["y_scale"]
    b, bx, by = spatial_basis(X, Y, c.shape[-1])
    stamp = c @ b
    if not grad:
        return stamp
    return stamp, (c @ bx) / psf["x_scale"], (c @ by) / psf["y_scale"]


def pad_psf_coefs(coefs: np.ndarray, tier: int) -> np.ndarray:
    """Zero-pad (T,T,nt) coefficients symmetrically to a larger odd tier,
    keeping the centre pixel at tier//2 (island_solver.py:193-206)."""
    H, W = coefs.shape[:2]
    if tier <= H and tier <= W:
        return coefs
    out = np.zeros((tier, tier) + coefs.shape[2:], dtype=coefs.dtype)
    py, px = (tier - H) // 2, (tier - W) // 2
    out[py:py + H, px:px + W] = coefs
    return out


# ---------------------------------------------------------------------------
# Lanczos-3  (lanczos.py:50-181)
# ---------------------------------------------------------------------------

def _l3(x: np.ndarray):
    """Lanczos-3 window value and derivative at x (float64)."""
    x = np.asarray(x, dtype=np.float64)
    val = np.zeros_like(x)
    der = np.zeros_like(x)
    inside = (np.abs(x) < 3.0) & (np.abs(x) >= 1e-12)
    xi = x[inside]
    px = np.pi * xi
    s1, c1 = np.sin(px), np.cos(px)
    s3, c3 = np.sin(px / 3.0), np.cos(px / 3.0)
    v = 3.0 * s1 * s3 / (px * px)
    val[inside] = v
    der[inside] = (3.0 * np.pi * c1 * s3 + np.pi * s1 * c3) / (px * px) - 2.0 * v / xi
    val[np.abs(x) < 1e-12] = 1.0
    return val, der


def lanczos_taps(d: float, grad: bool = False):
    """7 normalised taps at offsets k-3+d, k=0..6, and optionally their
    derivative with respect to d."""
    x = np.arange(-3, 4, dtype=np.float64) + float(d)
    v, dv = _l3(x)
    s = v.sum()
    taps = v / s
    if not grad:
        return taps
    ds = dv.sum()
    return taps, (dv * s - v * ds) / (s * s)


def correlate_x(img: np.ndarray, k: np.ndarray) -> np.ndarray:
    """out[i, j] = sum_k k[k] * img[i, j+k-3], zero outside (axis 1)."""
    n = img.shape[1]
    pad = np.pad(img, ((0, 0), (3, 3)))
    out = k[0] * pad[:, 0:n]
    for t in range(1, 7):
        out = out + k[t] * pad[:, t:t + n]
    return out


def correlate_y(img: np.ndarray, k: np.ndarray) -> np.ndarray:
    n = img.shape[0]
    pad = np.pad(img, ((3, 3), (0, 0)))
    out = k[0] * pad[0:n, :]
    for t in range(1, 7):
        out = out + k[t] * pad[t:t + n, :]
    return out


def lanczos_shift(img: np.ndarray, dx: float, dy: float) -> np.ndarray:
    return correlate_y(correlate_x(img, lanczos_taps(dx)), lanczos_taps(dy))


# ---------------------------------------------------------------------------
# Sersic mixture-of-Gaussians table (sersic_mog_gen.py:61-111, Shajib 2019)
# ---------------------------------------------------------------------------

K_MAX = 15


def _euler_weights(P: int = 10):
    """Weights/nodes of the Euler-summation inverse Laplace transform
    (Abate & Whitt 2006) used by Shajib (2019) to expand a profile in
    Gaussians."""
    from scipy.special import comb
    k = np.arange(2 * P + 1)
    beta = np.sqrt(2.0 * P * np.log(10.0) / 3.0 + 2.0j * np.pi * k)
    xi = np.zeros(2 * P + 1)
    xi[0] = 0.5
    xi[1:P + 1] = 1.0
    xi[2 * P] = 2.0 ** -P
    for j in range(1, P):
        xi[2 * P - j] = xi[2 * P - j + 1] + comb(P, j) * 2.0 ** -P
    eta = (-1.0) ** k * xi * 10.0 ** (P / 3.0) * 2.0 * np.sqrt(2.0 * np.pi)
    return eta, beta


def sersic_bn(n):
    """Sersic b_n approximation used by the table (pysersic form)."""
    n = np.asarray(n, dtype=np.float64)
    lead = 2.0 * n - 1.0 / 3.0 + 4.0 / (405.0 * n) + 46.0 / (25515.0 * n * n)
    a = (0.00018073182821237496 / (n + np.exp(n)) + 3.7026973571904255e-5)
    a = a / (-0.09149183119702775 - n) ** 2
    c = (np.log(n) + (2.6248718397705195 * n - 0.9727511612512357)
         / (n ** -4 + 94.78011643586419)) / (n - 0.006044236674273689)
    return lead + a * c


def _sersic_profile(r, n):
    from scipy.special import gammaln
    b = sersic_bn(n)
    ie = b ** (2 * n) / (2.0 * np.pi * n * np.exp(b + gammaln(2 * n)))
    return ie * np.exp(-b * (r ** (1.0 / n) - 1.0))


def shajib_row(n: float, K: int = K_MAX, P: int = 10):
    """Amplitudes and variances (units Re^2) of the K-Gaussian expansion of
    a unit-total-flux, untruncated Sersic profile of index n."""
    eta, beta = _euler_weights(P)
    s_hi = float(np.clip(2.5 + 12.5 * (n - 0.8) / 1.2, 2.5, 15.0))
    sig = np.logspace(np.log10(1e-2), np.log10(s_hi), K)
    f = np.sum(eta * _sersic_profile(np.outer(sig, beta), n).real, axis=1)
    dlog = np.abs(np.diff(np.log(sig)).mean())
    amps = f * dlog / np.sqrt(2.0 * np.pi)
    amps[0] *= 0.5
    amps[-1] *= 0.5
    return amps * 2.0 * np.pi * sig * sig, sig * sig


@lru_cache(maxsize=1)
def sersic_table():
    """The reference's two-family table (sersic_mog_gen.py:45-52, 114-186):
    The Tractor's rows for n in [0.29, 0.8], which exist only as data (read
    from the data file the package ships, uniphot_amd/data), then Shajib rows
    on a 0.005 grid over [0.8, 6.3] computed here."""
    from pathlib import Path
    low = np.load(Path(__file__).resolve().parents[1] / "uniphot_amd" / "data" /
                  "sersic_tractor_rows.npz", allow_pickle=False)
    kn = np.linspace(0.8, SERSIC_N_MAX, int(round((SERSIC_N_MAX - 0.8) / 0.005)) + 1)
    rows = [shajib_row(float(n)) for n in kn]
    return dict(n_knots=np.concatenate([low["n_knots"], kn]),
                amps=np.concatenate([low["amps"], np.array([a for a, _ in rows])]),
                log_vars=np.concatenate([low["log_vars"],
                                         np.log(np.array([v for _, v in rows]))]),
                core_amp=np.concatenate([low["core_amp"], np.zeros(kn.size)]),
                n_min=SERSIC_N_MIN, n_max=SERSIC_N_MAX)


def mog_params_at(n: float, grad: bool = False):
    """n -> (amps, vars, core) by linear interpolation inside one knot
    segment (sersic_mog.py:87-100).  With ``grad`` also the n-derivatives."""
    tb = sersic_table()
    kn = tb["n_knots"]
    nn = float(np.clip(n, tb["n_min"], tb["n_max"]))
    clipped = nn != float(n)
    i = int(np.searchsorted(kn, nn, side="right")) - 1
    i = min(max(i, 0), kn.size - 2)
    n0, n1 = kn[i], kn[i + 1]
    t = (nn - n0) / (n1 - n0) if n1 > n0 else 0.0
    a = (1 - t) * tb["amps"][i] + t * tb["amps"][i + 1]
    lv = (1 - t) * tb["log_vars"][i] + t * tb["log_vars"][i + 1]
    c = (1 - t) * tb["core_amp"][i] + t * tb["core_amp"][i + 1]
    v = np.exp(lv)
    if not grad:
        return a, v, c
    dt = 0.0 if (clipped or not n1 > n0) else 1.0 / (n1 - n0)
    da = dt * (tb["amps"][i + 1] - tb["amps"][i])
    dv = v * dt * (tb["log_vars"][i + 1] - tb["log_vars"][i])
    dc = dt * (tb["core_amp"][i + 1] - tb["core_amp"][i])
    return a, v, c, da, dv, dc


def sersic_u_to_n(u):
    return SERSIC_N_MIN + (SERSIC_N_MAX - SERSIC_N_MIN) / (1.0 + np.exp(-u))


def sersic_n_to_u(n):
    t = (np.clip(n, SERSIC_N_MIN + 1e-4, SERSIC_N_MAX - 1e-4) - SERSIC_N_MIN) / (
        SERSIC_N_MAX - SERSIC_N_MIN)
    return np.log(t / (1.0 - t))


# ---------------------------------------------------------------------------
# Ellipse covariance and the FFT galaxy renderer (sersic_mog.py:108-207)
# ---------------------------------------------------------------------------

def _cov_terms_c(logre, ee1, ee2):
    """(cxx, cyy, cxy) written with analytic operations only, so a complex
    step in one argument yields its exact derivative.  Equal to the
    reference's polar form (hypot/arctan2) away from the clip limits."""
    re = np.exp(logre)
    ee = np.sqrt(ee1 * ee1 + ee2 * ee2)
    q = np.exp(-ee)
    e = 1.0 - q
    inv_ba = (1.0 + e) / q
    s2 = 1.0 / (inv_ba * inv_ba)            # (b/a)^2
    cos2, sin2 = ee1 / ee, ee2 / ee
    half = re * re / 2.0
    cxx = half * ((1.0 + s2) - cos2 * (1.0 - s2))
    cyy = half * ((1.0 + s2) + cos2 * (1.0 - s2))
    cxy = half * sin2 * (1.0 - s2)
    return cxx, cyy, cxy


def ellipse_cov_terms(logre, ee1, ee2, grad: bool = False):
    """EllipseESoft (logre, ee1, ee2) -> (cxx, cyy, cxy), following the
    affine of sersic_mog.py:108-150 including its clips and the ee=0 guard.
    With ``grad`` returns (terms, d/dlogre, d/dee1, d/dee2), each a
    3-vector."""
    logre = float(logre)
    ee1 = float(ee1)
    ee2 = float(ee2)
    if ee1 == 0.0 and ee2 == 0.0:
        ee1 = 1e-4
    re = min(max(math.exp(logre), 0.1), 1000.0)
    ee = math.hypot(ee1, ee2)
    e = 1.0 - math.exp(-ee)
    inv_ba = min((1.0 + e) / max(1.0 - e, 1e-4), 1000.0)
    th = 0.5 * math.atan2(ee2, ee1)
    ct, st = math.cos(th), math.sin(th)
    a, b = re * ct / inv_ba, re * st
    c, d = -re * st / inv_ba, re * ct
    terms = np.array([a * a + b * b, c * c + d * d, a * c + b * d])
    if not grad:
        return terms
    h = 1e-30
    re_clip = not (0.1 < math.exp(logre) < 1000.0)
    ba_clip = not ((1.0 - e) > 1e-4 and (1.0 + e) / (1.0 - e) < 1000.0)
    out = []
    for k in range(3):
        args = [complex(logre), complex(ee1), complex(ee2)]
        args[k] += 1j * h
        vals = _cov_terms_c(*args)
        dvec = np.array([v.imag / h for v in vals])
        if k == 0 and re_clip:
            dvec[:] = 0.0
        if k > 0 and ba_clip:
            # the axis-ratio clip freezes (b/a); only the angle moves
            dvec = _clipped_axis_grad(logre, ee1, ee2, k, inv_ba, re)
        out.append(dvec)
    return terms, out[0], out[1], out[2]


def _clipped_axis_grad(logre, ee1, ee2, k, inv_ba, re):
    """Derivative when the axis ratio is clipped (rare; only the position
    angle depends on ee1/ee2)."""
    s2 = 1.0 / (inv_ba * inv_ba)
    ee = math.hypot(ee1, ee2)
    half = re * re / 2.0
    # cos2 = ee1/ee, sin2 = ee2/ee
    if k == 1:
        dcos2 = (ee * ee - ee1 * ee1) / ee ** 3
        dsin2 = -ee1 * ee2 / ee ** 3
    else:
        dcos2 = -ee1 * ee2 / ee ** 3
        dsin2 = (ee * ee - ee2 * ee2) / ee ** 3
    return np.array([-half * dcos2 * (1 - s2), half * dcos2 * (1 - s2),
                     half * dsin2 * (1 - s2)])


def fft_pad_for(stamp: int) -> int:
    """Smallest power of two >= stamp + stamp//2 (sersic_mog.py:197-207)."""
    n = 1
    while n < stamp + stamp // 2:
        n <<= 1
    return n


def _freq_grids(pad: int):
    v = np.fft.rfftfreq(pad)[None, :]
    w = np.fft.fftfreq(pad)[:, None]
    return v, w


def galaxy_fourier(amps, vars_, terms, pad: int):
    """F(v, w) = sum_k a_k exp(-2 pi^2 var_k (cxx v^2 + cyy w^2 + 2 cxy v w))
    on the rfft grid, and the exponent field."""
    v, w = _freq_grids(pad)
    cxx, cyy, cxy = terms
    expo = -2.0 * np.pi ** 2 * (cxx * v * v + cyy * w * w + 2.0 * cxy * v * w)
    F = np.zeros_like(expo)
    for a, s in zip(amps, vars_):
        F = F + a * np.exp(s * expo)
    return F, expo


def fft_convolve_crop(stamp: np.ndarray, F: np.ndarray, pad: int) -> np.ndarray:
    H, W = stamp.shape
    sp = np.zeros((pad, pad))
    sp[:H, :W] = stamp
    out = np.fft.irfft2(np.fft.rfft2(sp) * F, s=(pad, pad))
    return out[:H, :W]


def render_galaxy_raw(psf_stamp, amps, vars_, core, logre, ee1, ee2, pad):
    """Unshifted galaxy stamp: irfft2(rfft2(psf) * F)[:T,:T] + core * psf."""
    terms = ellipse_cov_terms(logre, ee1, ee2)
    F, _ = galaxy_fourier(amps, vars_, terms, pad)
    return fft_convolve_crop(psf_stamp, F, pad) + core * psf_stamp


# ---------------------------------------------------------------------------
# Member stamps (source_model.py:264-328, island_solver.py:441-562)
# ---------------------------------------------------------------------------

def _round_half_even(v: float) -> float:
    return float(np.round(v))


def member_stamp(psf: dict, x, y, x0, y0, stype, logre=0.0, ee1=0.0, ee2=0.0,
                 n_val=1.0, want=(), renorm=False, pad=None):
    """Unit-flux (T,T) stamp of one source at full-image (x, y), shifted by
    the sub-pixel remainder relative to origin (x0, y0).

    ``want`` selects derivative planes among "x", "y", "lr", "e1", "e2",
    "n" (derivative with respect to the Sersic index n itself).
    ``renorm`` applies the all-star linear-solve unit-sum renormalisation
    (forced_phot_sparse.py:156-158; no derivatives with renorm).
    Returns (stamp, {name: d stamp / d name}).
    """
    T = psf["coefs"].shape[0]
    need_xy = ("x" in want) or ("y" in want)
    if need_xy:
        raw_psf, dpx, dpy = eval_psf(psf, x, y, grad=True)
    else:
        raw_psf = eval_psf(psf, x, y)
    grads_raw = {}
    if stype == SOURCE_STAR:
        raw = raw_psf
        if renorm:
            raw = raw / max(raw.sum(), 1e-12)
        if need_xy:
            grads_raw["x"], grads_raw["y"] = dpx, dpy
    else:
        if ee1 == 0.0 and ee2 == 0.0:
            ee1 = 1e-4
        if pad is None:
            pad = fft_pad_for(T)
        a, v, c, da, dv, dc = mog_params_at(n_val, grad=True)
        if any(k in want for k in ("lr", "e1", "e2")):
            terms, dlr, de1, de2 = ellipse_cov_terms(logre, ee1, ee2, grad=True)
        else:
            terms = ellipse_cov_terms(logre, ee1, ee2)
        F, expo = galaxy_fourier(a, v, terms, pad)
        raw = fft_convolve_crop(raw_psf, F, pad) + c * raw_psf
        if need_xy:
            grads_raw["x"] = fft_convolve_crop(dpx, F, pad) + c * dpx
            grads_raw["y"] = fft_convolve_crop(dpy, F, pad) + c * dpy
        vv, ww = _freq_grids(pad)
        for name, dterm in (("lr", "lr"), ("e1", "e1"), ("e2", "e2")):
            if name not in want:
                continue
            dt = {"lr": dlr, "e1": de1, "e2": de2}[name]
            dexpo = -2.0 * np.pi ** 2 * (dt[0] * vv * vv + dt[1] * ww * ww
                                         + 2.0 * dt[2] * vv * ww)
            dF = np.zeros_like(expo)
            for ak, sk in zip(a, v):
                dF = dF + ak * sk * np.exp(sk * expo) * dexpo
            grads_raw[name] = fft_convolve_crop(raw_psf, dF, pad)
        if "n" in want:
            dF = np.zeros_like(expo)
            for ak, sk, dak, dsk in zip(a, v, da, dv):
                ex = np.exp(sk * expo)
                dF = dF + dak * ex + ak * ex * expo * dsk
            grads_raw["n"] = fft_convolve_crop(raw_psf, dF, pad) + dc * raw_psf
    xg = x - x0
    yg = y - y0
    dx = xg - _round_half_even(xg)
    dy = yg - _round_half_even(yg)
    tx, dtx = lanczos_taps(dx, grad=True)
    ty, dty = lanczos_taps(dy, grad=True)
    hx = correlate_x(raw, tx)
    stamp = correlate_y(hx, ty)
    grads = {}
    for name, g in grads_raw.items():
        grads[name] = correlate_y(correlate_x(g, tx), ty)
    if "x" in grads:
        grads["x"] = grads["x"] + correlate_y(correlate_x(raw, dtx), ty)
    if "y" in grads:
        grads["y"] = grads["y"] + correlate_y(hx, dty)
    return stamp, grads


# Exponential-profile mixture of the direct backend (source_model.py:59-75).
EXP_AMP_RAW = np.array([2.34853813e-03, 3.07995260e-02, 2.23364214e-01, 1.17949102e00,
                        4.33873750e00, 5.99820770e00])
EXP_VAR = np.array([1.20078965e-03, 8.84526493e-03, 3.91463084e-02, 1.39976817e-01,
                    4.60962500e-01, 1.50159566e00])


def exp_profile(T: int, logre, ee1, ee2, amps=None, vars_=None):
    """Unit-sum pixel-centre exponential profile on a T x T grid centred at
    T//2 (render_exp_profile_stamp_jax, source_model.py:218-251)."""
    amps = EXP_AMP_RAW / EXP_AMP_RAW.sum() if amps is None else np.asarray(amps, float)
    vars_ = EXP_VAR if vars_ is None else np.asarray(vars_, float)
    yy, xx = np.mgrid[:T, :T].astype(float) - T // 2
    re = min(max(math.exp(logre), 0.1), 1000.0)
    ee = math.hypot(ee1, ee2)
    e = 1.0 - math.exp(-ee)
    inv_ba = min((1.0 + e) / max(1.0 - e, 1e-4), 1000.0)
    th = 0.5 * math.atan2(ee2, ee1)
    ct, st = math.cos(th), math.sin(th)
    a, b, c, d = re * ct / inv_ba, re * st, -re * st / inv_ba, re * ct
    prof = np.zeros((T, T))
    for amp, var in zip(amps, vars_):
        c00, c01, c11 = var * (a * a + b * b), var * (a * c + b * d), var * (c * c + d * d)
        det = max(c00 * c11 - c01 * c01, 1e-12)
        r2 = (c11 * xx * xx - 2.0 * c01 * xx * yy + c00 * yy * yy) / det
        prof += amp / (2.0 * np.pi * math.sqrt(det)) * np.exp(-0.5 * np.minimum(r2, 200.0))
    tot = prof.sum()
    return prof / tot if tot > 0 else prof


def convolve_same(a: np.ndarray, k: np.ndarray) -> np.ndarray:
    """'same' 2-D convolution of equal odd-sized arrays by shift-and-add
    over the kernel's pixels: out[i, j] = sum a[p, q] k[i + h - p, j + h - q]."""
    T = a.shape[0]
    h = T // 2
    full = np.zeros((2 * T - 1, 2 * T - 1))
    for p in range(T):
        for q in range(T):
            if a[p, q] != 0.0:
                full[p:p + T, q:q + T] += a[p, q] * k
    return full[h:h + T, h:h + T]


def direct_stamp(psf: dict, x, y, x0, y0, stype, logre=0.0, ee1=0.0, ee2=0.0):
    """Unit-flux stamp of the direct backend (source_model.py:264-328,
    "direct" branch): STAR = PSF at (x, y); EXP = exp_profile convolved
    ('same') with it and renormalised to unit sum; Lanczos-3 shift."""
    raw_psf = eval_psf(psf, x, y)
    T = raw_psf.shape[0]
    if stype == SOURCE_EXP:
        conv = convolve_same(exp_profile(T, logre, ee1, ee2), raw_psf)
        s = conv.sum()
        raw = conv / s if s > 0 else conv
    elif stype == SOURCE_STAR:
        raw = raw_psf
    else:
        raise ValueError("the direct backend renders STAR and EXP only")
    xg, yg = x - x0, y - y0
    tx = lanczos_taps(xg - _round_half_even(xg))
    ty = lanczos_taps(yg - _round_half_even(yg))
    return correlate_y(correlate_x(raw, tx), ty)


def place(stamp: np.ndarray, ix_top: int, iy_top: int, B: int) -> np.ndarray:
    """(T,T) stamp -> (B,B) window at integer top-left offset
    (island_solver.py:469-478)."""
    T = stamp.shape[0]
    out = np.zeros((B, B))
    y0, x0 = max(0, iy_top), max(0, ix_top)
    y1, x1 = min(B, iy_top + T), min(B, ix_top + T)
    if y1 > y0 and x1 > x0:
        out[y0:y1, x0:x1] = stamp[y0 - iy_top:y1 - iy_top, x0 - ix_top:x1 - ix_top]
    return out


def place_window(stamp: np.ndarray, ix_top: int, iy_top: int, h: int, w: int) -> np.ndarray:
    """(T,T) stamp -> (h,w) window at integer top-left offset, clipped."""
    T = stamp.shape[0]
    out = np.zeros((h, w))
    y0, x0 = max(0, iy_top), max(0, ix_top)
    y1, x1 = min(h, iy_top + T), min(w, ix_top + T)
    if y1 > y0 and x1 > x0:
        out[y0:y1, x0:x1] = stamp[y0 - iy_top:y1 - iy_top, x0 - ix_top:x1 - ix_top]
    return out


def stamp_top_left(x, y, x0, y0, T):
    """round(pos - origin) - T//2 (forced_phot_sparse.py:255-264)."""
    return (int(_round_half_even(x - x0)) - T // 2,
            int(_round_half_even(y - y0)) - T // 2)
