"""FP64 restatement of the correlated-noise variance factor F_corr
(TEST INFRASTRUCTURE: the checker of the device path, never the product).

Reference: ``tract7dt/fit/grid_gpu/error_model.py`` --
``_radial_profile`` (:155-171: ring means over round(sqrt(dy^2 + dx^2))),
``_acf_shape_is_resampling`` (:174-244: q2 = [rho_r(2) / rho_r(1)] /
[w_r(2) / w_r(1)], accept below 0.5, degenerate cases accept),
``measure_correlated_noise_factor`` (:933-1113, the raw-sky estimator:
central crop of at most 4096 px, median / 1.4826 MAD of the finite
unmasked pixels, source mask med + 3 MAD / med - 5 MAD dilated twice with
the 4-neighbour cross, at least 20 % clean sky, mask-normalised circular
autocovariance by FFT, F = sum rho(d) w(d) / w(0) over |dy|, |dx| <= L with
w the PSF autocorrelation, shape gate, F = max(F, 1)) and
``measure_correlated_noise_factor_residual`` (:1116-1302: the same
estimator on the residual map restricted to the cleanest quartile of its
256-px tiles, each kept tile detrended by its own clean mean; symmetric
3-MAD clip; at least 8 scored tiles and 20000 clean pixels); the per-band
choice of ``fit/grid_gpu/pipeline.py:2411-2458`` (clean-pixel-weighted
mean of the grids' valid residual factors, else the raw-sky factor, else 1).

The empty-aperture box-sum cross-check of :1058-1084 is a logged
diagnostic that feeds no output, and is not restated.

Parity pins (tests/test_fcorr.py): the FFT autocovariance equals a direct
circular sum on a small map; white noise gives F = 1 within its sampling
error; noise smoothed by a known kernel h gives F = sum rho_h(d) w(d) /
w(0) with rho_h the kernel's own normalised autocorrelation, and a shape
gate that accepts; noise smoothed by the PSF itself (the confusion shape)
gives q2 near 1 and the fallback F = 1; the tile selection keeps the
quartile of lowest variance.
"""
from __future__ import annotations

import numpy as np

CLIP_NSIGMA = 3.0
LOW_NSIGMA = 5.0
DILATE = 2
MAX_CROP = 4096
MIN_SKY_FRAC = 0.20
RES_TILE = 256
RES_KEEP_Q = 0.25
RES_MIN_TILES = 8
RES_MIN_CLEAN = 20000


def radial_profile(a, L):
    """Ring means of a (2L+1)^2 lag map, ring k = round(|d|) (:155-171)."""
    a = np.asarray(a, np.float64)
    yy, xx = np.mgrid[-L:L + 1, -L:L + 1]
    rr = np.rint(np.sqrt(yy * yy + xx * xx)).astype(int)
    prof = np.full(L + 1, np.nan)
    for k in range(L + 1):
        v = a[rr == k]
        v = v[np.isfinite(v)]
        if v.size:
            prof[k] = v.mean()
    return prof


def shape_gate(rho, w, L):
    """(is_resampling, q2, label) of :174-244."""
    if L < 2:
        return True, np.nan, "window-too-small"
    rr = radial_profile(rho, L)
    wr = radial_profile(w, L)
    r1, w1, r2, w2 = rr[1], wr[1], rr[2], wr[2]
    if not (np.isfinite(r1) and np.isfinite(w1) and w1 > 0 and np.isfinite(r2)
            and np.isfinite(w2) and w2 > 0):
        return True, np.nan, "degenerate"
    if not r1 > 0:
        return True, np.nan, "no-short-lag-corr"
    q2 = float((r2 / r1) / (w2 / w1))
    ok = bool(np.isfinite(q2) and q2 < 0.5)
    return ok, q2, ("resampling" if ok else "confusion")


def autocov_window(x, m, L):
    """rho over |dy|, |dx| <= L from the circular, mask-normalised FFT
    autocovariance (:1027-1052); returns (rho, C0)."""
    H, W = x.shape
    Fx = np.fft.rfft2(x)
    Fm = np.fft.rfft2(m.astype(np.float64))
    A = np.fft.irfft2(Fx * np.conj(Fx), s=(H, W))
    K = np.fft.irfft2(Fm * np.conj(Fm), s=(H, W))
    C = A / np.maximum(K, 1.0)
    C0 = float(C[0, 0])
    d = np.arange(-L, L + 1)
    return C[np.mod(d[:, None], H), np.mod(d[None, :], W)] / C0, C0


def psf_autocorr(phi, L):
    """w(d) / w(0) over |dy|, |dx| <= L, w = phi (corr) phi (:1043-1047)."""
    from scipy.signal import fftconvolve
    phi = np.asarray(phi, np.float64)
    phi = phi / phi.sum()
    wf = fftconvolve(phi, phi[::-1, ::-1], mode="full")
    cy, cx = wf.shape[0] // 2, wf.shape[1] // 2
    return wf[cy - L:cy + L + 1, cx - L:cx + L + 1] / wf[cy, cx]


def _robust(v):
    med = float(np.median(v))
    mad = 1.4826 * float(np.median(np.abs(v - med)))
    if not np.isfinite(mad) or mad <= 0:
        mad = float(np.std(v))
    return med, mad


def _finish(x, m, phi, H, W, diag):
    """F from the masked map x, the mask m and the stamp phi."""
    psum = float(np.sum(phi))
    rho, C0 = autocov_window(x, m, int(min(np.shape(phi)[0] - 1, H // 2 - 1, W // 2 - 1)))
    if not (np.isfinite(C0) and C0 > 0):
        diag["fallback"] = "non-positive variance"
        return 1.0, diag
    if not (np.isfinite(psum) and psum > 0):
        diag["fallback"] = "invalid PSF stamp"
        return 1.0, diag
    L = int(min(np.shape(phi)[0] - 1, H // 2 - 1, W // 2 - 1))
    w = psf_autocorr(phi, L)
    F = float((rho * w).sum())
    diag.update(F_acf=F, rho_1px=0.5 * (float(rho[L, L + 1]) + float(rho[L + 1, L])), L=L)
    if not np.isfinite(F):
        diag["fallback"] = "non-finite F_corr"
        return 1.0, diag
    ok, q2, label = shape_gate(rho, w, L)
    diag.update(q2=q2, acf_shape=label)
    if not ok:
        diag["fallback"] = "confusion-shaped ACF"
        return 1.0, diag
    return float(max(F, 1.0)), diag


def fcorr_sky(img, bad, phi):
    """Raw-sky F_corr of one band (:933-1113); returns (F, diag)."""
    from scipy.ndimage import binary_dilation
    img = np.asarray(img)
    H, W = img.shape
    ch, cw = int(min(H, MAX_CROP)), int(min(W, MAX_CROP))
    y0, x0 = (H - ch) // 2, (W - cw) // 2
    sub = np.asarray(img[y0:y0 + ch, x0:x0 + cw], np.float64)
    badc = (np.asarray(bad[y0:y0 + ch, x0:x0 + cw], bool) if bad is not None
            else ~np.isfinite(sub))
    fin = np.isfinite(sub) & ~badc
    diag = {"fallback": None, "n_clean": 0}
    if fin.sum() < MIN_SKY_FRAC * fin.size:
        diag["fallback"] = "too few finite sky pixels"
        return 1.0, diag
    med, mad = _robust(sub[fin])
    if not np.isfinite(mad) or mad <= 0:
        diag["fallback"] = "degenerate sky"
        return 1.0, diag
    with np.errstate(invalid="ignore"):
        src = ~fin | (sub > med + CLIP_NSIGMA * mad) | (sub < med - LOW_NSIGMA * mad)
    g = ~binary_dilation(src, iterations=DILATE)
    diag.update(med=med, mad=mad, n_clean=int(g.sum()))
    if g.sum() < MIN_SKY_FRAC * g.size:
        diag["fallback"] = "too few clean sky pixels"
        return 1.0, diag
    x = np.where(g, sub - float(sub[g].mean()), 0.0)
    return _finish(x, g, phi, ch, cw, diag)


def tile_selection(r, clean):
    """Per-tile clean variance, the kept quartile and its mask (:1219-1254);
    returns (keep_tile, tile_var, n_finite)."""
    H, W = r.shape
    ts = RES_TILE
    ny, nx = (H + ts - 1) // ts, (W + ts - 1) // ts
    tv = np.full((ny, nx), np.inf)
    for ty in range(ny):
        for tx in range(nx):
            cm = clean[ty * ts:(ty + 1) * ts, tx * ts:(tx + 1) * ts]
            if int(cm.sum()) < 0.25 * cm.size:
                continue
            tv[ty, tx] = float(r[ty * ts:(ty + 1) * ts, tx * ts:(tx + 1) * ts][cm].var())
    fin = np.isfinite(tv)
    if int(fin.sum()) < RES_MIN_TILES:
        return None, tv, int(fin.sum())
    thr = float(np.quantile(tv[fin], RES_KEEP_Q))
    return fin & (tv <= thr), tv, int(fin.sum())


def fcorr_residual(resid, good, phi):
    """Residual-map F_corr of one (grid, band) (:1116-1302); returns
    (F, diag) with diag['n_clean'] (0 on an early fallback)."""
    from scipy.ndimage import binary_dilation
    r = np.asarray(resid, np.float64)
    H, W = r.shape
    g0 = (np.asarray(good, bool) & np.isfinite(r)) if good is not None else np.isfinite(r)
    diag = {"fallback": None, "n_clean": 0}
    if g0.sum() < RES_MIN_CLEAN:
        diag["fallback"] = "too few good residual pixels"
        return 1.0, diag
    med, mad = _robust(r[g0])
    if not np.isfinite(mad) or mad <= 0:
        diag["fallback"] = "degenerate residual"
        return 1.0, diag
    with np.errstate(invalid="ignore"):
        struct = ~g0 | (np.abs(r - med) > CLIP_NSIGMA * mad)
    clean = ~binary_dilation(struct, iterations=DILATE)
    keep, _, _ = tile_selection(r, clean)
    diag.update(med=med, mad=mad)
    if keep is None:
        diag["fallback"] = "too few clean tiles"
        return 1.0, diag
    ts = RES_TILE
    tile_ok = np.repeat(np.repeat(keep, ts, 0), ts, 1)[:H, :W]
    m = clean & tile_ok
    n_clean = int(m.sum())
    diag.update(n_clean=n_clean, n_tiles_kept=int(keep.sum()))
    if n_clean < RES_MIN_CLEAN:
        diag["fallback"] = "too few clean tile pixels"
        return 1.0, diag
    r_dt = r.copy()
    for ty, tx in zip(*np.nonzero(keep)):
        sy, sx = slice(ty * ts, (ty + 1) * ts), slice(tx * ts, (tx + 1) * ts)
        if m[sy, sx].any():
            r_dt[sy, sx] = r[sy, sx] - float(r[sy, sx][m[sy, sx]].mean())
    x = np.where(m, r_dt, 0.0)
    return _finish(x, m, phi, H, W, diag)


def choose_band_fcorr(per_grid, raw):
    """Final per-band F_corr (pipeline.py:2411-2458): ``per_grid`` is a list
    over grids of lists over bands of (F, n_clean, fallback) or None; ``raw``
    the raw-sky factor per band.  Returns (F, source) per band."""
    nb = len(raw)
    acc = np.zeros(nb)
    cnt = np.zeros(nb, np.int64)
    for fr in per_grid:
        for b in range(min(nb, len(fr or []))):
            if fr[b] is None:
                continue
            Fv, nc, fb = fr[b]
            if fb is not None or not np.isfinite(Fv) or Fv <= 1.0 or nc <= 0:
                continue
            acc[b] += float(Fv) * int(nc)
            cnt[b] += int(nc)
    out, src = [], []
    for b in range(nb):
        fres = acc[b] / cnt[b] if cnt[b] > 0 else np.nan
        if np.isfinite(fres) and fres > 1.0:
            out.append(float(fres)); src.append("residual")
        elif np.isfinite(raw[b]) and raw[b] > 1.0:
            out.append(float(raw[b])); src.append("raw-sky")
        else:
            out.append(1.0); src.append("none")
    return np.array(out), src
