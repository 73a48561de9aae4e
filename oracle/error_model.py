"""FP64 restatement of the confusion-noise flux error (TEST INFRASTRUCTURE).

Reference: ``tract7dt/fit/grid_gpu/error_model.py`` -- ``_boxsum`` /
``_annulus_mean`` (:124-145, edge-truncated square boxes, >= 50 valid
pixels), ``_corr_same`` (:148-152, 'same' cross-correlation), the per-band
numerics ``_confusion_band_numeric`` (:268-325: matched filter
MF = corr(r, phi) / corr(g, phi^2) and Vstat = corr(var, phi^2) /
corr(g, phi^2)^2 where corr(g, phi^2) > 0.6 sum(phi^2); annulus
[ceil(2 FWHM), ceil(6 FWHM)] with FWHM = 2.355 sqrt(1 / (4 pi sum(phi^2)));
excess = max(0, <MF^2> - <MF>^2 - <Vstat>) at the source pixel; the
extended-template factor R) and its driver ``compute_grid_confusion``
(:663-930: model = fluxes x unit stamps + wings + background, residual
on good pixels, phi = the band PSF at the grid centre over its sum).

Parity pins (tests/test_error_model.py): the FFT correlation equals a
direct 'same' correlation; an unmodelled point source of flux F shaped
like phi gives MF = F at its pixel; a zero or a constant residual gives
zero excess; an interior annulus holds (2 r_out + 1)^2 - (2 r_in + 1)^2
pixels and a star's R is 1; white noise of the stated variance gives an
excess that is small against Vstat.
"""
from __future__ import annotations

import math

import numpy as np

OCC_MIN = 0.6
R_IN_FWHM, R_OUT_FWHM = 2.0, 6.0
MIN_ANNULUS_PX = 50
GAUSS_HALF_LIGHT = 1.1774


def corr_same(a, k):
    """'same' cross-correlation: out[i, j] = sum a[i - c + u, j - c + v] k[u, v]."""
    from scipy.signal import fftconvolve
    return fftconvolve(np.asarray(a, np.float64), np.asarray(k, np.float64)[::-1, ::-1],
                       mode="same")


def boxsum(a, r: int):
    """Sum over the (2r + 1)^2 box around each pixel, clipped to the map."""
    H, W = a.shape
    c = np.zeros((H + 1, W + 1))
    c[1:, 1:] = np.cumsum(np.cumsum(a, axis=0), axis=1)
    y = np.arange(H)
    x = np.arange(W)
    y0, y1 = np.clip(y - r, 0, H), np.clip(y + r + 1, 0, H)
    x0, x1 = np.clip(x - r, 0, W), np.clip(x + r + 1, 0, W)
    return c[y1][:, x1] - c[y1][:, x0] - c[y0][:, x1] + c[y0][:, x0]


def annulus_mean(a, valid, r_in: int, r_out: int):
    n = boxsum(valid, r_out) - boxsum(valid, r_in)
    s = boxsum(a * valid, r_out) - boxsum(a * valid, r_in)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(n >= MIN_ANNULUS_PX, s / np.maximum(n, 1.0), np.nan), n


def psf_scale(phi):
    """(sum(phi^2) in float32, Gaussian-equivalent variance, r_in, r_out)."""
    sumsq = np.float32(np.sum(np.asarray(phi, np.float32) ** 2, dtype=np.float32))
    sp2 = 1.0 / (4.0 * math.pi * float(sumsq))
    fwhm = 2.355 * math.sqrt(sp2)
    return sumsq, sp2, int(math.ceil(R_IN_FWHM * fwhm)), int(math.ceil(R_OUT_FWHM * fwhm))


def confusion_maps(data, model, good, invvar, phi):
    """MF, Vstat and the valid mask of one band window."""
    good = np.asarray(good, bool)
    sumsq, _, _, _ = psf_scale(phi)
    phi = np.asarray(phi, np.float64)
    resid = np.where(good, np.asarray(data, np.float64) - np.asarray(model, np.float64), 0.0)
    iv = np.asarray(invvar, np.float64)
    var = np.where(good & (iv > 0), 1.0 / np.maximum(iv, 1e-30), 0.0)
    num = corr_same(resid, phi)
    den = corr_same(good.astype(np.float64), phi * phi)
    vnum = corr_same(var, phi * phi)
    valid = (den > OCC_MIN * float(sumsq)) & np.isfinite(num)
    ds = np.maximum(den, 1e-30)
    mf = np.where(valid, num / ds, 0.0)
    vs = np.where(valid, vnum / (ds * ds), 0.0)
    return mf, vs, valid


def confusion_sigma(data, model, good, invvar, phi, xi, yi, s_gal2, no_fit):
    """sigma_conf per source of one band window (NaN where the annulus has
    fewer than 50 valid pixels or the source has no fitted position)."""
    _, sp2, r_in, r_out = psf_scale(phi)
    mf, vs, valid = confusion_maps(data, model, good, invvar, phi)
    vf = valid.astype(np.float64)
    m1, _ = annulus_mean(mf, vf, r_in, r_out)
    m2, _ = annulus_mean(mf * mf, vf, r_in, r_out)
    vsl, _ = annulus_mean(vs, vf, r_in, r_out)
    with np.errstate(invalid="ignore"):
        excess = np.maximum(0.0, (m2 - m1 * m1) - vsl)
        exc = excess[np.asarray(yi), np.asarray(xi)]
        sg = sp2 + np.asarray(s_gal2, np.float64)
        R = 2.0 * sg * sg / ((sg + sp2) * sp2)
        sig = np.sqrt(R * exc)
    return np.where(np.asarray(no_fit, bool), np.nan, sig)


def template_variance(stype, logre):
    """(Re / 1.1774)^2 of extended sources with a finite log Re, else 0."""
    ext = np.asarray(stype) != 0
    lr = np.asarray(logre, np.float64)
    re = np.where(ext & np.isfinite(lr), np.exp(np.clip(lr, -5.0, 5.0)), 0.0)
    return (re / GAUSS_HALF_LIGHT) ** 2
