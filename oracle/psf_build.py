"""FP64 restatement of the PSF model build from training stars (TEST
INFRASTRUCTURE: the checker of the device path, never the product).

Reference: ``tract7dt/fit/grid_gpu/psf_from_stars.py`` --
``fourier_shift`` (:1479-1507, shift theorem on the image zero-padded by 16
px, cut back), ``_gw_centroid_vec`` (:1510-1575, Gaussian-weighted first
moments of the non-negative data, at most 8 iterations, all stars stop
together when the largest move is below 1e-3 px), ``_extract_centered_stamps_gw``
(:1578-1720: 61-px cutouts at the rounded positions, dropped off the frame or
on a bad pixel; sigma-clipped annulus median background; recentred cutout at
the rounded centroid, shifted by minus its fractional part, cropped to 55 px
and normalised to unit sum), ``_annulus_median`` (:1957-1984, float32 values
in r^2 in [20^2, 30^2], median, 1.4826 MAD, 3-MAD clip, at least 30 pixels),
``_fit_polynomial_one_iter`` / ``_per_star_chi2`` /
``fit_polynomial_with_outlier_rejection`` (:2019-2239: per-pixel least
squares on the (x, y) polynomial over (x - W/2) / (W/2), (y - H/2) / (H/2),
coefficients rounded to float32, chi = sqrt(sum of squared residuals), reject
chi > median + 3 x 1.4826 MAD, at most 4 passes) and
``_replace_model_wings_with_median`` (:2082-2147: beyond max(6, 2.5 FWHM) the
constant term blends over 4 px to the pixelwise median of the accepted
stamps, the spatial terms fade, the constant term keeps its sum; skipped
below 20 stars).

Parity pins (tests/test_psf_build.py): an integer shift is an exact move;
the Dirichlet form of the shift equals the FFT form; the centroid of a
noiseless Gaussian star lands on its centre; stamps rendered from a known
polynomial PSF are fitted back to it, with an injected bad star rejected;
the wing blend keeps the constant term's sum and leaves the core alone; the
annulus median of a constant sky is that constant.
"""
from __future__ import annotations

import numpy as np

FOURIER_PAD = 16


def terms_for_order(order: int):
    """(m, n) exponent pairs of x^m y^n, by total degree then descending m."""
    return [(d - k, k) for d in range(order + 1) for k in range(d + 1)]


def fourier_shift(img, sx, sy, pad=FOURIER_PAD):
    """out(p) = img(p - s) on the image zero-padded by ``pad`` (:1479-1507)."""
    img = np.asarray(img, np.float64)
    n0, n1 = img.shape
    m0, m1 = n0 + 2 * pad, n1 + 2 * pad
    P = np.zeros((m0, m1))
    P[pad:pad + n0, pad:pad + n1] = img
    fy = np.fft.fftfreq(m0)[:, None]
    fx = np.fft.fftfreq(m1)[None, :]
    out = np.fft.ifft2(np.fft.fft2(P) * np.exp(-2j * np.pi * (fx * sx + fy * sy))).real
    return out[pad:pad + n0, pad:pad + n1]


def annulus_median(image, ix, iy, rin, rout, sigma, minpix, bad):
    """Sigma-clipped annulus median in float32 (:1957-1984)."""
    H, W = image.shape
    ri = int(np.ceil(rout))
    y0, y1 = max(0, iy - ri), min(H, iy + ri + 1)
    x0, x1 = max(0, ix - ri), min(W, ix + ri + 1)
    sub = image[y0:y1, x0:x1]
    yy, xx = np.mgrid[:sub.shape[0], :sub.shape[1]]
    r2 = (yy - (iy - y0)) ** 2 + (xx - (ix - x0)) ** 2
    ann = (r2 >= rin * rin) & (r2 <= rout * rout) & ~bad[y0:y1, x0:x1] & np.isfinite(sub)
    vals = sub[ann]
    if vals.size < minpix:
        return float("nan")
    med = float(np.median(vals))
    mad = float(np.median(np.abs(vals - med))) * 1.4826
    if mad <= 0:
        return med
    keep = (vals >= med - sigma * mad) & (vals <= med + sigma * mad)
    if keep.sum() < minpix:
        return med
    return float(np.median(vals[keep]))


def gw_centroid(cutouts, init_xy, fwhm, max_iters=8, tol=1e-3):
    """(cx, cy, iterations) of :1510-1575."""
    N, H, W = cutouts.shape
    s = float(fwhm) / 2.354820045
    k = 1.0 / (2.0 * s * s)
    yy, xx = np.mgrid[:H, :W].astype(np.float64)
    cx = init_xy[:, 0].astype(np.float64).copy()
    cy = init_xy[:, 1].astype(np.float64).copy()
    d = np.maximum(cutouts.astype(np.float64), 0.0)
    it = 0
    for it in range(1, max_iters + 1):
        dx = xx[None] - cx[:, None, None]
        dy = yy[None] - cy[:, None, None]
        wd = np.exp(-(dx * dx + dy * dy) * k) * d
        S = wd.sum(axis=(1, 2))
        Ss = np.where(S > 0, S, 1.0)
        nx = np.where(S > 0, (wd * xx[None]).sum(axis=(1, 2)) / Ss, cx)
        ny = np.where(S > 0, (wd * yy[None]).sum(axis=(1, 2)) / Ss, cy)
        delta = float(np.max(np.abs(np.concatenate([nx - cx, ny - cy]))))
        cx, cy = nx, ny
        if delta < tol:
            break
    return cx, cy, it


def extract_centered_stamps(image, bad, star_x, star_y, *, cutout_size=61, psf_size=55,
                            fwhm=5.0, bkg=True, rin=20.0, rout=30.0, bsig=3.0, minpix=30):
    """(stamps, fitted_x, fitted_y, kept) of :1578-1720."""
    image = np.asarray(image, np.float32)
    bad = np.asarray(bad, bool)
    H, W = image.shape
    h = cutout_size // 2
    c0 = (cutout_size - psf_size) // 2
    ox = np.round(star_x).astype(np.int64)
    oy = np.round(star_y).astype(np.int64)
    inf = (ox - h >= 0) & (ox + h + 1 <= W) & (oy - h >= 0) & (oy + h + 1 <= H)
    keep = np.nonzero(inf)[0]
    keep = np.array([i for i in keep
                     if not bad[oy[i] - h:oy[i] + h + 1, ox[i] - h:ox[i] + h + 1].any()],
                    np.int64)
    empty = (np.zeros((0, psf_size, psf_size), np.float32), np.zeros(0), np.zeros(0),
             np.zeros(0, np.int64))
    if keep.size == 0:
        return empty
    cut = np.stack([image[oy[i] - h:oy[i] + h + 1, ox[i] - h:ox[i] + h + 1] for i in keep])
    bk = np.full(keep.size, np.nan)
    if bkg:
        bk = np.array([annulus_median(image, int(ox[i]), int(oy[i]), rin, rout, bsig, minpix,
                                      bad) for i in keep])
        ok = np.isfinite(bk)
        cut = cut.copy()
        cut[ok] -= bk[ok, None, None].astype(np.float32)
    cx, cy, _ = gw_centroid(cut, np.full((keep.size, 2), float(h)), fwhm)
    fx_all = cx - h + ox[keep]
    fy_all = cy - h + oy[keep]
    st, fx, fy, kk = [], [], [], []
    for k, i in enumerate(keep):
        xc, yc = float(fx_all[k]), float(fy_all[k])
        if not (np.isfinite(xc) and np.isfinite(yc)):
            continue
        ix, iy = int(round(xc)), int(round(yc))
        dx, dy = xc - ix, yc - iy
        if iy - h < 0 or ix - h < 0 or iy + h + 1 > H or ix + h + 1 > W:
            continue
        if bad[iy - h:iy + h + 1, ix - h:ix + h + 1].any():
            continue
        c = image[iy - h:iy + h + 1, ix - h:ix + h + 1].astype(np.float32, copy=True)
        if bkg:
            lb = (bk[k] if (ix == ox[i] and iy == oy[i])
                  else annulus_median(image, ix, iy, rin, rout, bsig, minpix, bad))
            if np.isfinite(lb):
                c = c - np.float32(lb)
        crop = fourier_shift(c, -dx, -dy)[c0:c0 + psf_size, c0:c0 + psf_size]
        s = float(crop.sum())
        if not np.isfinite(s) or s <= 0:
            continue
        st.append((crop / s).astype(np.float32))
        fx.append(xc)
        fy.append(yc)
        kk.append(int(i))
    if not st:
        return empty
    return np.stack(st), np.array(fx), np.array(fy), np.array(kk, np.int64)


def _design(x, y, shape, order):
    H, W = shape
    X = (np.asarray(x, np.float64) - 0.5 * W) / (0.5 * W)
    Y = (np.asarray(y, np.float64) - 0.5 * H) / (0.5 * H)
    return np.stack([X ** m * Y ** n for (m, n) in terms_for_order(order)], axis=1)


def fit_polynomial(star_x, star_y, stamps, shape, order=2, max_outer=4, kappa=3.0,
                   rcond=1e-9):
    """(coefs (Hp, Wp, n_terms) float32, accepted, iterations, chi2) of
    :2019-2239; the model's geometry is (W/2, W/2, H/2, H/2)."""
    n, Hp, Wp = stamps.shape
    nt = len(terms_for_order(order))
    acc = np.ones(n, bool)
    A_all = _design(star_x, star_y, shape, order)
    Yall = stamps.reshape(n, -1).astype(np.float64)
    coefs, chi2, it = None, np.zeros(n), 0
    for it in range(1, max_outer + 1):
        if acc.sum() < nt:            # the pass counts, as in the reference
            break
        cf, *_ = np.linalg.lstsq(A_all[acc], Yall[acc], rcond=rcond)
        coefs = cf.T.reshape(Hp, Wp, nt).astype(np.float32)
        model = A_all @ coefs.reshape(Hp * Wp, nt).astype(np.float64).T
        chi2 = np.sum((Yall - model) ** 2, axis=1)
        chi = np.sqrt(np.maximum(chi2, 0.0))
        ca = chi[acc]
        if ca.size < 3:
            break
        med = float(np.median(ca))
        mad = float(np.median(np.abs(ca - med))) * 1.4826
        if mad <= 0:
            break
        new = chi <= med + kappa * mad
        changed = bool(((acc & ~new) | (~acc & new)).any())
        acc = new
        if not changed:
            break
    if coefs is None:
        raise RuntimeError("polynomial fit failed (no iterations completed)")
    return coefs, acc, it, chi2


def wing_blend(coefs, stamps, accepted, fwhm, start_fwhm=2.5, width=4.0):
    """Median-wing surgery of :2082-2147 (returns float32 coefs)."""
    st = stamps[accepted] if accepted is not None else stamps
    if st.shape[0] < 20:
        return coefs
    med = np.median(st.astype(np.float64), axis=0)
    c = coefs.astype(np.float64).copy()
    Hp, Wp, nt = c.shape
    yy, xx = np.mgrid[0:Hp, 0:Wp]
    r = np.hypot(yy - (Hp - 1) / 2.0, xx - (Wp - 1) / 2.0)
    fw = float(fwhm) if np.isfinite(fwhm) and fwhm > 0 else 5.0
    r0 = max(6.0, start_fwhm * fw)
    w = np.clip((r - r0) / max(width, 1e-6), 0.0, 1.0)
    c0 = c[:, :, 0]
    target = float(c0.sum())
    new = (1.0 - w) * c0 + w * med
    s = float(new.sum())
    if s > 0 and np.isfinite(s):
        new *= target / s
    c[:, :, 0] = new
    if nt > 1:
        c[:, :, 1:] *= (1.0 - w)[:, :, None]
    return c.astype(np.float32)
