"""PSF model build from training stars: the producer of the per-band
``PsfPolynomial`` the fit reads.

Mirrors the numerical core of ``tract7dt/fit/grid_gpu/psf_from_stars.py``
``build_psf_polynomial_from_stars_for_band`` (:2242-2370) from the training
stars on: ``extract_centered_stamps`` with the default ``gw_centroid``
method (:1578-1720), ``fit_polynomial_with_outlier_rejection`` (:2150-2239)
and ``_replace_model_wings_with_median`` (:2082-2147).  The star selection
before it (SEP detection, the auto-FWHM mode, Gaia identity rows,
isolation and saturation cuts, :117-766) needs the SEP library and a
reference catalogue and is the caller's: it hands over the training star
positions and the band FWHM.

The device does the pixel work (``uph_psf_*``): cutouts with the
sigma-clipped annulus background, the Gaussian-weighted centroid
iterations, the recentred Fourier-shifted unit-sum stamps (the shift as the
separable Dirichlet kernel of the zero-padded box), the per-pixel polynomial
solve with the pseudo-inverse of the design matrix, the per-star residuals
and the pixelwise median of the accepted stamps.  The host takes the
iteration count at which every star's centroid has settled, the (n_stars x
n_terms) pseudo-inverse by SVD with numpy's lstsq cutoff, the median / MAD
star rejection and the wing blend of the 55 x 55 constant term.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .psf_polynomial import PsfPolynomial, terms_for_order

FOURIER_PAD = 16


@dataclass
class PsfBuildResult:
    psf: PsfPolynomial
    stamps: np.ndarray        # (n_kept, psf_size, psf_size) float32 unit-sum stamps
    fitted_x: np.ndarray      # (n_kept,) centroids in image pixels
    fitted_y: np.ndarray
    kept: np.ndarray          # (n_kept,) indices into the input stars
    accepted: np.ndarray      # (n_kept,) bool: used by the final polynomial fit
    n_iters: int              # outer rejection passes
    chi2: np.ndarray          # (n_kept,) per-star sum of squared residuals
    centroid_iters: int


def _design(x, y, H, W, order):
    X = (np.asarray(x, np.float64) - 0.5 * W) / (0.5 * W)
    Y = (np.asarray(y, np.float64) - 0.5 * H) / (0.5 * H)
    return np.stack([X ** m * Y ** n for (m, n) in terms_for_order(order)], axis=1)


def build_psf_from_stars(image, bad, star_x, star_y, *, fwhm: float = 5.0,
                         polynomial_order: int = 4, cutout_size: int = 61, psf_size: int = 55,
                         local_bkg: bool = True, bkg_rin: float = 20.0, bkg_rout: float = 30.0,
                         bkg_sigma: float = 3.0, bkg_minpix: int = 30, max_outer_iters: int = 4,
                         kappa: float = 3.0, rcond: float = 1e-9, wing_from_median: bool = True,
                         wing_blend_start_fwhm: float = 2.5, centroid_iters: int = 8,
                         centroid_tol: float = 1e-3) -> PsfBuildResult:
    """The band's PsfPolynomial from training stars at (star_x, star_y):
    ``image`` (H, W) float32 and ``bad`` (H, W) uint8 device tensors.  Raises
    RuntimeError when fewer stars than polynomial terms survive the
    extraction (psf_from_stars.py:2343-2347)."""
    import torch
    from . import _lib
    L = _lib.lib()
    P_ = _lib.ptr
    dev = image.device
    H, W = int(image.shape[0]), int(image.shape[1])
    stream = _lib.stream_ptr(None)
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    half = cutout_size // 2
    C = 2 * half + 1
    sx = np.asarray(star_x, np.float64)
    sy = np.asarray(star_y, np.float64)
    n = sx.size
    ox = np.rint(sx).astype(np.int64)
    oy = np.rint(sy).astype(np.int64)
    nt = len(terms_for_order(polynomial_order))
    if n == 0:
        raise RuntimeError(f"only 0 stars extracted; need at least {nt} for order "
                           f"{polynomial_order}")
    # int32 pixel positions; far-off stars are off the frame either way
    oxi = np.clip(ox, -(1 << 30), 1 << 30).astype(np.int32)
    oyi = np.clip(oy, -(1 << 30), 1 << 30).astype(np.int32)
    d_ox, d_oy = up(oxi, np.int32), up(oyi, np.int32)
    st0 = torch.empty(n, dtype=torch.uint8, device=dev)
    bkg = torch.empty(n, dtype=torch.float64, device=dev)
    cut = torch.empty((n, C, C), dtype=torch.float32, device=dev)
    _lib.check(L.uph_psf_cutouts(n, H, W, P_(image), P_(bad), P_(d_ox), P_(d_oy), half,
                                 bkg_rin, bkg_rout, bkg_sigma, bkg_minpix, int(local_bkg),
                                 P_(st0), P_(bkg), P_(cut), stream))
    k0 = np.nonzero(st0.cpu().numpy() == 0)[0]
    if k0.size == 0:
        raise RuntimeError(f"only 0 stars extracted; need at least {nt} for order "
                           f"{polynomial_order}")
    d_k0 = torch.from_numpy(k0).to(dev)
    cut0 = cut.index_select(0, d_k0).contiguous()
    hist = torch.empty((k0.size, centroid_iters, 3), dtype=torch.float64, device=dev)
    _lib.check(L.uph_psf_centroids(int(k0.size), half, float(fwhm), centroid_iters, P_(cut0),
                                   P_(hist), stream))
    hs = hist.cpu().numpy()
    # every star stops at the first iteration whose largest move is below the tolerance
    settled = np.nonzero(hs[:, :, 2].max(axis=0) < centroid_tol)[0]
    n_it = int(settled[0]) + 1 if settled.size else centroid_iters
    fx0 = hs[:, n_it - 1, 0] - half + ox[k0]
    fy0 = hs[:, n_it - 1, 1] - half + oy[k0]
    st1 = torch.empty(k0.size, dtype=torch.uint8, device=dev)
    stamps = torch.empty((k0.size, psf_size, psf_size), dtype=torch.float32, device=dev)
    keep = [up(oxi[k0], np.int32), up(oyi[k0], np.int32), up(fx0, np.float64),
            up(fy0, np.float64), bkg.index_select(0, d_k0).contiguous()]
    _lib.check(L.uph_psf_stamps(int(k0.size), H, W, P_(image), P_(bad), P_(keep[0]),
                                P_(keep[1]), P_(keep[2]), P_(keep[3]), P_(keep[4]), half,
                                psf_size, FOURIER_PAD, bkg_rin, bkg_rout, bkg_sigma, bkg_minpix,
                                int(local_bkg), P_(st1), P_(stamps), stream))
    ok = np.nonzero(st1.cpu().numpy() == 0)[0]
    if ok.size < nt:
        raise RuntimeError(f"only {ok.size} stars extracted; need at least {nt} for order "
                           f"{polynomial_order}")
    d_ok = torch.from_numpy(ok).to(dev)
    stk = stamps.index_select(0, d_ok).contiguous()
    fx, fy, kept = fx0[ok], fy0[ok], k0[ok]
    nk = ok.size
    Pn = psf_size * psf_size
    A_all = _design(fx, fy, H, W, polynomial_order)
    d_A = up(A_all, np.float64)
    coefs = torch.empty((Pn, nt), dtype=torch.float32, device=dev)
    chi2_d = torch.empty(nk, dtype=torch.float64, device=dev)
    acc = np.ones(nk, bool)
    chi2 = np.zeros(nk)
    have, it = False, 0
    for it in range(1, max_outer_iters + 1):
        if acc.sum() < nt:            # the pass counts, as in the reference
            break
        sel = np.nonzero(acc)[0]
        pinv = np.linalg.pinv(A_all[sel], rcond=rcond)           # (nt, ns), lstsq's cutoff
        d_sel, d_pinv = up(sel, np.int32), up(pinv, np.float64)
        _lib.check(L.uph_psf_polyfit(int(sel.size), nt, Pn, P_(d_sel), P_(d_pinv), P_(stk),
                                     P_(coefs), stream))
        _lib.check(L.uph_psf_chi2(nk, nt, Pn, P_(d_A), P_(stk), P_(coefs), P_(chi2_d), stream))
        have = True
        chi2 = chi2_d.cpu().numpy().copy()
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
    if not have:
        raise RuntimeError("polynomial fit failed (no iterations completed)")
    cf = coefs.cpu().numpy().reshape(psf_size, psf_size, nt)
    if wing_from_median and int(acc.sum()) >= 20:
        sel = np.nonzero(acc)[0]
        med = torch.empty(Pn, dtype=torch.float64, device=dev)
        d_sel = up(sel, np.int32)
        _lib.check(L.uph_psf_pixel_median(int(sel.size), Pn, P_(d_sel), P_(stk), P_(med),
                                          stream))
        cf = _wing_blend(cf, med.cpu().numpy().reshape(psf_size, psf_size), fwhm,
                         wing_blend_start_fwhm)
    psf = PsfPolynomial(cf.astype(np.float32), 0.5 * W, 0.5 * W, 0.5 * H, 0.5 * H, psf_size,
                        psf_size)
    return PsfBuildResult(psf, stk.cpu().numpy(), fx, fy, kept.astype(np.int64), acc, it, chi2,
                          n_it)


def _wing_blend(coefs, med, fwhm, start_fwhm, width=4.0):
    """Constant term blended to the pixelwise median beyond max(6, start x
    FWHM) over ``width`` px, spatial terms faded, constant-term sum kept
    (psf_from_stars.py:2082-2147)."""
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


def save_psf_polynomial(out_dir, band: str, result: PsfBuildResult, polynomial_order: int):
    """``{out_dir}/{band}/psf_polynomial.npz`` in the layout the reference's
    stage writes and its loader reads (psf_pipeline_stage.py:555-567,
    :1066-1084): coefs f32, the four geometry scalars f32, H_psf / W_psf /
    polynomial_order i32 and the accepted training stars f64."""
    from pathlib import Path
    d = Path(out_dir) / band
    d.mkdir(parents=True, exist_ok=True)
    p = result.psf
    np.savez(d / "psf_polynomial.npz", coefs=p.coefs.astype(np.float32),
             x_origin=np.float32(p.x_origin), x_scale=np.float32(p.x_scale),
             y_origin=np.float32(p.y_origin), y_scale=np.float32(p.y_scale),
             H_psf=np.int32(p.H_psf), W_psf=np.int32(p.W_psf),
             polynomial_order=np.int32(polynomial_order),
             star_x=np.asarray(result.fitted_x[result.accepted], np.float64),
             star_y=np.asarray(result.fitted_y[result.accepted], np.float64))
    return d / "psf_polynomial.npz"


def load_psf_polynomial_for_band(psf_dir, band: str) -> PsfPolynomial:
    """The band's PsfPolynomial from ``{psf_dir}/{band}/psf_polynomial.npz``
    (psf_pipeline_stage.py:1066-1084); FileNotFoundError when absent."""
    from pathlib import Path
    path = Path(psf_dir) / band / "psf_polynomial.npz"
    if not path.is_file():
        raise FileNotFoundError(f"PSF polynomial not found for band {band}: {path}")
    with np.load(path, allow_pickle=False) as z:
        return PsfPolynomial(z["coefs"].astype(np.float32), float(z["x_origin"]),
                             float(z["x_scale"]), float(z["y_origin"]), float(z["y_scale"]),
                             int(z["H_psf"]), int(z["W_psf"]))
