"""Post-fit flux-error model: the confusion term and the error columns.

Mirrors ``tract7dt/fit/grid_gpu/error_model.py`` for the confusion
measurement of each (grid, band) after pass 2 (``compute_grid_confusion``,
:663-930, and its per-band numerics ``_confusion_band_numeric``,
:268-325): the residual of the converged model (fluxes x unit stamps,
the big sources' wings, the background plane) is matched-filtered with the
band's central PSF and the variance of that map in an annulus around each
source, above the statistical expectation, is the source's confusion
variance, scaled by the extended-template factor.  The device does the
whole per-problem computation (``uph_confusion``).

``add_error_columns`` follows ``error_columns.py:112-363`` for the flux
columns: FLUXERR_CONF_{b}_fit, FLUXERR_TOT_{b}_fit =
sqrt(FLUXERR^2 + CONF^2) and FLUXERR_MARG_TOT_{b}_fit =
sqrt(FLUXERR^2 + dvar + CONF^2) x sqrt(max(1, rchisq)).  The correlated-
noise factor F_corr (measured by :mod:`uniphot_amd.fcorr`) and the
per-island chi2_nu scale the geometric error columns (``*_err_corr``,
``*_err_tot``, ``add_geometric_error_columns``, error_columns.py:153-246)
and the marginal increment inside FLUXERR_MARG_TOT.  The source-Poisson
term needs the images' gain and is not part of this build: its variance is
0, as the reference's is when that leg is off.
"""
from __future__ import annotations

import math

import numpy as np

from . import forced_phot_sparse as fps
from .source_model import SOURCE_STAR

OCC_MIN = 0.6
R_IN_FWHM, R_OUT_FWHM = 2.0, 6.0
GAUSS_HALF_LIGHT = 1.1774
MAX_TEMPLATE = 63  # largest stamp the correlation launch stages in LDS


def psf_scale(phi32: np.ndarray):
    """sum(phi^2) (float32), the Gaussian-equivalent PSF variance and the
    annulus radii ceil(2 FWHM), ceil(6 FWHM) (error_model.py:293-300)."""
    sumsq = np.float32(np.sum(np.square(phi32, dtype=np.float32), dtype=np.float32))
    sp2 = 1.0 / (4.0 * math.pi * float(sumsq))
    fwhm = 2.355 * math.sqrt(sp2)
    return sumsq, sp2, int(math.ceil(R_IN_FWHM * fwhm)), int(math.ceil(R_OUT_FWHM * fwhm))


def confusion_sigma(fld, T, problems, owners, res, x, y, stype, logre, ee1, ee2, sn, cfg,
                    bg_geo, wing=None, use_wing=None, fcorr_out=None):
    """sigma_conf of every fit source of each (grid, band) problem after the
    final solve ``res`` (a list of per-problem arrays, NaN where the annulus
    is too masked or the position is not finite).  ``wing`` (nb, H, W) is
    subtracted from the data of the problems flagged in ``use_wing``.  With
    ``fcorr_out`` (a dict) the residual correlated-noise estimate of problem
    j (error_model.py:1116-1302 on the same residual and good maps and the
    same unit-sum window PSF) is stored under j as (F, n_clean, fallback)."""
    import torch
    from . import _lib
    from .pipeline import window_psf_stamps
    if T > MAX_TEMPLATE:
        raise ValueError(f"confusion: stamp {T} px exceeds the correlation's {MAX_TEMPLATE} px")
    L = _lib.lib()
    P = _lib.ptr
    dev = fld.data.device
    nb, H, W = fld.nb, fld.H, fld.W
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    cs, gs = fld.psf(T)
    out = [None] * len(problems)
    n_of = np.array([owners[q][0].fit_indices.size for q in range(len(problems))])
    cost = n_of * T * T * 4.0 + np.array([pr.h * pr.w * 24.0 for pr in problems])
    cut = np.concatenate([[0], np.nonzero(np.diff(np.floor(np.cumsum(cost) /
                                                         cfg.stamp_budget)))[0] + 1,
                          [len(problems)]])
    for a, z in zip(cut[:-1], cut[1:]):
        qs = np.arange(a, z)
        if qs.size == 0:
            continue
        src = np.concatenate([owners[q][0].fit_indices for q in qs]).astype(np.int64)
        counts = n_of[qs]
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        band = np.repeat([owners[q][1] for q in qs], counts).astype(np.int32)
        ox = np.repeat([problems[q].x0 for q in qs], counts).astype(np.float32)
        oy = np.repeat([problems[q].y0 for q in qs], counts).astype(np.float32)
        ext = stype[src] != SOURCE_STAR
        stamps = fps.render_unit_stamps(
            cs, gs, T, fld.n_terms, band, x[src], y[src], ox, oy, stype[src],
            np.where(ext, logre[src], 0.0), np.where(ext, ee1[src], 1.0),
            np.where(ext, ee2[src], 0.0), sn[src])
        ix, iy = fps.stamp_top_left(x[src], y[src], ox, oy, T, T)
        flux = np.concatenate([res[q].fluxes for q in qs]).astype(np.float32)
        flux = np.where(np.isfinite(flux), flux, 0.0).astype(np.float32)
        bg = np.stack([res[q].bg_coefs for q in qs]).astype(np.float32)
        win = np.array([[owners[q][1], problems[q].x0, problems[q].y0, problems[q].h,
                         problems[q].w] for q in qs], np.int32)
        fps.check_windows(win, nb, H, W, "confusion_sigma")
        phi = np.empty((qs.size, T, T), np.float32)
        sumsq = np.empty(qs.size, np.float32)
        sp2 = np.empty(qs.size, np.float64)
        radii = np.empty((qs.size, 2), np.int32)
        for j, t in enumerate(window_psf_stamps(fld, problems, owners, qs, T)):
            t = t.astype(np.float32)
            phi[j] = t / np.sum(t, dtype=np.float32)
            sumsq[j], sp2[j], radii[j, 0], radii[j, 1] = psf_scale(phi[j])
        # rounded window-local source pixel, clipped into the window
        wq = np.repeat([problems[q].w for q in qs], counts)
        hq = np.repeat([problems[q].h for q in qs], counts)
        with np.errstate(invalid="ignore"):
            xi = np.clip(np.rint(x[src].astype(np.float32) - ox), 0, wq - 1)
            yi = np.clip(np.rint(y[src].astype(np.float32) - oy), 0, hq - 1)
        no_fit = ~np.isfinite(x[src]) | ~np.isfinite(y[src])
        xi = np.where(no_fit, 0, xi).astype(np.int32)
        yi = np.where(no_fit, 0, yi).astype(np.int32)
        lr = np.asarray(logre[src], np.float64)
        re = np.where(ext & np.isfinite(lr), np.exp(np.clip(lr, -5.0, 5.0)), 0.0)
        sgal2 = (re / GAUSS_HALF_LIGHT) ** 2
        max_pix = int(max(problems[q].h * problems[q].w for q in qs))
        nw = int(L.uph_confusion_work_floats(int(qs.size), max_pix))
        work = torch.empty(max(nw, 1), dtype=torch.float32, device=dev)
        sig = torch.empty(max(int(src.size), 1), dtype=torch.float32, device=dev)
        uw = (np.asarray(use_wing, np.uint8)[qs] if use_wing is not None
              else np.zeros(qs.size, np.uint8))
        keep = [up(off, np.int64), up(win, np.int32), up(ix, np.int32), up(iy, np.int32),
                up(flux, np.float32), up(bg, np.float32), up(phi, np.float32),
                up(uw, np.uint8), up(sumsq, np.float32), up(radii, np.int32),
                up(sp2, np.float64), up(xi, np.int32), up(yi, np.int32),
                up(sgal2, np.float64), up(no_fit, np.uint8)]
        _lib.check(L.uph_confusion(
            int(qs.size), int(src.size), T, cfg.bg_order, P(keep[0]), P(keep[1]), H, W,
            P(fld.data), P(fld.invvar), P(fld.good), P(wing) if wing is not None else None,
            P(keep[7]), P(stamps), P(keep[2]), P(keep[3]), P(keep[4]), P(keep[5]),
            *map(float, bg_geo), P(keep[6]), P(keep[8]), P(keep[9]), P(keep[10]),
            P(keep[11]), P(keep[12]), P(keep[13]), P(keep[14]), P(work), nw, max_pix, P(sig),
            _lib.stream_ptr(None)))
        s = sig.cpu().numpy()
        for j, q in enumerate(qs):
            out[q] = s[off[j]:off[j + 1]].copy()
        if fcorr_out is not None:
            from .fcorr import fcorr_residual_windows
            for j, r in enumerate(fcorr_residual_windows(work, max_pix, win, phi)):
                fcorr_out[int(qs[j])] = r
    return out


GEOMETRIC_ERR_BASES = ("x_pix_white_err", "y_pix_white_err", "Re_err", "ELL_err", "THETA_err",
                       "sersic_n_err")


def fcorr_effective(out, bandnames, fcorr):
    """Per-source F_corr_eff: the SNR^2-weighted mean of the band factors
    over the bands with a finite flux and a positive flux error, else the
    median band factor (error_columns.py:166-187)."""
    import pandas as pd
    num = lambda c: pd.to_numeric(out[c], errors="coerce").to_numpy(dtype=np.float64)
    n = len(out)
    wsum, fnum = np.zeros(n), np.zeros(n)
    for b, bn in enumerate(bandnames):
        fl, fe = num(f"FLUX_{bn}_fit"), num(f"FLUXERR_{bn}_fit")
        with np.errstate(invalid="ignore", divide="ignore"):
            snr2 = np.where(np.isfinite(fl) & np.isfinite(fe) & (fe > 0), (fl / fe) ** 2, 0.0)
        wsum += snr2
        fnum += snr2 * float(fcorr[b])
    fmed = float(np.median(np.asarray(fcorr, np.float64)))
    return np.where(wsum > 0, fnum / np.maximum(wsum, 1e-30), fmed)


def add_geometric_error_columns(out, bandnames, fcorr, chi2_nu) -> None:
    """The correlated-noise and reduced-chi2 legs of the geometric errors,
    in place (error_columns.py:153-246, pipeline.py:2548-2562, :2877-2881):
    chi2_nu_nl = the island's reduced chi2; {base}_corr = sqrt(max(1,
    F_corr_eff)) x {base} (with ``fcorr``) and {base}_tot = sqrt(max(1,
    F_corr_eff)) x sqrt(max(1, chi2_nu_nl)) x {base} (F_corr_eff = 1 without
    ``fcorr``; the chi2 factor 1 where chi2_nu_nl is not finite) for the
    position, shape and Sersic-n errors."""
    import pandas as pd
    num = lambda c: pd.to_numeric(out[c], errors="coerce").to_numpy(dtype=np.float64)
    n = len(out)
    cn = (np.asarray(chi2_nu, np.float64) if chi2_nu is not None else np.full(n, np.nan))
    out["chi2_nu_nl"] = cn.astype(np.float32)
    if fcorr is not None:
        c_fac = np.sqrt(np.maximum(fcorr_effective(out, bandnames, fcorr), 1.0))
    else:
        c_fac = np.ones(n)
    d_fac = np.sqrt(np.where(np.isfinite(cn), np.maximum(cn, 1.0), 1.0))
    for base in GEOMETRIC_ERR_BASES:
        if base not in out.columns:
            continue
        raw = num(base)
        if fcorr is not None:
            out[f"{base}_corr"] = (c_fac * raw).astype(np.float32)
        out[f"{base}_tot"] = (c_fac * d_fac * raw).astype(np.float32)


def add_error_columns(out, bandnames, flux_err_conf, *, flux_marg_var=None,
                      rchisq_scale_marg_tot: bool = True, fcorr=None) -> None:
    """FLUXERR_CONF / FLUXERR_TOT / FLUXERR_MARG_TOT per band, in place
    (error_columns.py:134-150, :250-363; the source-Poisson term, which needs
    the images' gain, at its neutral value).  ``flux_err_conf`` (n, nb);
    ``flux_marg_var`` (n, nb) the shape-marginalisation increment (0 for
    stars, NaN off-island), scaled in FLUXERR_MARG_TOT by the geometric
    inflation G = max(1, F_corr_eff) x max(1, chi2_nu_nl) (``fcorr`` the
    band factors, or None for F_corr_eff = 1; the chi2 factor where the
    chi2_nu_nl column exists)."""
    import pandas as pd
    num = lambda c: pd.to_numeric(out[c], errors="coerce").to_numpy(dtype=np.float64)
    conf = np.asarray(flux_err_conf, np.float64)
    n = len(out)
    g_corr = (np.maximum(fcorr_effective(out, bandnames, fcorr), 1.0) if fcorr is not None
              else np.ones(n))
    if "chi2_nu_nl" in out.columns:
        cnu = num("chi2_nu_nl")
        g_corr = g_corr * np.where(np.isfinite(cnu), np.maximum(cnu, 1.0), 1.0)
    scale = None
    if rchisq_scale_marg_tot and "rchisq" in out.columns:
        rq = num("rchisq")
        scale = np.sqrt(np.where(np.isfinite(rq), np.maximum(rq, 1.0), 1.0))
    for b, bn in enumerate(bandnames):
        fe = num(f"FLUXERR_{bn}_fit")
        fc = conf[:, b]
        out[f"FLUXERR_CONF_{bn}_fit"] = fc.astype(np.float32)
        fc2 = np.where(np.isfinite(fc), fc * fc, 0.0)
        out[f"FLUXERR_TOT_{bn}_fit"] = np.sqrt(fe * fe + fc2).astype(np.float32)
        if flux_marg_var is None:
            continue
        dv = np.asarray(flux_marg_var, np.float64)[:, b] if n else np.zeros(0)
        tot = np.where(dv == 0.0, np.sqrt(fe * fe + fc2), np.sqrt(fe * fe + dv * g_corr + fc2))
        if scale is not None:
            tot = tot * scale
        out[f"FLUXERR_MARG_TOT_{bn}_fit"] = tot.astype(np.float32)
