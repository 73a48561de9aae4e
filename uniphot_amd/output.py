"""Catalogue columns of an island-mode field fit.

Mirrors the island-architecture writeback of
``tract7dt/fit/grid_gpu/pipeline.py`` (:2498-2725 per-grid merge, :2730-2785
bound flags, :2785-2910 island columns) and the statistical
shape-marginalised error columns of ``error_columns.py`` (:255-341):
  * fitted positions (and RA/DEC through a caller's WCS), solver flags from
    the island stage, ``FLUX_{b}_fit`` = pass-2 flux, ``FLUXERR_{b}_fit`` =
    sqrt(pass-2 variance) where positive and finite;
  * owning grid, type name, fitted shape (Re, axis ratio, angle; ELL = 1 -
    ab, THETA = phi - 90), Sersic n of DEV/SERSIC rows and its freeze flag;
  * island membership, stall, divergence and plateau flags;
  * catalog-frame 1-sigma errors of the fitted parameters, per-source model
    metrics and window fluxes (post-fit pass);
  * FLUXERR_MARG_{b}_fit = sqrt(FLUXERR^2 + increment) and
    FLUX_MARG_FRAC_{b}_fit = increment / FLUXERR^2 (0 for stars);
  * flag_bound_stalled, flag_shape_bound, flag_slow_converge;
  * with ``band_psfs`` given, the band-level contract
    (:mod:`uniphot_amd.band_flags`, ``pipeline.py:3084-3130``): no-data
    cells blanked, ``FLAGS_{b}_fit`` and ``n_bands_usable``;
  * ``fit_is_reliable`` / ``fit_quality_issues`` (``merge.py:16-344``).
  * the error model's flux columns (FLUXERR_CONF, FLUXERR_TOT,
    FLUXERR_MARG_TOT; :mod:`uniphot_amd.error_model`) when the fit carries
    the confusion term;
  * chi2_nu_nl and the geometric error columns scaled by the correlated-
    noise factor and the island's reduced chi2 (``*_err_corr``,
    ``*_err_tot``; error_columns.py:153-246).
"""
from __future__ import annotations

import numpy as np

from .island_plan import _catalog_errors
from .source_model import (SERSIC_N_MAX, SERSIC_N_MIN, SOURCE_DEV, SOURCE_EXP, SOURCE_SERSIC,
                           SOURCE_STAR, ellipse_soft_to_re_ab_phi_np, sersic_n_to_u)

_TYPE_NAMES = {SOURCE_STAR: "star", SOURCE_EXP: "exp", SOURCE_DEV: "dev",
               SOURCE_SERSIC: "sersic"}


def slow_converge_flag(n_steps, has_fit, max_steps_main, max_steps) -> np.ndarray:
    """opt_niters > max_steps_main (the main pass / re-queue boundary); with
    no re-queue (max_steps_main unset, 0 or >= max_steps) opt_niters >=
    max_steps (pipeline.py:1294-1325)."""
    n = np.asarray(n_steps, float)
    fit = np.asarray(has_fit, bool)
    ms = int(max_steps) if max_steps is not None else 0
    mm = int(max_steps_main) if max_steps_main else 0
    if ms <= 0:
        return np.zeros(n.size, bool)
    if 0 < mm < ms:
        return fit & (n > mm)
    return fit & (n >= ms)


def write_fit_columns(catalog, fit, bandnames, *, x_init, y_init, source_type,
                      pos_max_shift_pix: float = 5.0, psf_stamp_size: int = 55, wcs=None,
                      param_errors: bool = True, source_metrics: bool = True,
                      marginal_flux_err: bool = True, max_steps_main: int = 8,
                      max_steps: int = 30, band_psfs=None, band_nodata=None,
                      band_saturated=None, device=None, reliability: bool = True):
    """A copy of ``catalog`` (a DataFrame, rows aligned with the fit's
    sources) with the fit's columns added.  ``fit`` is a FitResult of
    :func:`uniphot_amd.pipeline.fit_field`; ``x_init``, ``y_init`` the
    catalogue seeds in the fit frame; ``wcs`` (optional) converts fitted
    pixels with ``all_pix2world(x, y, 0)``.  ``band_psfs`` (one
    PsfPolynomial or None per band) turns on the band flags, with the
    optional per-band no-data masks ``band_nodata`` and saturation hits
    ``band_saturated`` ({band: (n,) bool}); their counts are in
    ``out.attrs["band_flags"]``.  ``reliability`` adds ``fit_is_reliable``
    and ``fit_quality_issues`` (merge.py:16-344, default modes)."""
    out = catalog.copy()
    n = len(out)
    st = np.asarray(source_type, np.int32)
    if st.size != n or np.asarray(fit.x).size != n:
        raise ValueError("catalog, source_type and the fit must have one row per source")
    giant = np.asarray(fit.island_giant, bool)
    in_isl = np.asarray(fit.island_id) >= 0
    fitted = np.isfinite(np.asarray(fit.x, float)) & in_isl & ~giant
    out["x_pix_white_fit"] = np.asarray(fit.x, np.float64)
    out["y_pix_white_fit"] = np.asarray(fit.y, np.float64)
    out["RA_fit"] = np.nan
    out["DEC_fit"] = np.nan
    if wcs is not None:
        ra, dec = wcs.all_pix2world(np.asarray(fit.x), np.asarray(fit.y), 0)
        out["RA_fit"], out["DEC_fit"] = ra, dec
    # island-stage convergence; giant members keep their seeds and count as
    # converged (island_plan.py:1019-1088)
    conv = np.where(fitted, np.asarray(fit.island_conv, bool), giant)
    out["opt_converged"] = conv
    out["opt_niters"] = np.where(fitted, np.asarray(fit.island_nst, int), 0)
    out["opt_hit_max_iters"] = fitted & ~conv
    out["position_stranded"] = fitted & ~conv
    out["opt_last_dlnp"] = np.nan
    var = np.asarray(fit.flux_var if fit.flux_var is not None else np.square(fit.flux_err),
                     np.float64)
    fe = np.where(np.isfinite(var) & (var > 0), np.sqrt(np.where(var > 0, var, 1.0)), np.nan)
    for b, bn in enumerate(bandnames):
        out[f"FLUX_{bn}_fit"] = np.asarray(fit.flux[:, b], np.float64)
        out[f"FLUXERR_{bn}_fit"] = fe[:, b].astype(np.float32).astype(np.float64)
    out["grid_id_fit"] = (np.asarray(fit.grid_id, int) if fit.grid_id is not None
                          else np.full(n, -1))
    out["stype_fit"] = np.array([_TYPE_NAMES.get(int(t), "star") for t in st], dtype=object)
    ext = st != SOURCE_STAR
    re, ab, phi = ellipse_soft_to_re_ab_phi_np(fit.logre, fit.ee1, fit.ee2)
    for col, v in (("re_pix_fit", re), ("ab_fit", ab), ("phi_deg_fit", phi), ("Re_fit", re),
                   ("ELL_fit", 1.0 - ab), ("THETA_fit", phi - 90.0)):
        out[col] = np.where(ext, v.astype(np.float64), np.nan)
    has_n = (st == SOURCE_DEV) | (st == SOURCE_SERSIC)
    out["sersic_n_fit"] = np.where(has_n, np.asarray(fit.sersic_n, np.float64), np.nan)
    free_n = (np.asarray(fit.free_n, bool) if fit.free_n is not None
              else np.zeros(n, bool))
    out["sersic_n_frozen"] = has_n & ~free_n
    out["shape_frozen"] = False
    # island membership and LM flags
    nsrc = np.bincount(np.asarray(fit.island_id)[in_isl], minlength=1)
    out["island_id"] = np.asarray(fit.island_id, int)
    out["island_nsrc"] = np.where(in_isl, nsrc[np.maximum(fit.island_id, 0)], 0)
    out["island_giant"] = giant
    out["island_stalled"] = np.asarray(fit.island_stall, bool)
    out["island_diverged"] = (np.asarray(fit.island_diverged, bool)
                              if fit.island_diverged is not None else False)
    if fit.island_stall_dchi2 is not None:
        out["island_stall_dchi2"] = np.asarray(fit.island_stall_dchi2, np.float64)
        out["island_plateau"] = np.asarray(fit.island_plateau, bool)
    out["island_railed"] = False
    # catalog-frame parameter errors (post-fit pass)
    if param_errors and fit.param_var is not None:
        pv = np.asarray(fit.param_var, np.float64)
        p = np.stack([fit.x, fit.y, fit.logre, fit.ee1, fit.ee2,
                      sersic_n_to_u(np.clip(np.asarray(fit.sersic_n, np.float64),
                                            SERSIC_N_MIN, SERSIC_N_MAX))], axis=1)
        keep = np.zeros((n, 6), bool)
        keep[:, :2] = fitted[:, None]
        keep[:, 2:5] = (fitted & ext)[:, None]
        keep[:, 5] = fitted & free_n
        xe, ye, ree, elle, the, ne = _catalog_errors(p, pv[:, :6], pv[:, 6], keep, free_n,
                                                     np.asarray(fit.sersic_n, np.float32))
        for col, v in (("x_pix_white_err", xe), ("y_pix_white_err", ye), ("Re_err", ree),
                       ("ELL_err", elle), ("THETA_err", the)):
            out[col] = np.where(fitted & ~giant, v, np.nan)
        out["sersic_n_err"] = np.where(fitted & free_n, ne, np.nan)
    else:
        out["sersic_n_err"] = np.nan
    if source_metrics and fit.metrics is not None:
        for q, col in enumerate(("rchisq", "fracflux", "fracin", "fracmasked")):
            out[col] = np.asarray(fit.metrics[:, q], np.float64)
        if fit.flux_window is not None:
            for b, bn in enumerate(bandnames):
                out[f"FLUX_WIN_{bn}_fit"] = np.asarray(fit.flux_window[:, b], np.float64)
    if marginal_flux_err and fit.flux_marg_var is not None:
        dv_all = np.asarray(fit.flux_marg_var, np.float64)
        for b, bn in enumerate(bandnames):
            f_e = out[f"FLUXERR_{bn}_fit"].to_numpy(np.float64)
            dv = dv_all[:, b]
            marg = np.where(dv == 0.0, f_e, np.sqrt(f_e * f_e + dv))
            out[f"FLUXERR_MARG_{bn}_fit"] = marg.astype(np.float32).astype(np.float64)
            with np.errstate(divide="ignore", invalid="ignore"):
                frac = np.where(dv == 0.0, 0.0, np.where(f_e > 0.0, dv / (f_e * f_e), np.nan))
            out[f"FLUX_MARG_FRAC_{bn}_fit"] = frac.astype(np.float32)
    # bound flags (pipeline.py:2730-2783)
    dx = np.asarray(fit.x, float) - np.asarray(x_init, float)
    dy = np.asarray(fit.y, float) - np.asarray(y_init, float)
    ok = np.isfinite(dx) & np.isfinite(dy)
    out["flag_bound_stalled"] = ok & ((np.abs(np.abs(dx) - pos_max_shift_pix) < 1e-4)
                                      | (np.abs(np.abs(dy) - pos_max_shift_pix) < 1e-4))
    cap = np.full(n, 0.75 * float(psf_stamp_size))
    if fit.re_upper_pix is not None:
        up = np.asarray(fit.re_upper_pix, float)
        cap = np.where(np.isfinite(up) & (up > 0), up, cap)
    re_v = out["re_pix_fit"].to_numpy(float)
    out["flag_shape_bound"] = np.isfinite(re_v) & (np.abs(re_v - cap) < 1e-2 * cap)
    has_fit = np.isfinite(out["x_pix_white_fit"].to_numpy(float))
    out["flag_slow_converge"] = slow_converge_flag(out["opt_niters"].to_numpy(), has_fit,
                                                   max_steps_main, max_steps)
    from .error_model import add_error_columns, add_geometric_error_columns
    fcorr = getattr(fit, "fcorr", None)
    add_geometric_error_columns(out, list(bandnames), fcorr, getattr(fit, "chi2_nu", None))
    if getattr(fit, "flux_err_conf", None) is not None:
        add_error_columns(out, list(bandnames), fit.flux_err_conf,
                          flux_marg_var=fit.flux_marg_var if marginal_flux_err else None,
                          fcorr=fcorr)
    if band_psfs is not None:
        from .band_flags import apply_band_flags
        out.attrs["band_flags"] = apply_band_flags(out, list(bandnames), list(band_psfs),
                                                   nodata=band_nodata, saturated=band_saturated,
                                                   device=device)
    if reliability:
        from .band_flags import compute_reliability_flag
        compute_reliability_flag(out)
    return out
