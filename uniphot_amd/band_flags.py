"""Band-level output flags of a fitted catalogue.

Mirrors ``tract7dt/fit/grid_gpu/band_flags.py``: no-data band cells carry
no value (:75-112), the per-band reason bitmask ``FLAGS_{b}_fit`` with the
usable-band count (:199-259), the PSF-model negative-lobe zone (:115-137)
and the partial-coverage fraction (:140-196), in the order the reference's
writeback applies them (``fit/grid_gpu/pipeline.py:3084-3130``).

The two PSF diagnostics evaluate the band's PSF model at every fitted
position; here that is one device launch over all (band, source) pairs
(``uph_band_psf_flags``), which returns each stamp's peak, minimum and
uncovered fraction.  The column rules are host code on the DataFrame.
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .psf_polynomial import psf_device_arrays

FLAG_NO_DATA = 1
FLAG_PSF_LOBE = 2
FLAG_ZP_EXTRAP = 4  # reserved (zero point extrapolated); never set by the fit
FLAG_SATURATED = 8
FLAG_PARTIAL_COVERAGE = 16

# more than this fraction of the PSF's positive part on no-data pixels
PARTIAL_COVERAGE_MAX = 0.1
# the PSF model dipping below this fraction of its peak is a broken model
PSF_LOBE_FRAC = 0.002

BAND_FLUX_COLUMN_PREFIXES = (
    "FLUX_", "FLUXERR_", "FLUXERR_CONF_", "FLUX_WIN_", "FLUXERR_TOT_",
    "FLUXERR_MARG_", "FLUX_MARG_FRAC_", "FLUXERR_MARG_TOT_",
)


def _num(out, col):
    import pandas as pd
    return pd.to_numeric(out[col], errors="coerce").to_numpy(dtype=np.float64)


def _band_cols(out, band):
    return [f"{p}{band}_fit" for p in BAND_FLUX_COLUMN_PREFIXES if f"{p}{band}_fit" in out.columns]


def blank_nodata_band_cells(out, bandnames) -> dict:
    """Set every per-band flux column of a cell to NaN where its
    ``FLUXERR_{b}_fit`` is not finite (no information from the band's linear
    solve, or no fit at all), in place.  Returns, per band present, the
    number of such cells that still carried a finite ``FLUX_{b}_fit``."""
    counts = {}
    for band in bandnames:
        ecol = f"FLUXERR_{band}_fit"
        if ecol not in out.columns:
            continue
        bad = ~np.isfinite(_num(out, ecol))
        fcol = f"FLUX_{band}_fit"
        if fcol in out.columns:
            counts[band] = int(np.sum(bad & np.isfinite(_num(out, fcol))))
        if bad.any():
            for col in _band_cols(out, band):
                out.loc[bad, col] = np.nan
    return counts


def psf_flag_values(psfs, x, y, nodata=None, *, device=None, stream=None):
    """Peak, minimum and uncovered fraction of every band's PSF model at
    every position, in one launch.

    ``psfs``: one PsfPolynomial per band (square stamps, one n_terms);
    ``x``, ``y``: (n,) positions in the image frame; ``nodata``: None, or a
    sequence with one entry per band, each None or an (H, W) boolean mask of
    pixels without data (all masks the same shape; a torch tensor on the
    device is used in place).  Returns numpy arrays peak, vmin (n_bands, n)
    float32 and uncovered (n_bands, n) float64 (0 for bands without a
    mask)."""
    import torch
    dev = torch.device(device or "cuda")
    x = np.ascontiguousarray(np.asarray(x, np.float64).ravel())
    y = np.ascontiguousarray(np.asarray(y, np.float64).ravel())
    if x.shape != y.shape:
        raise ValueError(f"x and y differ in shape: {x.shape} vs {y.shape}")
    nb, n = len(psfs), x.size
    sizes = np.array([int(p.coefs.shape[0]) for p in psfs], np.int32)
    for p in psfs:
        if p.coefs.shape[0] != p.coefs.shape[1] or p.coefs.shape[0] % 2 == 0:
            raise ValueError(f"PSF stamps must be square and odd-sized; got {p.coefs.shape[:2]}")
    T = int(sizes.max())
    coefs, _ = psf_device_arrays(psfs, T)
    geo = np.array([(p.x_origin, p.x_scale, p.y_origin, p.y_scale) for p in psfs], np.float64)
    if nodata is not None and len(nodata) != nb:
        raise ValueError(f"nodata has {len(nodata)} entries for {nb} bands")
    masks = [] if nodata is None else [m for m in nodata]
    have = np.array([m is not None for m in masks] if masks else [False] * nb, np.uint8)
    H = W = 0
    nd = hv = None
    if have.any():
        shapes = {tuple(m.shape) for m in masks if m is not None}
        if len(shapes) != 1 or len(next(iter(shapes))) != 2:
            raise ValueError(f"nodata masks must share one (H, W) shape; got {sorted(shapes)}")
        H, W = next(iter(shapes))
        nd = torch.zeros((nb, H, W), dtype=torch.uint8, device=dev)
        for b, m in enumerate(masks):
            if m is None:
                continue
            if isinstance(m, torch.Tensor):
                nd[b] = m.to(device=dev).to(torch.uint8)
            else:
                nd[b] = torch.from_numpy(np.asarray(m, bool).astype(np.uint8)).to(dev)
        hv = torch.from_numpy(have).to(dev)
    tz = lambda a: torch.from_numpy(a).to(dev)
    d_size, d_coefs, d_geo, d_x, d_y = map(tz, (sizes, coefs, geo, x, y))
    peak = torch.empty((nb, n), dtype=torch.float32, device=dev)
    vmin = torch.empty_like(peak)
    unc = torch.empty((nb, n), dtype=torch.float64, device=dev)
    P = _lib.ptr
    _lib.check(_lib.lib().uph_band_psf_flags(
        n, nb, int(coefs.shape[1]), T, P(d_size), P(d_coefs), P(d_geo), P(d_x), P(d_y),
        int(H), int(W), P(nd), P(hv), P(peak), P(vmin), P(unc), _lib.stream_ptr(stream)))
    return peak.cpu().numpy(), vmin.cpu().numpy(), unc.cpu().numpy()


def lobe_from_values(peak, vmin, lobe_frac: float = PSF_LOBE_FRAC):
    """True where the stamp has no positive peak or dips below -lobe_frac
    of it; never at a non-finite position (NaN values)."""
    pk = np.asarray(peak, np.float64)
    mn = np.asarray(vmin, np.float64)
    ok = np.isfinite(pk) & np.isfinite(mn)
    return ok & ((pk <= 0.0) | (mn < -float(lobe_frac) * pk))


def psf_lobe_zone(psf, x, y, *, lobe_frac: float = PSF_LOBE_FRAC, device=None) -> np.ndarray:
    """Negative-lobe zone of one band's PSF model at (x, y)
    (band_flags.py:115-137)."""
    pk, mn, _ = psf_flag_values([psf], x, y, device=device)
    return lobe_from_values(pk[0], mn[0], lobe_frac)


def psf_uncovered_fraction(psf, nodata, x, y, *, device=None) -> np.ndarray:
    """Fraction of one band's PSF (positive part) at (x, y) on pixels
    without data (band_flags.py:140-196)."""
    _, _, u = psf_flag_values([psf], x, y, [nodata], device=device)
    return u[0]


def compose_band_flags(out, bandnames, *, psf_zone: dict | None = None,
                       saturated: dict | None = None, partial: dict | None = None) -> dict:
    """Write ``FLAGS_{b}_fit`` (int16) and ``n_bands_usable`` in place.

    Bits: no data (FLUXERR not finite), PSF lobe (``psf_zone[b]``), partial
    coverage (``partial[b]``, value kept) and saturated (``saturated[b]``,
    every per-band flux column of the cell blanked).  A band counts as
    usable where its flags are 0 and its flux is finite.  Returns per band
    the no-data count and, for the other reasons, the count among cells
    that had data."""
    psf_zone, saturated, partial = psf_zone or {}, saturated or {}, partial or {}
    usable = np.zeros(len(out), np.int16)
    counts = {}
    for band in bandnames:
        ecol = f"FLUXERR_{band}_fit"
        if ecol not in out.columns:
            continue
        nodata = ~np.isfinite(_num(out, ecol))
        flags = np.where(nodata, FLAG_NO_DATA, 0).astype(np.int16)
        got = {}
        for key, bit, table in (("psf_lobe", FLAG_PSF_LOBE, psf_zone),
                                ("partial_coverage", FLAG_PARTIAL_COVERAGE, partial),
                                ("saturated", FLAG_SATURATED, saturated)):
            m = table.get(band)
            if m is None:
                got[key] = 0
                continue
            m = np.asarray(m, bool)
            flags |= np.where(m, bit, 0).astype(np.int16)
            got[key] = int(np.sum(m & ~nodata))
            if bit == FLAG_SATURATED and m.any():
                for col in _band_cols(out, band):
                    out.loc[m, col] = np.nan
        out[f"FLAGS_{band}_fit"] = flags
        fcol = f"FLUX_{band}_fit"
        flux = _num(out, fcol) if fcol in out.columns else np.full(len(out), np.nan)
        usable += ((flags == 0) & np.isfinite(flux)).astype(np.int16)
        counts[band] = {"no_data": int(nodata.sum()), "psf_lobe": got["psf_lobe"],
                        "saturated": got["saturated"],
                        "partial_coverage": got["partial_coverage"]}
    out["n_bands_usable"] = usable
    return counts


def apply_band_flags(out, bandnames, psfs, *, nodata=None, saturated=None, device=None,
                     lobe_frac: float = PSF_LOBE_FRAC,
                     coverage_max: float = PARTIAL_COVERAGE_MAX) -> dict:
    """The writeback's band-level sequence (pipeline.py:3084-3130) on a
    catalogue that already holds the fit's columns: blank the no-data
    cells, evaluate every band's PSF at ``x_pix_white_fit`` /
    ``y_pix_white_fit`` (lobe zone; uncovered fraction where ``nodata``
    gives the band a mask) and compose the flags.  ``psfs``: one
    PsfPolynomial per band of ``bandnames`` (None for a band without a
    polynomial PSF: no PSF bit); ``nodata``: per band None or an (H, W)
    mask; ``saturated``: {band: (n,) bool}.  Returns {"blanked": ...,
    "flags": ...} counts."""
    blanked = blank_nodata_band_cells(out, bandnames)
    fx, fy = _num(out, "x_pix_white_fit"), _num(out, "y_pix_white_fit")
    have = [i for i, p in enumerate(psfs) if p is not None and getattr(p, "coefs", None) is not None]
    zone, part = {}, {}
    if have:
        masks = None if nodata is None else [nodata[i] for i in have]
        pk, mn, unc = psf_flag_values([psfs[i] for i in have], fx, fy, masks, device=device)
        for j, i in enumerate(have):
            band = bandnames[i]
            zone[band] = lobe_from_values(pk[j], mn[j], lobe_frac)
            if masks is not None and masks[j] is not None:
                part[band] = unc[j] > coverage_max
    flags = compose_band_flags(out, bandnames, psf_zone=zone, saturated=saturated, partial=part)
    return {"blanked": blanked, "flags": flags}


def _bool_col(out, col, default=False):
    if col not in out.columns:
        return None
    return out[col].astype(object).where(out[col].notna(), default).astype(bool).to_numpy()


def compute_reliability_flag(out, *, island_stalled_mode: str = "fail",
                             island_giant_mode: str = "fail", high_rchisq_mode: str = "warn",
                             blend_unresolved_mode: str = "fail",
                             island_plateau_mode: str = "stalled") -> None:
    """``fit_is_reliable`` and ``fit_quality_issues`` in place, the shared
    reliability contract of ``tract7dt/merge.py:16-344`` for a catalogue of
    this build's columns.

    Hard reasons (unreliable): excluded, no_fit_result, hit_max_iters,
    bound_stalled, bright_mask, island_stalled, blend_unresolved,
    giant_frozen_geometry, no_usable_band (``n_bands_usable`` = 0 on a
    fitted source), high_rchisq (in "fail" mode).  Soft tags: island_plateau
    (plateau-certified stalls in "converged" mode), island_diverged,
    blend_suspect, blend_resolved, not_converged, all_psf_fallback,
    shape_frozen, and the mode-selectable reasons in "warn" mode.  Tags are
    joined with commas, hard reasons first, in the order listed.  A missing
    column is no signal."""
    import pandas as pd
    for name, v, ok in (("island_stalled_mode", island_stalled_mode, ("fail", "warn", "ignore")),
                        ("island_plateau_mode", island_plateau_mode, ("stalled", "converged")),
                        ("blend_unresolved_mode", blend_unresolved_mode,
                         ("fail", "warn", "ignore")),
                        ("island_giant_mode", island_giant_mode, ("fail", "warn", "ignore")),
                        ("high_rchisq_mode", high_rchisq_mode, ("fail", "warn", "ignore"))):
        if v not in ok:
            raise ValueError(f"{name} must be one of {'|'.join(ok)}; got {v!r}")
    n = len(out)
    if n == 0:
        out["fit_is_reliable"] = pd.Series(dtype=bool)
        out["fit_quality_issues"] = pd.Series(dtype=str)
        return
    zero = np.zeros(n, bool)
    hard, soft = [], []

    def moded(tag, col, mode, fail="fail"):
        m = _bool_col(out, col)
        if m is None or mode == "ignore":
            return
        (hard if mode == fail else soft).append((tag, m))

    excluded = _bool_col(out, "excluded_any")
    excluded = zero if excluded is None else excluded
    hard.append(("excluded", excluded))
    if "x_pix_white_fit" in out.columns:
        no_fit = out["x_pix_white_fit"].isna().to_numpy(bool)
    else:
        per_band = [c for c in out.columns
                    if c.startswith("x_pix_white_") and c.endswith("_fit") and c != "x_pix_white_fit"]
        no_fit = (out[per_band].isna().all(axis=1).to_numpy(bool) if per_band
                  else np.ones(n, bool))
    no_fit &= ~excluded
    hard.append(("no_fit_result", no_fit))
    hmi = _bool_col(out, "opt_hit_max_iters")
    if hmi is None:
        cols = [c for c in out.columns if c.startswith("opt_hit_max_iters_")]
        hmi = (out[cols].fillna(False).astype(bool).any(axis=1).to_numpy() if cols else zero)
    hard.append(("hit_max_iters", hmi))
    bst = _bool_col(out, "flag_bound_stalled")
    if bst is None:
        cols = [c for c in out.columns
                if c.startswith("flag_bound_stalled_") and "which_parameter" not in c]
        bst = (out[cols].fillna(False).astype(bool).any(axis=1).to_numpy() if cols else zero)
    hard.append(("bound_stalled", bst))
    bm = _bool_col(out, "affected_by_bright_mask")
    if bm is not None:
        hard.append(("bright_mask", bm))
    stalled = _bool_col(out, "island_stalled")
    if island_stalled_mode != "ignore" and stalled is not None:
        plat = _bool_col(out, "island_plateau")
        if island_plateau_mode == "converged" and plat is not None and plat.any():
            soft.append(("island_plateau", plat & stalled))
            stalled = stalled & ~plat
        if stalled.any():
            (soft if island_stalled_mode == "warn" else hard).append(("island_stalled", stalled))
    for tag, col in (("island_diverged", "island_diverged"), ("blend_suspect", "flag_blend_suspect"),
                     ("blend_resolved", "blend_resolved")):
        m = _bool_col(out, col)
        if m is not None and m.any():
            soft.append((tag, m))
    moded("blend_unresolved", "blend_unresolved", blend_unresolved_mode)
    moded("giant_frozen_geometry", "island_giant", island_giant_mode)
    if "n_bands_usable" in out.columns:
        nbu = pd.to_numeric(out["n_bands_usable"], errors="coerce").fillna(0).to_numpy(float)
        hard.append(("no_usable_band", (nbu <= 0) & ~excluded & ~no_fit))
    moded("high_rchisq", "flag_high_rchisq", high_rchisq_mode)
    conv = _bool_col(out, "opt_converged", default=True)
    if conv is None:
        cols = [c for c in out.columns if c.startswith("opt_converged_")]
        conv = (out[cols].fillna(True).astype(bool).all(axis=1).to_numpy() if cols
                else np.ones(n, bool))
    nc = ~conv & ~hmi & ~excluded & ~no_fit
    if nc.any():
        soft.append(("not_converged", nc))
    if "psf_fallback_band_count" in out.columns and "psf_used_epsf_band_count" in out.columns:
        fb = pd.to_numeric(out["psf_fallback_band_count"], errors="coerce").fillna(0).to_numpy()
        used = pd.to_numeric(out["psf_used_epsf_band_count"], errors="coerce").fillna(0).to_numpy()
        allfb = (fb > 0) & (used == 0)
        if allfb.any():
            soft.append(("all_psf_fallback", allfb))
    shp = _bool_col(out, "shape_frozen")
    if shp is not None and shp.any():
        soft.append(("shape_frozen", shp))
    items = hard + soft
    mat = np.column_stack([m for _, m in items])
    out["fit_is_reliable"] = ~mat[:, :len(hard)].any(axis=1)
    tags = [t for t, _ in items]
    out["fit_quality_issues"] = [",".join(t for t, on in zip(tags, row) if on) for row in mat]
