"""Per-grid operator: one grid's multi-band linear solve as a task.

Mirrors ``tract7dt/fit/grid_gpu/runner.py`` (``PerGridTask``,
``PerGridResult``, ``run_one_grid``, :354-786) for the island architecture,
where a grid pass is the S1 seed solve (pass 1) or the final S1' solve at
the island fit's parameters (pass 2) of ``solve_grid_3stage``
(``solver_grid.py:2306-3016``): per band, unit stamps and the joint flux +
polynomial background solve, wing peeling of large galaxies, and in pass 2
(``blank_response_correction``) the blank-response correction of STAR
fluxes.  The task's per-band arrays are the grid's padded window; the fit
runs on the device over a one-grid field in window coordinates (positions,
PSF origins and the background origin shifted by the window corner, which
leaves every rendered stamp and basis value unchanged).  As in the
reference, any failure becomes ``error`` plus NaN fluxes, variances and
background, with the catalogue seeds echoed.
"""
from __future__ import annotations

import dataclasses
import time
from dataclasses import dataclass, field
from typing import Any

import numpy as np

from .psf_polynomial import n_terms_for_order


@dataclass
class PerGridTask:
    """Input of one grid (runner.py:354-453).  Fields the island
    architecture never reads (per-source position solver knobs, shared
    memory slices, error-model switches) are accepted and ignored."""
    grid_id: int
    grid_row: int
    grid_col: int
    bandnames: list
    data_per_band: list
    invvar_per_band: list
    pix_good_per_band: list
    psfs_per_band: list
    psf_shape: tuple
    grid_x0: int
    grid_y0: int
    source_indices: np.ndarray
    pos_x_init: np.ndarray
    pos_y_init: np.ndarray
    pos_err_pix: np.ndarray
    source_type: np.ndarray
    logre_init: np.ndarray
    ee1_init: np.ndarray
    ee2_init: np.ndarray
    pos_max_shift_pix: float
    bg_x_origin: float
    bg_x_scale: float
    bg_y_origin: float
    bg_y_scale: float
    bg_poly_order: int = 2
    single_stage: bool = False
    n_opt_iters: int = 5
    pos_step_max_pix: float = 1.0
    dlnp_stop_pix: float = 0.01
    position_solver: str = "per_source"
    joint_cg_max_iters: int = 20
    joint_cg_tol: float = 1.0e-6
    joint_use_preconditioner: bool = True
    pos_stop_unconverged_frac: float = 0.0
    chi2_ftol: float = 0.1
    sersic_n_init: np.ndarray | None = None
    render_backend: str = "direct"
    fit_sersic_n: bool = True
    n_fit_min_snr: float = 30.0
    wing_peeling: bool = False
    wing_peel_tier: int = 221
    wing_peel_max_sources: int = 64
    blank_response_correction: bool = False
    compute_confusion: bool = False
    compute_fcorr_resid: bool = False
    n_grid_slots: int = 1
    shared_data_per_band: list = field(default_factory=list)
    shared_invvar_per_band: list = field(default_factory=list)
    shared_pix_good_per_band: list = field(default_factory=list)
    owned_source_mask: np.ndarray | None = None


@dataclass
class PerGridResult:
    """Output of one grid (runner.py:456-499), rows of the owned sources."""
    grid_id: int
    n_sources: int
    pos_x: np.ndarray
    pos_y: np.ndarray
    logre: np.ndarray
    ee1: np.ndarray
    ee2: np.ndarray
    fluxes: np.ndarray                   # (n_sources, n_bands)
    bg_coefs: np.ndarray                 # (n_bands, M)
    flux_var: np.ndarray
    bg_var: np.ndarray
    pos_var: np.ndarray                  # (n_sources, 2)
    source_indices: np.ndarray
    s1_converged: bool
    s2_converged: bool
    s3_converged: bool
    n_outer_iters: int
    sec_total: float
    error: str | None = None
    position_stranded: np.ndarray | None = None
    sersic_n_fit: np.ndarray | None = None
    sersic_n_frozen: np.ndarray | None = None
    sersic_n_err: np.ndarray | None = None
    shape_frozen: np.ndarray | None = None
    fluxerr_conf: np.ndarray | None = None
    fcorr_resid: list | None = None
    blank_response_bhat: np.ndarray | None = None


def _unsupported(task: PerGridTask) -> str | None:
    if task.compute_confusion or task.compute_fcorr_resid:
        return "the post-fit error model (confusion, correlated noise) is not part of this build"
    st = np.asarray(task.source_type)
    if task.render_backend == "direct" and np.any(st != 0):
        return ("render_backend='direct' for extended sources is not part of this build; the "
                "reference's 'auto' resolves to 'fft' whenever extended sources are present")
    return None


def run_one_grid(task: PerGridTask) -> PerGridResult:
    """Fit one grid (runner.py:502-786): per band the joint flux and
    background solve over the grid's sources (wing peeling when
    ``wing_peeling``; the blank-response correction when
    ``blank_response_correction``); owned rows returned.  Positions and
    shapes are the task's (the island architecture fits them in the island
    stage).  Any exception becomes ``error`` + NaN (runner.py:756-786)."""
    t0 = time.time()
    try:
        why = _unsupported(task)
        if why is not None:
            raise NotImplementedError(why)
        return _fit_grid(task, t0)
    except Exception as exc:                                    # noqa: BLE001
        n = int(np.asarray(task.pos_x_init).shape[0])
        nb = len(task.bandnames)
        M = n_terms_for_order(int(task.bg_poly_order))
        nan = lambda *s: np.full(s, np.nan, np.float32)        # noqa: E731
        return PerGridResult(
            grid_id=task.grid_id, n_sources=n,
            pos_x=np.array(task.pos_x_init, copy=True), pos_y=np.array(task.pos_y_init, copy=True),
            logre=np.array(task.logre_init, copy=True), ee1=np.array(task.ee1_init, copy=True),
            ee2=np.array(task.ee2_init, copy=True), fluxes=nan(n, nb), bg_coefs=nan(nb, M),
            flux_var=nan(n, nb), bg_var=nan(nb, M), pos_var=nan(n, 2),
            source_indices=task.source_indices, s1_converged=False, s2_converged=False,
            s3_converged=False, n_outer_iters=0, sec_total=time.time() - t0,
            error=f"{type(exc).__name__}: {exc}")


def _fit_grid(task: PerGridTask, t0: float) -> PerGridResult:
    from .grid import GridDef
    from .pipeline import DeviceField, FitConfig, linear_pass
    from .source_model import SOURCE_DEV, SOURCE_SERSIC
    gx0, gy0 = int(task.grid_x0), int(task.grid_y0)
    data = np.stack([np.asarray(a, np.float32) for a in task.data_per_band])
    iv = np.stack([np.asarray(a, np.float32) for a in task.invvar_per_band])
    good = np.stack([np.asarray(a, bool) for a in task.pix_good_per_band])
    nb, h, w = data.shape
    if iv.shape != data.shape or good.shape != data.shape:
        raise ValueError("per-band data, invvar and pix_good must share one shape")
    # window coordinates: the grid's (0, 0) is full-image (gx0, gy0)
    psfs = [dataclasses.replace(p, x_origin=p.x_origin - gx0, y_origin=p.y_origin - gy0)
            for p in task.psfs_per_band]
    fld = DeviceField(data, iv, good, psfs)
    n = int(np.asarray(task.pos_x_init).shape[0])
    owned = (np.ones(n, bool) if task.owned_source_mask is None
             else np.asarray(task.owned_source_mask, bool))
    if owned.shape != (n,):
        raise ValueError(f"owned_source_mask shape {owned.shape} does not match "
                         f"n_sources={n}")
    # the field holds this one grid (id 0 in the field's own numbering)
    g = GridDef(0, task.grid_row, task.grid_col, 0, w, 0, h, 0, w, 0, h, 0)
    g.source_indices = np.flatnonzero(owned).astype(np.int64)
    g.halo_source_indices = np.flatnonzero(~owned).astype(np.int64)
    x = (np.asarray(task.pos_x_init, np.float64) - gx0).astype(np.float32)
    y = (np.asarray(task.pos_y_init, np.float64) - gy0).astype(np.float32)
    st = np.asarray(task.source_type, np.int32)
    sn = (np.where(st == SOURCE_DEV, 4.0, np.where(st == SOURCE_SERSIC, 2.0, 1.0))
          if task.sersic_n_init is None else np.asarray(task.sersic_n_init)).astype(np.float32)
    cfg = FitConfig(bg_order=int(task.bg_poly_order), wing_peeling=bool(task.wing_peeling),
                    wing_tier=int(task.wing_peel_tier),
                    wing_max_sources=int(task.wing_peel_max_sources),
                    blank_response=bool(task.blank_response_correction))
    bg_geo = (float(task.bg_x_origin) - gx0, float(task.bg_x_scale),
              float(task.bg_y_origin) - gy0, float(task.bg_y_scale))
    T = int(task.psf_shape[0])
    flux, var, bg, bgv = linear_pass(fld, [g], x, y, st,
                                     np.asarray(task.logre_init, np.float32),
                                     np.asarray(task.ee1_init, np.float32),
                                     np.asarray(task.ee2_init, np.float32), sn, cfg,
                                     {0: T},
                                     final=bool(task.blank_response_correction),
                                     bg_geo=bg_geo, with_bg_var=True)
    M = n_terms_for_order(int(task.bg_poly_order))
    fl, fv = flux[owned], var[owned]
    ok = bool(np.isfinite(fl).all() and np.isfinite(bg).all())
    # position-variance placeholder of the island-mode grid solve
    # (solver_grid.py:2972-2980)
    pe = np.asarray(task.pos_err_pix, np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        pv = 1.0 / np.maximum(pe ** 2 + 1.0 / np.maximum(np.sum(flux, axis=1) ** 2, 1e-30),
                              1e-30)
    free_n = np.zeros(n, bool)
    return PerGridResult(
        grid_id=task.grid_id, n_sources=int(owned.sum()),
        pos_x=np.asarray(task.pos_x_init)[owned], pos_y=np.asarray(task.pos_y_init)[owned],
        logre=np.asarray(task.logre_init)[owned], ee1=np.asarray(task.ee1_init)[owned],
        ee2=np.asarray(task.ee2_init)[owned], fluxes=fl, bg_coefs=bg[0].reshape(nb, M),
        flux_var=fv, bg_var=bgv[0].reshape(nb, M),
        pos_var=np.stack([pv, pv], axis=1)[owned].astype(np.float32),
        source_indices=np.asarray(task.source_indices)[owned], s1_converged=ok,
        s2_converged=True, s3_converged=True, n_outer_iters=0,
        sec_total=time.time() - t0, position_stranded=np.zeros(int(owned.sum()), bool),
        sersic_n_fit=sn[owned],
        sersic_n_frozen=(((st == SOURCE_SERSIC) | (st == SOURCE_DEV)) & ~free_n)[owned],
        shape_frozen=np.zeros(int(owned.sum()), bool))
