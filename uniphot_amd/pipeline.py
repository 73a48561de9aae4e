"""Island-architecture forced-photometry fit of one field on one GPU.

The single-GPU restatement of ``run_grid_gpu_fit`` in island mode
(``tract7dt/fit/grid_gpu/pipeline.py:1352-2725``) and
``run_island_prestage`` (``island_stage.py:1338-3363``):

  pass 1  per-grid linear S1 solve at the catalog seeds (fluxes + plane),
          wing peeling of large galaxies
  islands flux-aware coupling radii (or a fixed linking radius), connected
          components, (B, N) buckets, star/mixed classes and tiers
  stacks  padded field stacks: data - per-grid S1 plane on each grid core,
          every catalogued star subtracted at its S1 flux (scene subtraction)
  fit     band-split VarPro-LM over launch sequences of island cutouts (own
          member stars added back), post-fit errors and source metrics
  pass 2  per-grid linear S1' solve at the fitted positions and shapes,
          blank-response correction of the star fluxes

Every default-on knob of the reference has a switch in FitConfig; catalog
priors are not built (the run record lists what is off).
Inputs stay resident on the device for the whole fit.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import forced_phot_sparse as fps
from .error_model import confusion_sigma
from .grid import assign_sources_to_grids, divide_image_into_grids
from .island_plan import (DEFAULT_MAX_STEPS, JOINT_NSRC_CAP, class_host_arrays, free_n_gate,
                          plan_island_classes)
from .island_solver import (FP64_REFIT_KAPPA, PLATEAU_DLNP_TOL, fit_float64,
                           make_class_kernels, psf_tree_device, refit_ill_conditioned,
                           run_class_chunk, stall_probes)
from .islands import (ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES, bucket_islands,
                      build_islands, flux_aware_radii, psf_radial_profile)
from .psf_polynomial import evaluate_psf_at_np, psf_device_arrays
from .scene import SceneStars, island_mosaic, mosaic_bytes, star_flux, subtract_stars
from .source_model import SOURCE_STAR, sersic_u_to_n


# float32 fits send islands above this flux contrast to the float64 kernels
FP64_CONTRAST = 4.0


@dataclass
class FitConfig:
    n_rows: int = 1
    n_cols: int = 1
    halo: int = 30
    bg_order: int = 1
    kappa: float = 0.5
    link_radius: float | None = None   # fixed coupling radius (px) instead of flux-aware
    pos_max_shift_pix: float = 5.0
    max_steps: int = DEFAULT_MAX_STEPS
    reg: float = 1e-9
    k_max: int = 32
    precision: str = "fp32"
    stamp_tiers: tuple = ()      # whole-grid stamp-tier promotion ladder; () = native only
    wing_peeling: bool = True    # large galaxies: subtract tier-wing_tier wings (native solve)
    wing_tier: int = 221
    wing_max_sources: int = 64
    blank_response: bool = True  # pass 2: subtract the blank-response median from STAR fluxes
    scene_subtract: bool = True  # island fit: non-member stars removed at their S1 flux
    island_split: bool = True    # band-split island LM (uph_island_fit_split)
    galaxy_prerender: bool = True  # split path: galaxies rendered by their own launches
    post_fit: bool = True        # parameter variances, marginal flux errors, source metrics
    gram: str = "fp32"           # island Gram operands: "fp32" or "bf16" (matrix cores)
    stamp_budget: float = 24e9   # bytes of rendered stamps per linear-solve launch
    island_workspace: float = 40 * 2 ** 30  # bytes of island-fit workgroup workspace (cap)
    # float32 fits: islands whose scaled Gauss-Newton condition number
    # exceeds this are refitted on the float64 kernels (None: never)
    fp64_refit_kappa: float | None = FP64_REFIT_KAPPA
    # float32 fits: islands whose flux contrast (the brightest member's flux
    # over the faintest member's max(|flux|, sigma), in any band, from the
    # pass-1 solve) exceeds this are fitted on the float64 kernels, where
    # float32 model error (~1e-7 of the bright member's light) would move
    # the faint members by more than 1e-5 (None: never)
    fp64_contrast: float | None = FP64_CONTRAST
    # post-hoc diverged / plateau probes of stalled islands (the reference
    # builds and runs both by default, island_solver.py:1762-1764, :2984-3009)
    stall_probes: bool = True
    # persistent workgroups of the band-split launch sequences per precision
    # ({"fp32": n, "fp64": n}); None: measured per workload by the
    # autotuner (uniphot_amd/autotune.py) and cached, or MAX_BLOCKS with
    # autotune off
    island_blocks: dict | None = None
    autotune: bool = True
    # post-fit error model: confusion flux error of every source after
    # pass 2 (error_model.py:663-930, on by default in the reference)
    confusion: bool = True
    # correlated-noise factor F_corr per band (error_model.py:933-1302,
    # pipeline.py:2411-2458; correlated_noise_enabled, on by default): the
    # residual estimate of each pass-2 confusion window, else the raw-sky one
    correlated_noise: bool = True


@dataclass
class FitResult:
    flux: np.ndarray          # (n, nb) pass-2 fluxes
    flux_err: np.ndarray      # (n, nb) sqrt(pass-2 variance)
    s1_flux: np.ndarray       # (n, nb)
    s1_var: np.ndarray
    x: np.ndarray             # fitted positions / shapes
    y: np.ndarray
    logre: np.ndarray
    ee1: np.ndarray
    ee2: np.ndarray
    sersic_n: np.ndarray
    island_id: np.ndarray     # (n,) label, -1 never
    island_conv: np.ndarray   # (n,) bool per source (of its island)
    island_stall: np.ndarray
    island_nst: np.ndarray
    island_giant: np.ndarray
    timings: dict = field(default_factory=dict)
    record: dict = field(default_factory=dict)
    # island post-fit pass (NaN where not computed):
    param_var: np.ndarray = None      # (n, 7) var x, y, logre, ee1, ee2, u, cov(ee1, ee2)
    island_flux: np.ndarray = None    # (n, nb) island VarPro fluxes
    island_flux_err_marg: np.ndarray = None  # (n, nb) shape-marginalised island flux error
    metrics: np.ndarray = None        # (n, 4) rchisq, fracflux, fracin, fracmasked
    flux_window: np.ndarray = None    # (n, nb)
    island_diverged: np.ndarray = None    # (n,) stall probes (island_solver.py:1335-1465)
    island_stall_dchi2: np.ndarray = None
    island_plateau: np.ndarray = None     # stalled with no damped step better than -2 tol
    island_fp64: np.ndarray = None        # (n,) island fitted on the float64 kernels
    flux_var: np.ndarray = None           # (n, nb) pass-2 flux variance
    flux_marg_var: np.ndarray = None      # (n, nb) shape-marginalisation variance increment
    re_upper_pix: np.ndarray = None       # (n,) fitted-Re bound 0.75 x the island's tier
    free_n: np.ndarray = None             # (n,) Sersic n fitted
    grid_id: np.ndarray = None            # (n,) grid owning the source
    flux_err_conf: np.ndarray = None      # (n, nb) confusion flux error (error model)
    fcorr: np.ndarray = None              # (nb,) correlated-noise variance factor F_corr
    chi2_nu: np.ndarray = None            # (n,) reduced chi2 of the source's island


class DeviceField:
    """Per-band images resident on the device plus the band PSFs."""

    def __init__(self, data, invvar, good, psfs, device="cuda"):
        import torch
        self.data = torch.from_numpy(np.ascontiguousarray(data, np.float32)).to(device)
        iv = np.where(good, invvar, 0.0).astype(np.float32)
        self.invvar = torch.from_numpy(iv).to(device)
        self.good = torch.from_numpy(np.ascontiguousarray(good, np.uint8)).to(device)
        self.psfs = list(psfs)
        self.nb, self.H, self.W = data.shape
        self.native = int(psfs[0].H_psf)
        self.n_terms = int(psfs[0].n_terms)
        self._psf_dev = {}
        self.device = device
        self.sigma_sky = float(np.median([
            np.median(1.0 / np.sqrt(v[::16, ::16][v[::16, ::16] > 0])) for v in iv]))

    @classmethod
    def from_device(cls, data, invvar, good, psfs, psf_dev=None):
        """Wrap images already on the device (no copy); ``psf_dev`` maps
        stamp tier -> (coefs, geo) device arrays to use instead of building
        them from ``psfs``."""
        import torch
        self = cls.__new__(cls)
        self.data = data
        self.invvar = torch.where(good.bool(), invvar, torch.zeros((), dtype=invvar.dtype,
                                                                  device=invvar.device))
        self.good = good
        self.psfs = list(psfs)
        self.nb, self.H, self.W = (int(v) for v in data.shape)
        self.native = int(psfs[0].H_psf)
        self.n_terms = int(psfs[0].n_terms)
        self._psf_dev = dict(psf_dev or {})
        self.device = data.device
        iv = self.invvar[:, ::16, ::16]
        med = [torch.median(1.0 / torch.sqrt(v[v > 0])).item() for v in iv]
        self.sigma_sky = float(np.median(med))
        return self

    def psf(self, T):
        import torch
        if T not in self._psf_dev:
            c, g = psf_device_arrays(self.psfs, T)
            self._psf_dev[T] = (torch.from_numpy(c).to(self.device),
                                torch.from_numpy(g).to(self.device))
        return self._psf_dev[T]


def _src_tier(stype, logre, sersic_n, tiers):
    nre = np.where(sersic_n < 1.5, 4.0, np.where(sersic_n < 3.0, 6.0, 8.0))
    req = np.where(stype != SOURCE_STAR, nre * np.exp(logre.astype(np.float64)), 0.0)
    tier = np.full(stype.size, tiers[-1], np.int32)
    for t in reversed(tiers):
        tier = np.where(req <= t // 2, t, tier)
    return tier


def grid_tiers(fld: DeviceField, grids, stype, logre, sn, tiers):
    """Stamp tier of each grid: the largest requirement of its fit sources
    at the catalog seeds (pipeline.py:1609-1621, :2014-2019)."""
    src_tier = _src_tier(stype, logre, sn, tiers)
    return {g.grid_id: (int(src_tier[g.fit_indices].max()) if g.fit_indices.size else
                        fld.native) for g in grids}


def wing_sources(stype, logre, sersic_n, native: int, max_sources: int) -> np.ndarray:
    """Indices of the "big" sources whose wings are peeled: extended and
    nre(n) Re > native / 2 with nre = 4 / 6 / 8 for n < 1.5 / < 3 / >= 3,
    the largest ``max_sources`` by nre Re (solver_grid.py:1517-1548)."""
    n = np.asarray(sersic_n, np.float64)
    nre = np.where(n < 1.5, 4.0, np.where(n < 3.0, 6.0, 8.0))
    req = nre * np.exp(np.asarray(logre, np.float64))
    big = (np.asarray(stype) != SOURCE_STAR) & np.isfinite(req) & (req > 0.5 * float(native))
    idx = np.where(big)[0]
    if idx.size > int(max_sources):
        idx = np.sort(idx[np.argsort(-req[idx], kind="stable")[:int(max_sources)]])
    return idx.astype(np.int64)


def _disjoint_colors(grids):
    """Colour grids so that the padded windows of one colour are disjoint."""
    colors, boxes = {}, []
    for g in grids:
        c = 0
        while any(not (g.x1_pad <= b[0] or b[2] <= g.x0_pad or g.y1_pad <= b[1] or
                       b[3] <= g.y0_pad) for b in boxes if b[4] == c):
            c += 1
        colors[g.grid_id] = c
        boxes.append((g.x0_pad, g.y0_pad, g.x1_pad, g.y1_pad, c))
    return colors


def blank_response_correct(fld, T, problems, owners, res, x, y, stype, logre, ee1, ee2, sn,
                           cfg: FitConfig, bg_geo, wing=None, use_wing=None):
    """Blank-response (empty-aperture) correction of the final solve of
    pass 2 (solver_grid.py:2855-2943): per (grid, band), the median
    matched-filter response of the final residual -- corr(w r, phi) /
    corr(w, phi^2), phi the band PSF at the window centre clipped >= 0 and
    unit-sum -- is subtracted from the STAR fluxes.  The model uses unit
    stamps without renormalisation (the STAR rows of render_source_stamps)
    and the final fluxes and background; with wing peeling the residual is
    taken against data - wing (``wing`` (nb, H, W) for the problems flagged
    in ``use_wing``).  Runs on the device (``uph_blank_response``) in
    launches bounded by the stamp budget; returns the corrected results and
    b_hat per problem."""
    import torch
    from . import _lib
    from .psf_polynomial import evaluate_psf_at_np
    L = _lib.lib()
    P = _lib.ptr
    dev = fld.data.device
    nb, H, W = fld.nb, fld.H, fld.W
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    cs, gs = fld.psf(T)
    out = list(res)
    b_hat = np.full(len(problems), np.nan)
    n_of = np.array([owners[q][0].fit_indices.size for q in range(len(problems))])
    cost = n_of * T * T * 4.0 + np.array([pr.h * pr.w * 12.0 for pr in problems])
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
        flux = np.concatenate([out[q].fluxes for q in qs]).astype(np.float32)
        flux = np.where(np.isfinite(flux), flux, 0.0).astype(np.float32)
        bg = np.stack([out[q].bg_coefs for q in qs]).astype(np.float32)
        win = np.array([[owners[q][1], problems[q].x0, problems[q].y0, problems[q].h,
                         problems[q].w] for q in qs], np.int32)
        fps.check_windows(win, nb, H, W, "blank_response_correct")
        phi = np.empty((qs.size, T, T), np.float32)
        for j, t in enumerate(window_psf_stamps(fld, problems, owners, qs, T)):
            t = np.clip(t.astype(np.float64), 0.0, None)
            phi[j] = (t / t.sum()).astype(np.float32)
        max_pix = int(max(problems[q].h * problems[q].w for q in qs))
        nw = int(L.uph_blank_work_floats(int(qs.size), max_pix))
        work = torch.empty(max(nw, 1), dtype=torch.float32, device=dev)
        bh = torch.empty(qs.size, dtype=torch.float64, device=dev)
        uw = (np.asarray(use_wing, np.uint8)[qs] if use_wing is not None
              else np.zeros(qs.size, np.uint8))
        keep = [up(off, np.int64), up(win, np.int32), up(ix, np.int32), up(iy, np.int32),
                up(flux, np.float32), up(bg, np.float32), up(phi, np.float32), up(uw, np.uint8)]
        _lib.check(L.uph_blank_response(
            int(qs.size), T, cfg.bg_order, P(keep[0]), P(keep[1]), H, W, P(fld.data),
            P(fld.invvar), P(fld.good), P(wing) if wing is not None else None, P(keep[7]),
            P(stamps), P(keep[2]), P(keep[3]), P(keep[4]), P(keep[5]), *map(float, bg_geo),
            P(keep[6]), P(work), nw, max_pix, P(bh), _lib.stream_ptr(None)))
        bq = bh.cpu().numpy()
        for j, q in enumerate(qs):
            b_hat[q] = bq[j]
            if not np.isfinite(bq[j]):
                continue
            star = stype[owners[q][0].fit_indices] == SOURCE_STAR
            r0 = out[q]
            f = r0.fluxes.copy()
            f[star] = f[star] - np.float32(bq[j])
            out[q] = fps.SparseForcedPhotResult(f, r0.bg_coefs, r0.flux_var, r0.bg_var,
                                                r0.converged)
    return out, b_hat


def window_psf_stamps(fld, problems, owners, qs, T):
    """The band PSF at the centre of each window ``qs`` (float32, padded to
    T x T), evaluated per band in one product."""
    from .psf_polynomial import evaluate_psf_at_batch_np
    out = [None] * len(qs)
    bands = np.array([owners[q][1] for q in qs])
    for b in np.unique(bands):
        sel = np.nonzero(bands == b)[0]
        xc = np.array([problems[qs[j]].x0 + 0.5 * problems[qs[j]].w for j in sel])
        yc = np.array([problems[qs[j]].y0 + 0.5 * problems[qs[j]].h for j in sel])
        st = evaluate_psf_at_batch_np(fld.psfs[int(b)], xc, yc)
        for k, j in enumerate(sel):
            out[j] = _pad_center(st[k], T)
    return out


def _pad_center(t, T):
    """Zero-pad (or keep) a square stamp to T x T, centred."""
    n = t.shape[0]
    if n == T:
        return t
    o = (T - n) // 2
    out = np.zeros((T, T), t.dtype)
    out[o:o + n, o:o + n] = t
    return out


def wing_peel(fld, T, problems, owners, res, x, y, stype, logre, ee1, ee2, sn,
              cfg: FitConfig, bg_geo, timer=None, blank=False, post=None):
    """Wing peeling of a grid pass at the native stamp (solver_grid.py:
    2495-2546, :2826-2846): per grid, the big sources' wings -- their S1
    flux times the tier-wing_tier render minus the native render -- are
    subtracted from the data before the final solve.  The final solve is
    linear in the data, so it equals the S1 solve minus the same solve of
    the wing image: the wing images are built on the device
    (``uph_wing_images``, grids of one colour at a time so windows are
    disjoint), solved, and their fluxes and background subtracted.  With
    ``blank`` (pass 2) the blank-response correction follows, against data
    minus each grid's own wing; ``post(sel, results, wing)`` (pass 2) is then
    called with the corrected results of the local problems ``sel`` and the
    wing image they saw.  Returns (results, problems corrected)."""
    import torch
    from . import _lib
    done = np.zeros(len(problems), bool)
    if T != fld.native:
        return res, done
    nb, H, W = fld.nb, fld.H, fld.W
    Tb, Ts = int(cfg.wing_tier), int(fld.native)
    big_of = {}
    for (g, b) in owners:
        if g.grid_id not in big_of:
            fi = g.fit_indices
            big_of[g.grid_id] = wing_sources(stype[fi], logre[fi], sn[fi], Ts,
                                             cfg.wing_max_sources)
    qs = [q for q, (g, b) in enumerate(owners) if big_of[g.grid_id].size]
    if not qs:
        return res, done
    colors = _disjoint_colors([g for g, _ in {id(g): (g, 0) for g, _ in owners}.values()])
    L = _lib.lib()
    P = _lib.ptr
    dev = fld.data.device
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    cb, gb = fld.psf(Tb)
    cs, gs = fld.psf(Ts)
    out = list(res)
    for c in sorted({colors[owners[q][0].grid_id] for q in qs}):
        sel = [q for q in qs if colors[owners[q][0].grid_id] == c]
        rows = []   # (problem, local index in its fit set, global source index)
        for q in sel:
            g, b = owners[q]
            for k in big_of[g.grid_id]:
                rows.append((q, int(k), int(g.fit_indices[k])))
        qa = np.array([r[0] for r in rows])
        src = np.array([r[2] for r in rows])
        band = np.array([owners[q][1] for q in qa], np.int32)
        ox = np.array([owners[q][0].x0_pad for q in qa], np.float32)
        oy = np.array([owners[q][0].y0_pad for q in qa], np.float32)
        cat = (x[src], y[src], stype[src], logre[src], ee1[src], ee2[src], sn[src])
        s_big = fps.render_unit_stamps(cb, gb, Tb, fld.n_terms, band, cat[0], cat[1], ox, oy,
                                       *cat[2:])
        s_small = fps.render_unit_stamps(cs, gs, Ts, fld.n_terms, band, cat[0], cat[1], ox, oy,
                                         *cat[2:])
        ix, iy = fps.stamp_top_left(cat[0], cat[1], ox, oy, Tb, Tb)
        fl = np.array([res[q].fluxes[k] for q, k, _ in rows], np.float32)
        fl = np.where(np.isfinite(fl), fl, 0.0).astype(np.float32)
        counts = np.bincount(np.searchsorted(sel, qa), minlength=len(sel))
        off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        win = np.array([[owners[q][1], problems[q].x0, problems[q].y0, problems[q].h,
                         problems[q].w] for q in sel], np.int32)
        fps.check_windows(win, nb, H, W, "wing_peel")
        wing = torch.zeros((nb, H, W), dtype=torch.float32, device=dev)
        keep = [up(off, np.int64), up(win, np.int32), up(ix, np.int32), up(iy, np.int32),
                up(fl, np.float32)]
        _lib.check(L.uph_wing_images(len(sel), Tb, Ts, P(keep[0]), P(keep[1]), P(s_big),
                                     P(s_small), P(keep[2]), P(keep[3]), P(keep[4]), H, W,
                                     P(wing), _lib.stream_ptr(None)))
        rw = fps.solve_problems([problems[q] for q in sel], cs, gs, Ts, fld.n_terms, wing,
                                fld.invvar, fld.good, bg_order=cfg.bg_order, bg_geo=bg_geo,
                                reg=cfg.reg, k_max=cfg.k_max, timer=timer)
        for q, r in zip(sel, rw):
            r0 = out[q]
            out[q] = fps.SparseForcedPhotResult(r0.fluxes - r.fluxes, r0.bg_coefs - r.bg_coefs,
                                                r0.flux_var, r0.bg_var,
                                                r0.converged and r.converged)
        if blank:
            use = np.zeros(len(problems), np.uint8)
            use[sel] = 1
            sub = [problems[q] for q in sel]
            corr, _ = blank_response_correct(fld, T, sub, [owners[q] for q in sel],
                                             [out[q] for q in sel], x, y, stype, logre, ee1,
                                             ee2, sn, cfg, bg_geo, wing=wing,
                                             use_wing=np.ones(len(sel), np.uint8))
            for q, r in zip(sel, corr):
                out[q] = r
            done[sel] = True
        if post is not None:
            post(sel, [out[q] for q in sel], wing)
            done[sel] = True
    return out, done


def linear_pass(fld: DeviceField, grids, x, y, stype, logre, ee1, ee2, sn, cfg: FitConfig,
                gtier: dict, timer=None, final=False, bg_geo=None, with_bg_var=False,
                conf_out=None, fcorr_out=None):
    """Per-grid, per-band linear solve; returns fluxes/var (n, nb) for each
    source from the grid that owns it, and bg coefs (n_grid, nb, M).  The
    grids, their fit sets and stamp tiers are those of pass 1 (pass 2 only
    replaces the source parameters, island_stage.py:3331-3363).  ``bg_geo``
    (x origin, x scale, y origin, y scale) of the background basis defaults
    to the field's centre and half-size (pipeline.py:1693-1696).  With
    ``final`` and ``conf_out`` (an (n, nb) array) the confusion flux error
    of each owned source is written there (error_model.confusion_sigma,
    after the blank-response correction; native stamps only); with
    ``fcorr_out`` (a dict) the residual correlated-noise estimate of each
    such (grid, band) window is stored there under (grid_id, band) as
    (F, n_clean, fallback)."""
    n = x.size
    nb = fld.nb
    M = (cfg.bg_order + 1) * (cfg.bg_order + 2) // 2
    flux = np.full((n, nb), np.nan, np.float32)
    var = np.full((n, nb), np.nan, np.float32)
    bg = np.zeros((len(grids), nb, M), np.float32)
    bgv = np.full((len(grids), nb, M), np.nan, np.float32)
    if bg_geo is None:
        bg_geo = (0.5 * (fld.W - 1), max(0.5 * fld.W, 1.0), 0.5 * (fld.H - 1),
                  max(0.5 * fld.H, 1.0))
    by_tier = {}
    for g in grids:
        by_tier.setdefault(gtier[g.grid_id], []).append(g)
    for T, gl in sorted(by_tier.items()):
        problems, owners = [], []
        conf_q = {}
        for g in gl:
            fi = g.fit_indices
            nbr, _ = fps.build_neighbor_list(x[fi], y[fi], T, T, cfg.k_max)
            # the bands of a grid share its source arrays
            px, py = x[fi].astype(np.float32), y[fi].astype(np.float32)
            gst, glr, ge1, ge2, gsn = stype[fi], logre[fi], ee1[fi], ee2[fi], sn[fi]
            x0, y0, h, w = g.x0_pad, g.y0_pad, g.pad_h, g.pad_w
            for b in range(nb):
                problems.append(fps.BandProblem(
                    band=b, x0=x0, y0=y0, h=h, w=w, pos_x=px, pos_y=py, nbr_idx=nbr,
                    stype=gst, logre=glr, ee1=ge1, ee2=ge2, sersic_n=gsn))
                owners.append((g, b))
        coefs, geo = fld.psf(T)
        # launches of at most stamp_budget bytes of rendered stamps each
        cost = np.array([pr.pos_x.size * T * T * 4.0 for pr in problems])
        cut = np.concatenate([[0], np.nonzero(np.diff(np.floor(np.cumsum(cost) /
                                                             cfg.stamp_budget)))[0] + 1,
                              [len(problems)]])
        res = []
        for a, z in zip(cut[:-1], cut[1:]):
            if z > a:
                part = fps.solve_problems(problems[a:z], coefs, geo, T, fld.n_terms, fld.data,
                                          fld.invvar, fld.good, bg_order=cfg.bg_order,
                                          bg_geo=bg_geo, reg=cfg.reg, k_max=cfg.k_max,
                                          timer=timer)
                blank = final and cfg.blank_response and T == fld.native
                conf = final and conf_out is not None and T == fld.native
                done = np.zeros(len(part), bool)

                def post(sel, rs, wing, a=a):
                    fq = {} if fcorr_out is not None else None
                    sig = confusion_sigma(fld, T, [problems[a + q] for q in sel],
                                          [owners[a + q] for q in sel], rs, x, y, stype, logre,
                                          ee1, ee2, sn, cfg, bg_geo, wing=wing,
                                          use_wing=(None if wing is None
                                                    else np.ones(len(sel), np.uint8)),
                                          fcorr_out=fq)
                    for j, (q, sg) in enumerate(zip(sel, sig)):
                        conf_q[a + q] = sg
                        if fq is not None:
                            g_, b_ = owners[a + q]
                            fcorr_out[(g_.grid_id, b_)] = fq[j]

                if cfg.wing_peeling:
                    part, done = wing_peel(fld, T, problems[a:z], owners[a:z], part, x, y,
                                           stype, logre, ee1, ee2, sn, cfg, bg_geo, timer,
                                           blank=blank, post=post if conf else None)
                if (blank or conf) and not done.all():
                    rest = np.nonzero(~done)[0]
                    if blank:
                        corr, _ = blank_response_correct(
                            fld, T, [problems[a + q] for q in rest],
                            [owners[a + q] for q in rest], [part[q] for q in rest], x, y,
                            stype, logre, ee1, ee2, sn, cfg, bg_geo)
                        for q, r in zip(rest, corr):
                            part[q] = r
                    if conf:
                        post(list(rest), [part[q] for q in rest], None)
                res += part
        for q, ((g, b), r) in enumerate(zip(owners, res)):
            own = g.source_indices.size
            if conf_out is not None and q in conf_q:
                conf_out[g.source_indices, b] = conf_q[q][:own]
            flux[g.source_indices, b] = r.fluxes[:own]
            var[g.source_indices, b] = r.flux_var[:own]
            bg[g.grid_id, b] = r.bg_coefs
            bgv[g.grid_id, b] = r.bg_var
    if with_bg_var:
        return flux, var, bg, bgv
    return flux, var, bg


def padded_stacks(fld: DeviceField, grids, bg, bg_order, pad):
    """(dstack, wstack) (nb, H+2 pad, W+2 pad): data minus each grid's S1
    background on its core (``uph_subtract_bg_planes``), inverse variance;
    zero border."""
    import torch
    nb, H, W = fld.nb, fld.H, fld.W
    dev = fld.data.device
    d = torch.zeros((nb, H + 2 * pad, W + 2 * pad), dtype=torch.float32, device=dev)
    w = torch.zeros_like(d)
    d[:, pad:pad + H, pad:pad + W] = fld.data
    w[:, pad:pad + H, pad:pad + W] = fld.invvar
    xo, xs = 0.5 * (W - 1), max(0.5 * W, 1.0)
    yo, ys = 0.5 * (H - 1), max(0.5 * H, 1.0)
    from . import _lib
    if grids:
        win = np.array([[g.x0_core, g.y0_core, g.y1_core - g.y0_core, g.x1_core - g.x0_core]
                        for g in grids], np.int32)
        fps.check_windows(np.concatenate([np.zeros((len(grids), 1), np.int32), win], 1),
                          1, H, W, "padded_stacks")
        coef = np.stack([np.asarray(bg[g.grid_id], np.float32) for g in grids])
        t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (win, coef)]
        max_pix = int((win[:, 2] * win[:, 3]).max())
        _lib.check(_lib.lib().uph_subtract_bg_planes(
            len(grids), max_pix, nb, int(bg_order), _lib.ptr(t[0]), _lib.ptr(t[1]), int(pad),
            xo, xs, yo, ys, H + 2 * pad, W + 2 * pad, _lib.ptr(d), _lib.stream_ptr(None)))
    return d, w


@dataclass
class IslandSetFit:
    """Per-source results of the island LM over an island set."""
    x: np.ndarray
    y: np.ndarray
    logre: np.ndarray
    ee1: np.ndarray
    ee2: np.ndarray
    sersic_n: np.ndarray      # fitted n of free-n members, the catalog n otherwise
    p: np.ndarray             # (n, 6) fitted parameter vector (x, y, logre, ee1, ee2, u)
    flux: np.ndarray          # (n, nb) island VarPro fluxes, NaN where not fitted
    flux_err: np.ndarray
    conv: np.ndarray
    stall: np.ndarray
    nst: np.ndarray
    fitted: np.ndarray        # (n,) bool: member of a jointly fitted island
    n_fitted_islands: int
    classes: list             # (plan, host) per class
    chunks: list              # (n_islands, seconds) per launch sequence; (-n, 0) marks n
                              # islands refitted on float32 Gram operands
    # post-fit pass (cfg.post_fit; NaN where not computed)
    param_var: np.ndarray = None      # (n, 7) var x, y, logre, ee1, ee2, u, cov(ee1, ee2)
    flux_err_marg: np.ndarray = None  # (n, nb) sqrt(flux_err^2 + shape-marginal increment)
    metrics: np.ndarray = None        # (n, 4) rchisq, fracflux, fracin, fracmasked
    flux_window: np.ndarray = None    # (n, nb) flux times the unit model's window sum
    fp64_refit: np.ndarray = None     # (n,) bool: island refitted on the float64 kernels
    n_fp64_islands: int = 0
    kappa: np.ndarray = None          # (n,) the island's scaled Gauss-Newton condition number
    class_results: list = None        # (plan, host, rows, result dict) per launch part
    flux_marg_var: np.ndarray = None  # (n, nb) shape-marginalisation variance increment
    re_upper_pix: np.ndarray = None   # (n,) exp(upper bound of logre)
    diverged: np.ndarray = None       # (n,) stalled on identical (clip-pinned) proposals
    stall_dchi2: np.ndarray = None    # (n,) best damped-step chi2 change of a stalled island
    chi2_nu: np.ndarray = None        # (n,) the island's chi2 / (good pixels - free parameters)


def metrics_from_sums(sums):
    """Per-source (rchisq, fracflux, fracin, fracmasked) and the per-band
    window flux from the post-fit pass's per-band sums (..., nb, 6) = [sum m,
    sum (f m)^2, sum (model - f m) f m, sum m on zero-weight pixels,
    sum m w (d - model)^2, f], combined over bands as the reference does
    (island_solver.py:1672-1713): fracflux and fracmasked fin-weighted with
    fin = clip(sum m, 0, 1), fracin and rchisq as ratios of band sums."""
    sm, own2, nown, msk, rc, f = (sums[..., q].astype(np.float64) for q in range(6))
    fin = np.clip(sm, 0.0, 1.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        ffr = np.where(own2 > 0, nown / own2, 0.0)
        ff_n, ff_d = (fin * ffr).sum(-1), fin.sum(-1)
        fin_n, fin_d = np.abs(sm).sum(-1), np.ones_like(sm).sum(-1)
        fm_n = (fin * np.clip(msk, 0.0, 1.0)).sum(-1)
        rc_n, rc_d = np.abs(rc).sum(-1), np.abs(sm).sum(-1)
        fracflux = np.where(ff_d > 0, ff_n / ff_d, 0.0)
        fracin = np.where(fin_d > 0, fin_n / fin_d, np.nan)
        fracmasked = np.where(ff_d > 0, fm_n / ff_d, 0.0)
        rchisq = np.where(rc_d > 0, rc_n / rc_d, np.nan)
    return np.stack([rchisq, fracflux, fracin, fracmasked], axis=-1), f * sm


def _chunk_batch(batch, nb, budget_bytes, mosaic=False):
    """Split a class batch [(kernels, host, rows)] into launch sequences whose
    pre-render planes (and, with ``mosaic``, the island cutout mosaic) fit
    ``budget_bytes``; classes are taken in order of their per-workgroup
    scratch so large classes share sequences."""
    from . import _lib
    L = _lib.lib()
    keyed = []
    for kern, host, rows in batch:
        per = L.uph_island_scratch_per_block(kern.N, kern.B, kern.T, kern.n_live_max, 0)
        side = min(kern.T, kern.B)
        st = host["st"][rows]
        mm = host["mmask"][rows].astype(bool)
        planes = np.where(st == SOURCE_STAR, 3, 7) * mm
        cost = nb * side * side * 4.0 * planes.sum(axis=1)
        if mosaic:
            cost = cost + mosaic_bytes(nb, kern.B)
        keyed.append((per, kern, host, np.asarray(rows), cost))
    keyed.sort(key=lambda t: t[0])
    chunks, cur, used = [], [], 0.0
    for per, kern, host, rows, cost in keyed:
        start = 0
        while start < rows.size:
            room = budget_bytes - used
            take = int(np.searchsorted(np.cumsum(cost[start:]), room, side="right"))
            if take == 0 and cur:
                chunks.append(cur)
                cur, used = [], 0.0
                continue
            take = max(take, 1)
            cur.append((kern, host, rows[start:start + take]))
            used += float(cost[start:start + take].sum())
            start += take
    if cur:
        chunks.append(cur)
    return chunks


def island_contrast(plan, rows, flux, sigma) -> np.ndarray:
    """Per island row of a class: max over bands of the brightest member's
    |flux| over the faintest member's max(|flux|, sigma); 1 for a single
    member.  ``flux``, ``sigma`` (n_sources, n_bands); NaN counts as 0 flux
    and infinite sigma."""
    rows = np.asarray(rows)
    mm = plan.mmask[rows].astype(bool)
    idx = np.where(mm, plan.midx[rows], 0)
    f = np.abs(np.nan_to_num(np.asarray(flux, np.float64)[idx], nan=0.0))
    s = np.nan_to_num(np.asarray(sigma, np.float64)[idx], nan=np.inf)
    den = np.where(mm[:, :, None], np.maximum(f, s), np.inf)
    num = np.where(mm[:, :, None], f, 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        c = num.max(axis=1) / den.min(axis=1)
    c = np.where(np.isfinite(c), c, 1.0)
    return np.maximum(c.max(axis=1), 1.0) if c.size else np.ones(rows.size)


def _island_blocks(cfg, part32, part64, trees, nb, ds, ws, sersic_dev, post, sync):
    """Persistent workgroups per precision for one launch sequence: the
    configured ones, or the autotuner's choice for this workload (measured
    on the sequence itself the first time, then cached)."""
    from . import autotune
    from .island_solver import fit_float64, run_island_batch
    if cfg.island_blocks is not None or not cfg.autotune or not cfg.island_split:
        return dict(cfg.island_blocks or {})
    import torch
    sync = sync or torch.cuda.synchronize
    out = {}

    def timed(fn):
        sync()
        t0 = time.perf_counter()
        fn()
        sync()
        return time.perf_counter() - t0

    for prec, part in ((cfg.precision, part32), ("fp64", part64)):
        if not part:
            continue
        n_isl = int(sum(np.asarray(r).size for _, _, r in part))
        tiers = sorted({int(k.T) for k, _, _ in part})
        mixed = any(k.kind != "star" for k, _, _ in part)
        key = autotune.workload_key(prec, nb, tiers, mixed, n_isl, ds.device)
        if prec == "fp64" and part is part64:
            run = lambda c: timed(lambda: fit_float64(
                part64, trees, cfg.max_steps, ds, ws, sersic_dev=sersic_dev, post=False,
                workspace_bytes=cfg.island_workspace, max_blocks=c))
        else:
            run = lambda c: timed(lambda: run_island_batch(
                part32, trees, cfg.max_steps, ds, ws, precision=cfg.precision,
                sersic_dev=sersic_dev, split=True, prerender=cfg.galaxy_prerender, post=post,
                gram=cfg.gram, workspace_bytes=cfg.island_workspace, max_blocks=c))
        out[prec] = autotune.choose(key, run)
    return out


def fit_island_set(psf_of, nb, n_terms, native, dstack, wstack, pad, iset, plans, x, y, stype,
                   logre, ee1, ee2, sn, free_n, cfg: FitConfig, timer=None,
                   budget_bytes: float = 48e9, sync=None, p_start=None,
                   scene: SceneStars | None = None, flux_scale=None,
                   keep_class_results: bool = False) -> IslandSetFit:
    """Island VarPro-LM over the planned classes of an island set
    (island_solver.py:3233-3595, one device): class host arrays, the device
    Sersic table and PSF trees, then the band-split fit in launch sequences
    of at most ``budget_bytes`` of pre-rendered planes each.  ``psf_of(T)``
    gives the device (coefs, geo) of stamp tier T.  ``p_start`` (n, 6), when
    given, replaces the starting parameters of the real members after the
    classes (stamp tiers, bounds) were planned from the catalog seeds.  With
    ``scene`` (scene subtraction), dstack already has every catalogued star
    subtracted at its S1 flux and each launch sequence fits a mosaic of its
    island cutouts with the members' own stars added back.  ``flux_scale``
    = (flux, sigma) per source and band (the pass-1 solve) routes the
    islands above ``cfg.fp64_contrast`` to the float64 kernels."""
    from .island_solver import PsfTree, run_island_batch
    n = x.size
    fx, fy = x.copy(), y.copy()
    flr, fe1, fe2 = logre.copy(), ee1.copy(), ee2.copy()
    fsn = np.asarray(sn, np.float32).copy()
    pfit = np.full((n, 6), np.nan, np.float32)
    flux = np.full((n, nb), np.nan, np.float32)
    ferr = np.full((n, nb), np.nan, np.float32)
    conv = np.zeros(n, bool)
    stall = np.zeros(n, bool)
    nst = np.zeros(n, np.int32)
    fitted = np.zeros(n, bool)
    pvar = np.full((n, 7), np.nan, np.float32)
    ferr_m = np.full((n, nb), np.nan, np.float32)
    mets = np.full((n, 4), np.nan, np.float32)
    fwin = np.full((n, nb), np.nan, np.float32)
    dvar = np.full((n, nb), np.nan, np.float32)
    re_up = np.full(n, np.nan, np.float32)
    f64src = np.zeros(n, bool)
    kap = np.full(n, np.nan, np.float32)
    div = np.zeros(n, bool)
    sdc = np.full(n, np.nan, np.float32)
    c2nu = np.full(n, np.nan, np.float32)
    classes, batch, trees = [], [], {}
    for plan in plans:
        host = class_host_arrays(plan, pos_x=x, pos_y=y, source_type=stype, logre=logre,
                                 ee1=ee1, ee2=ee2, sersic_n=sn, free_n_all=free_n,
                                 pos_max_shift_pix=cfg.pos_max_shift_pix, pad=pad)
        kern = make_class_kernels(plan, nb, n_terms, reg=cfg.reg, precision=cfg.precision)
        if p_start is not None:
            mm = plan.mmask.astype(bool)
            host["p0"][mm] = np.asarray(p_start, np.float32)[plan.midx[mm]]
        if plan.tier not in trees:
            coefs, geo = psf_of(plan.tier)
            trees[plan.tier] = PsfTree(coefs, geo, plan.tier, n_terms)
        batch.append((kern, host, np.arange(plan.n_islands)))
        classes.append((plan, host))
    sersic_dev = None
    if any(p.kind != "star" for p in plans):
        from .sersic_table import device_table
        sersic_dev = device_table(dstack.device)
    if cfg.island_split and cfg.galaxy_prerender and cfg.precision != "fp64":
        chunks = _chunk_batch(batch, nb, budget_bytes, mosaic=scene is not None)
    else:
        chunks = [batch]
    plan_of = {id(h): p for p, h in classes}
    timing = []
    # the float64 refit reads the conditioning the post-fit pass measures
    refit_on = (cfg.fp64_refit_kappa is not None and cfg.precision == "fp32" and
                cfg.island_split and cfg.galaxy_prerender)
    route_on = (refit_on and cfg.fp64_contrast is not None and flux_scale is not None)
    n_fp64 = 0
    kept = []
    for chunk in chunks:
        if not chunk:
            continue
        t0 = time.perf_counter()
        ds, ws = dstack, wstack
        if scene is not None:
            ds, ws, moved = island_mosaic(chunk, plan_of, dstack, wstack, scene)
            for (_, h2, _), (_, h, _) in zip(moved, chunk):
                plan_of[id(h2)] = plan_of[id(h)]
            chunk = moved
        part64 = []
        if route_on:
            part32 = []
            for kern, host, rows in chunk:
                rows = np.asarray(rows)
                hi = island_contrast(plan_of[id(host)], rows, *flux_scale) > cfg.fp64_contrast
                if (~hi).any():
                    part32.append((kern, host, rows[~hi]))
                if hi.any():
                    part64.append((kern, host, rows[hi]))
            chunk = part32
        blocks = _island_blocks(cfg, chunk, part64, trees, nb, ds, ws, sersic_dev, refit_on,
                                sync)
        results = run_island_batch(chunk, trees, cfg.max_steps, ds, ws,
                                   precision=cfg.precision, timer=timer, sersic_dev=sersic_dev,
                                   split=cfg.island_split, prerender=cfg.galaxy_prerender,
                                   post=cfg.post_fit or refit_on, gram=cfg.gram,
                                   workspace_bytes=cfg.island_workspace,
                                   max_blocks=blocks.get(cfg.precision))
        redo = [(kern, host, rows[res["gram_unsettled"]])
                for (kern, host, rows), res in zip(chunk, results) if "gram_unsettled" in res
                and res["gram_unsettled"].any()]
        if redo:
            # islands whose bfloat16 Gram refinement did not settle, refitted
            # on float32 operands
            again = run_island_batch(redo, trees, cfg.max_steps, ds, ws,
                                     precision=cfg.precision, sersic_dev=sersic_dev,
                                     split=cfg.island_split, prerender=cfg.galaxy_prerender,
                                     post=cfg.post_fit, gram="fp32",
                                     workspace_bytes=cfg.island_workspace, timer=timer,
                                     max_blocks=blocks.get(cfg.precision))
            pos = {id(h): j for j, (_, h, _) in enumerate(redo)}
            merged = []
            for (kern, host, rows), res in zip(chunk, results):
                if id(host) in pos and "gram_unsettled" in res and res["gram_unsettled"].any():
                    sel = np.nonzero(res["gram_unsettled"])[0]
                    ra = again[pos[id(host)]]
                    res = {k: (v.copy() if isinstance(v, np.ndarray) else v)
                           for k, v in res.items()}
                    for k, v in ra.items():
                        if isinstance(v, np.ndarray) and k in res and res[k].shape[0] == rows.size:
                            res[k][sel] = v
                merged.append(res)
            results = merged
            n_redo = sum(int(r.size) for _, _, r in redo)
            timing.append((-n_redo, 0.0))
        if refit_on:
            results, n64 = refit_ill_conditioned(chunk, results, trees, cfg.max_steps, ds, ws,
                                                 kappa=cfg.fp64_refit_kappa,
                                                 sersic_dev=sersic_dev,
                                                 workspace_bytes=cfg.island_workspace,
                                                 timer=timer, max_blocks=blocks.get("fp64"))
            n_fp64 += n64
        if part64:
            results = results + fit_float64(part64, trees, cfg.max_steps, ds, ws,
                                             sersic_dev=sersic_dev, post=cfg.post_fit,
                                             workspace_bytes=cfg.island_workspace, timer=timer,
                                             max_blocks=blocks.get("fp64"))
            chunk = chunk + part64
            n_fp64 += sum(int(r.size) for _, _, r in part64)
        if cfg.stall_probes and any("post_hg" in r for r in results):
            stall_probes(chunk, results, trees, ds, ws, sersic_dev=sersic_dev,
                         workspace_bytes=cfg.island_workspace, timer=timer)
        if sync is not None:
            sync()
        timing.append((int(sum(np.asarray(r).size for _, _, r in chunk)),
                       time.perf_counter() - t0))
        for (kern, host, rows), res in zip(chunk, results):
            plan = plan_of[id(host)]
            if keep_class_results:
                kept.append((plan, host, np.asarray(rows), res))
            mm = plan.mmask[rows]
            src = plan.midx[rows][mm]
            fx[src] = res["p"][:, :, 0][mm]
            fy[src] = res["p"][:, :, 1][mm]
            ext = (stype[src] != SOURCE_STAR)
            flr[src[ext]] = res["p"][:, :, 2][mm][ext]
            fe1[src[ext]] = res["p"][:, :, 3][mm][ext]
            fe2[src[ext]] = res["p"][:, :, 4][mm][ext]
            fnm = host["fn_flag"][rows][mm].astype(bool)
            fsn[src[fnm]] = sersic_u_to_n(res["p"][:, :, 5][mm][fnm]).astype(np.float32)
            pfit[src] = res["p"][mm]
            flux[src] = res["fluxes"].transpose(0, 2, 1)[mm]
            ferr[src] = res["ferr"].transpose(0, 2, 1)[mm]
            r_of = np.repeat(np.arange(rows.size), mm.sum(axis=1))
            conv[src] = res["conv"][r_of]
            stall[src] = res["stall"][r_of]
            nst[src] = res["nst"][r_of]
            fitted[src] = True
            re_up[src] = np.exp(host["upper"][rows][:, :, 2][mm].astype(np.float64))
            if "fp64_refit" in res:
                f64src[src] = res["fp64_refit"][r_of]
            if "post_kappa" in res:
                kap[src] = res["post_kappa"][r_of]
            if "stall_dchi2" in res:
                div[src] = res["diverged"][r_of]
                sdc[src] = res["stall_dchi2"][r_of]
            if "n_good" in res:
                # reduced chi2 of the island (island_plan.py:888-911): dof = good pixels -
                # fitted nonlinear parameters - n_bands x members
                n_par = (host["keep"][rows] & mm[:, :, None]).reshape(rows.size, -1).sum(axis=1)
                dof = (np.asarray(res["n_good"], np.float64) - n_par
                       - float(res["fluxes"].shape[1]) * mm.sum(axis=1))
                c2 = np.asarray(res["chi2"], np.float64)
                with np.errstate(invalid="ignore", divide="ignore"):
                    cn = np.where((dof > 0) & np.isfinite(c2), c2 / np.maximum(dof, 1.0), np.nan)
                c2nu[src] = cn[r_of]
            if "post_var" in res:
                pvar[src] = res["post_var"][mm]
                dv = res["post_dvar"].transpose(0, 2, 1)[mm]
                ferr_m[src] = np.sqrt(np.square(ferr[src]) + dv)
                dvar[src] = dv
                m4, fw = metrics_from_sums(res["post_metrics"][mm])
                mets[src] = m4
                fwin[src] = fw
    n_fit = sum(p.n_islands for p in plans)
    return IslandSetFit(fx, fy, flr, fe1, fe2, fsn, pfit, flux, ferr, conv, stall, nst, fitted,
                        n_fit, classes, timing, pvar, ferr_m, mets, fwin, fp64_refit=f64src,
                        n_fp64_islands=n_fp64, kappa=kap, diverged=div, stall_dchi2=sdc,
                        class_results=kept if keep_class_results else None,
                        flux_marg_var=dvar, re_upper_pix=re_up, chi2_nu=c2nu)


def fit_field(fld: DeviceField, x, y, stype, logre, ee1, ee2, sersic_n, cfg: FitConfig,
              sync=None, timer=None, keep_artifacts=False) -> FitResult:
    """Full island-mode fit of one field (catalog seeds in, fitted fluxes
    out).  ``sync`` is called between stages for timing (e.g.
    torch.cuda.synchronize)."""
    sync = sync or (lambda: None)
    tm = {}
    t0 = time.perf_counter()
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    stype = np.asarray(stype, np.int32)
    logre = np.asarray(logre, np.float32)
    ee1 = np.asarray(ee1, np.float32)
    ee2 = np.asarray(ee2, np.float32)
    sn = np.asarray(sersic_n, np.float32)
    n = x.size
    tiers = tuple(sorted({int(t) for t in cfg.stamp_tiers} | {fld.native}))
    halo = cfg.halo
    if np.any(stype != SOURCE_STAR):
        halo = max(halo, int(_src_tier(stype, logre, sn, tiers).max()) // 2 + 3)
    grids = divide_image_into_grids(fld.H, fld.W, cfg.n_rows, cfg.n_cols, halo)
    assign_sources_to_grids(grids, x, y)
    gtier = grid_tiers(fld, grids, stype, logre, sn, tiers)
    s1_flux, s1_var, bg = linear_pass(fld, grids, x, y, stype, logre, ee1, ee2, sn, cfg, gtier,
                                      timer)
    sync()
    tm["pass1"] = time.perf_counter() - t0
    # ---- islands ----
    t1 = time.perf_counter()
    r_psf = fld.native / 2.0
    if cfg.link_radius is not None:
        radii = np.full(n, float(cfg.link_radius))
    else:
        ref = fld.nb // 2
        stamp = evaluate_psf_at_np(fld.psfs[ref], fld.W / 2.0, fld.H / 2.0)
        pr, pp = psf_radial_profile(stamp)
        with np.errstate(invalid="ignore"):
            fmed = np.nanmedian(s1_flux, axis=1)
        radii = flux_aware_radii(stype, logre, sn, fmed, fld.sigma_sky, cfg.kappa, pr, pp,
                                 r_psf)
    iset = build_islands(x, y, stype, logre, sn, r_psf, radius=radii)
    buckets = bucket_islands(iset, ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES)
    plans, capped = plan_island_classes(iset, buckets, stype, logre, sn, fld.native)
    giants = np.concatenate([buckets.giant_ids, np.asarray(capped, np.int64)]).astype(np.int64)
    free_n = free_n_gate(stype, np.nan_to_num(s1_flux, nan=0.0), np.nan_to_num(s1_var, nan=1e30))
    pad = int(max((p.bbox_class for p in plans), default=32)) + 8
    dstack, wstack = padded_stacks(fld, grids, bg, cfg.bg_order, pad)
    dstack_raw = dstack.clone() if keep_artifacts and cfg.scene_subtract else dstack
    scene = None
    if cfg.scene_subtract:
        coefs, geo = fld.psf(fld.native)
        subtract_stars(dstack, pad, coefs, geo, fld.native, fld.n_terms, x, y, stype, s1_flux)
        scene = SceneStars(x, y, stype, star_flux(s1_flux), fld.psf, fld.n_terms)
    sync()
    tm["islands"] = time.perf_counter() - t1
    # ---- island LM ----
    t2 = time.perf_counter()
    isf = fit_island_set(fld.psf, fld.nb, fld.n_terms, fld.native, dstack, wstack, pad, iset,
                         plans, x, y, stype, logre, ee1, ee2, sn, free_n, cfg, timer=timer,
                         scene=scene, flux_scale=(s1_flux, np.sqrt(np.maximum(s1_var, 0.0))))
    fx, fy, flr, fe1, fe2 = isf.x, isf.y, isf.logre, isf.ee1, isf.ee2
    fsn = isf.sersic_n
    conv, stall, nst = isf.conv, isf.stall, isf.nst
    giant = np.zeros(n, bool)
    for gid in giants:
        giant[iset.members[int(gid)]] = True
    n_fitted_islands = isf.n_fitted_islands
    classes = isf.classes
    sync()
    tm["island_fit"] = time.perf_counter() - t2
    # ---- pass 2 ----
    t3 = time.perf_counter()
    conf = np.full((n, fld.nb), np.nan, np.float32) if cfg.confusion else None
    fc_res = {} if (cfg.confusion and cfg.correlated_noise) else None
    flux, var, _ = linear_pass(fld, grids, fx, fy, stype, flr, fe1, fe2, fsn, cfg, gtier, timer,
                               final=True, conf_out=conf, fcorr_out=fc_res)
    fcorr = None
    if cfg.correlated_noise:
        from .fcorr import choose_band_fcorr, fcorr_sky
        raw = fcorr_sky(fld)
        per_grid = []
        for g in grids:
            per_grid.append([(fc_res or {}).get((g.grid_id, b)) for b in range(fld.nb)])
        fcorr, fc_src, fc_resid = choose_band_fcorr(per_grid, np.array([f for f, _ in raw]))
    sync()
    tm["pass2"] = time.perf_counter() - t3
    tm["total"] = time.perf_counter() - t0
    owner = np.full(n, -1, np.int64)
    for g in grids:
        owner[g.source_indices] = g.grid_id
    rec = dict(n_sources=n, n_islands=int(iset.n_islands), n_fitted_islands=n_fitted_islands,
               n_giant_islands=int(giants.size), n_fp64_refit_islands=int(isf.n_fp64_islands),
               classes=[
                   (p.bbox_class, p.nsrc_class, p.kind, p.tier, p.n_islands) for p in plans],
               knobs_off=["priors"] + ([] if cfg.scene_subtract else ["scene_subtract"])
               + ([] if cfg.blank_response else ["blank_response_correction"])
               + ([] if cfg.wing_peeling else ["wing_peeling"])
               + ([] if cfg.confusion else ["confusion"])
               + ([] if cfg.correlated_noise else ["correlated_noise"])
               + ([] if cfg.post_fit else ["param_errors", "source_metrics",
                                           "marginal_flux_err"]))
    if fcorr is not None:
        rec["fcorr"] = dict(F=[float(f) for f in fcorr], source=fc_src,
                            residual=[float(f) for f in fc_resid],
                            raw_sky=[float(f) for f, _ in raw],
                            raw_sky_fallback=[d.get("fallback") for _, d in raw])
    if keep_artifacts:
        rec["_artifacts"] = dict(classes=classes, dstack=dstack, dstack_raw=dstack_raw,
                                 scene=scene, wstack=wstack, pad=pad,
                                 grids=grids, island_set=iset, gtier=gtier, free_n=free_n,
                                 giants=giants)
    return FitResult(flux, np.sqrt(np.maximum(var, 0.0)), s1_flux, s1_var, fx, fy, flr, fe1,
                     fe2, fsn, iset.labels.astype(np.int64), conv, stall, nst, giant, tm, rec,
                     isf.param_var, isf.flux, isf.flux_err_marg, isf.metrics, isf.flux_window,
                     isf.diverged, isf.stall_dchi2,
                     stall & (isf.stall_dchi2 > -2.0 * PLATEAU_DLNP_TOL), isf.fp64_refit,
                     flux_var=var, flux_marg_var=isf.flux_marg_var,
                     re_upper_pix=isf.re_upper_pix, free_n=np.asarray(free_n, bool),
                     grid_id=owner, flux_err_conf=conf, fcorr=fcorr, chi2_nu=isf.chi2_nu)


def psf_bank(psfs, device="cuda"):
    """``psf_of(T)``: device (coefs, geo) of the band PSFs padded to tier T,
    built once per tier."""
    import torch
    cache = {}

    def psf_of(T):
        if T not in cache:
            c, g = psf_device_arrays(psfs, T)
            cache[T] = (torch.from_numpy(c).to(device), torch.from_numpy(g).to(device))
        return cache[T]
    return psf_of


def fit_cutouts(cut, cfg: FitConfig, timer=None, sync=None, joint_nsrc_cap=None,
                budget_bytes: float = 48e9, psf_of=None, island_ids=None,
                p_start=None) -> IslandSetFit:
    """Batched island fit of independent island cutouts (each island in its
    own zero-weight-bordered cell of a mosaic; see scenes/cutouts.py): class
    planning with every island fitted jointly up to 64 members
    (``joint_nsrc_cap=None``), then :func:`fit_island_set`.  With no linear
    pass before it, the free-n gate's S/N is the true flux over the
    sky-limited noise of a 50 px^2 aperture.  ``island_ids`` restricts the
    fit to those islands."""
    iset = cut.island_set
    if island_ids is not None:
        iset = sub_island_set(iset, island_ids)
    native = int(cut.psfs[0].H_psf)
    buckets = bucket_islands(iset, ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES)
    plans, capped = plan_island_classes(iset, buckets, cut.stype, cut.logre, cut.sersic_n,
                                        native, joint_nsrc_cap=joint_nsrc_cap)
    var = (np.asarray(cut.sky, np.float64) + 25.0)[None, :] * 50.0
    free_n = free_n_gate(cut.stype, cut.true_flux, np.broadcast_to(var, cut.true_flux.shape))
    psf_of = psf_of or psf_bank(cut.psfs, cut.dstack.device)
    return fit_island_set(psf_of, cut.n_bands, int(cut.psfs[0].n_terms), native, cut.dstack,
                          cut.wstack, cut.pad, iset, plans, cut.x, cut.y, cut.stype, cut.logre,
                          cut.ee1, cut.ee2, cut.sersic_n, free_n, cfg, timer=timer,
                          budget_bytes=budget_bytes, sync=sync, p_start=p_start,
                          flux_scale=cutout_flux_scale(cut))


def cutout_flux_scale(cut):
    """(flux, sigma) per source and band of island cutouts for the float64
    routing: with no linear pass before the fit, the true flux and the
    sky-limited noise of a 50 px^2 aperture (the free-n gate's S/N)."""
    var = (np.asarray(cut.sky, np.float64) + 25.0)[None, :] * 50.0
    f = np.asarray(cut.true_flux, np.float64)
    return f, np.sqrt(np.broadcast_to(var, f.shape))


def _rel_stats(f, ref, ref_err):
    rel = np.abs(f - ref) / np.maximum(np.abs(ref), ref_err)
    r = np.nanmax(rel, axis=1) if rel.ndim == 2 else rel
    r = r[np.isfinite(r)]
    if r.size == 0:
        return dict(n=0)
    return dict(n=int(r.size), p50=float(np.percentile(r, 50)), p99=float(np.percentile(r, 99)),
                max=float(r.max()), frac_le_1e5=float((r <= 1e-5).mean()))


def sub_island_set(iset, island_ids):
    """The islands ``island_ids`` of an island set (members, boxes, sizes)."""
    from .islands import IslandSet
    ids = np.asarray(island_ids, np.int64)
    return IslandSet(n_sources=iset.n_sources, n_islands=ids.size, labels=iset.labels,
                     radius=iset.radius, members=[iset.members[i] for i in ids],
                     n_src=iset.n_src[ids], bbox=iset.bbox[ids], r_psf=iset.r_psf)


def island_accuracy(psf_of, nb, n_terms, native, dstack, wstack, pad, iset, island_ids, cat,
                    free_n, cfg: FitConfig, grams=("fp32",), sync=None,
                    joint_nsrc_cap=JOINT_NSRC_CAP, scene: SceneStars | None = None,
                    flux_scale=None) -> dict:
    """Flux error of float32 island fits (one per Gram operand precision in
    ``grams``) against the float64 fit of the same islands, per source as
    max over bands of |f - f64| / max(|f64|, sigma64):
      * ``fit``: full LM fits.  Where the chi2 landscape of a blend is flat,
        the trajectories may stop at different positions and the fluxes
        follow them.
      * ``flux_solve``: the flux solve (no LM step) at the float64 fit's
        parameters -- the arithmetic precision of the fluxes.
    ``cat`` = (x, y, stype, logre, ee1, ee2, sersic_n) catalog seeds;
    ``scene`` and ``flux_scale`` as for :func:`fit_island_set` (the float32
    fits then route their high-contrast islands to the float64 kernels, as
    the field fit does)."""
    import dataclasses
    x, y, st, lr, e1, e2, sn = cat
    sub = sub_island_set(iset, island_ids)
    buckets = bucket_islands(sub, ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES)
    plans, _ = plan_island_classes(sub, buckets, st, lr, sn, native,
                                   joint_nsrc_cap=joint_nsrc_cap)

    def run(c, p_start=None):
        return fit_island_set(psf_of, nb, n_terms, native, dstack, wstack, pad, sub, plans, x,
                              y, st, lr, e1, e2, sn, free_n, c, sync=sync, p_start=p_start,
                              scene=scene, flux_scale=flux_scale)
    c64 = dataclasses.replace(cfg, precision="fp64", gram="fp32", post_fit=False)
    f64 = run(c64)
    fz64 = run(dataclasses.replace(c64, max_steps=0), f64.p)
    out = dict(n_islands=int(sub.n_islands),
               flux_solve_fp64_self=None)
    ok = f64.fitted
    out["flux_solve_fp64_self"] = _rel_stats(fz64.flux[ok], f64.flux[ok], f64.flux_err[ok])
    for g in grams:
        c32 = dataclasses.replace(cfg, precision="fp32", gram=g, post_fit=False)
        f32 = run(c32)
        fz = run(dataclasses.replace(c32, max_steps=0), f64.p)
        ok = f32.fitted & f64.fitted
        same = ok & (f32.nst == f64.nst) & (f32.conv == f64.conv) & (f32.stall == f64.stall)
        out[g] = dict(same_trajectory=float(same.sum() / max(ok.sum(), 1)),
                      fit=_rel_stats(f32.flux[ok], f64.flux[ok], f64.flux_err[ok]),
                      flux_solve=_rel_stats(fz.flux[ok], f64.flux[ok], f64.flux_err[ok]),
                      position_max_px=float(np.max(np.hypot(f32.x[ok] - f64.x[ok],
                                                            f32.y[ok] - f64.y[ok]),
                                                   initial=0.0)))
    return out


def cutout_accuracy(cut, cfg: FitConfig, island_ids, sync=None, psf_of=None,
                    grams=("fp32",)) -> dict:
    """:func:`island_accuracy` on island cutouts (every island fitted
    jointly, free-n gate from the true flux as in :func:`fit_cutouts`)."""
    psf_of = psf_of or psf_bank(cut.psfs, cut.dstack.device)
    var = (np.asarray(cut.sky, np.float64) + 25.0)[None, :] * 50.0
    free_n = free_n_gate(cut.stype, cut.true_flux, np.broadcast_to(var, cut.true_flux.shape))
    acc = island_accuracy(psf_of, cut.n_bands, int(cut.psfs[0].n_terms),
                          int(cut.psfs[0].H_psf), cut.dstack, cut.wstack, cut.pad,
                          cut.island_set, island_ids, (cut.x, cut.y, cut.stype, cut.logre,
                                                       cut.ee1, cut.ee2, cut.sersic_n),
                          free_n, cfg, grams=grams, sync=sync, joint_nsrc_cap=None,
                          flux_scale=cutout_flux_scale(cut))
    if tuple(grams) == ("fp32",):
        acc.update(acc.pop("fp32"))   # flat layout for the single-precision comparison
    return acc
