"""Island-architecture forced-photometry fit of one field on one GPU.

The single-GPU restatement of ``run_grid_gpu_fit`` in island mode
(``tract7dt/fit/grid_gpu/pipeline.py:1352-2725``) and
``run_island_prestage`` (``island_stage.py:1338-3363``):

  pass 1  per-grid linear S1 solve at the catalog seeds (fluxes + plane)
  islands flux-aware coupling radii (or a fixed linking radius), connected
          components, (B, N) buckets, star/mixed classes and tiers
  stacks  padded field stacks: data - per-grid S1 plane on each grid core
  fit     per-class persistent VarPro-LM kernel
  pass 2  per-grid linear S1' solve at the fitted positions and shapes

Knobs the reference turns on by default that this build leaves off (and
says so in the run record): scene subtraction, blank-response correction,
wing peeling, post-fit error/metric passes, catalog priors.
Inputs stay resident on the device for the whole fit.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import forced_phot_sparse as fps
from .grid import assign_sources_to_grids, divide_image_into_grids
from .island_plan import (DEFAULT_MAX_STEPS, class_host_arrays, free_n_gate,
                          plan_island_classes)
from .island_solver import make_class_kernels, psf_tree_device, run_class_chunk
from .islands import (ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES, bucket_islands,
                      build_islands, flux_aware_radii, psf_radial_profile)
from .psf_polynomial import evaluate_psf_at_np, psf_device_arrays
from .source_model import SOURCE_STAR


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
    stamp_tiers: tuple = (55, 111, 221)
    island_split: bool = True    # band-split island LM (uph_island_fit_split)
    galaxy_prerender: bool = True  # split path: galaxies rendered by their own launches


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


def linear_pass(fld: DeviceField, grids, x, y, stype, logre, ee1, ee2, sn, cfg: FitConfig,
                gtier: dict, timer=None):
    """Per-grid, per-band linear solve; returns fluxes/var (n, nb) for each
    source from the grid that owns it, and bg coefs (n_grid, nb, M).  The
    grids, their fit sets and stamp tiers are those of pass 1 (pass 2 only
    replaces the source parameters, island_stage.py:3331-3363)."""
    n = x.size
    nb = fld.nb
    M = (cfg.bg_order + 1) * (cfg.bg_order + 2) // 2
    flux = np.full((n, nb), np.nan, np.float32)
    var = np.full((n, nb), np.nan, np.float32)
    bg = np.zeros((len(grids), nb, M), np.float32)
    bg_geo = (0.5 * (fld.W - 1), max(0.5 * fld.W, 1.0), 0.5 * (fld.H - 1), max(0.5 * fld.H, 1.0))
    by_tier = {}
    for g in grids:
        by_tier.setdefault(gtier[g.grid_id], []).append(g)
    for T, gl in sorted(by_tier.items()):
        problems, owners = [], []
        for g in gl:
            fi = g.fit_indices
            nbr, _ = fps.build_neighbor_list(x[fi], y[fi], T, T, cfg.k_max)
            for b in range(nb):
                problems.append(fps.BandProblem(
                    band=b, x0=g.x0_pad, y0=g.y0_pad, h=g.pad_h, w=g.pad_w,
                    pos_x=x[fi].astype(np.float32), pos_y=y[fi].astype(np.float32),
                    nbr_idx=nbr, stype=stype[fi], logre=logre[fi], ee1=ee1[fi], ee2=ee2[fi],
                    sersic_n=sn[fi]))
                owners.append((g, b))
        coefs, geo = fld.psf(T)
        res = fps.solve_problems(problems, coefs, geo, T, fld.n_terms, fld.data, fld.invvar,
                                 fld.good, bg_order=cfg.bg_order, bg_geo=bg_geo, reg=cfg.reg,
                                 k_max=cfg.k_max, timer=timer)
        for (g, b), r in zip(owners, res):
            own = g.source_indices.size
            flux[g.source_indices, b] = r.fluxes[:own]
            var[g.source_indices, b] = r.flux_var[:own]
            bg[g.grid_id, b] = r.bg_coefs
    return flux, var, bg


def padded_stacks(fld: DeviceField, grids, bg, bg_order, pad):
    """(dstack, wstack) (nb, H+2 pad, W+2 pad): data minus each grid's S1
    background on its core, inverse variance; zero border."""
    import torch
    nb, H, W = fld.nb, fld.H, fld.W
    dev = fld.data.device
    d = torch.zeros((nb, H + 2 * pad, W + 2 * pad), dtype=torch.float32, device=dev)
    w = torch.zeros_like(d)
    d[:, pad:pad + H, pad:pad + W] = fld.data
    w[:, pad:pad + H, pad:pad + W] = fld.invvar
    xo, xs = 0.5 * (W - 1), max(0.5 * W, 1.0)
    yo, ys = 0.5 * (H - 1), max(0.5 * H, 1.0)
    from .psf_polynomial import terms_for_order
    terms = terms_for_order(bg_order)
    for g in grids:
        h, ww = g.y1_core - g.y0_core, g.x1_core - g.x0_core
        X = ((np.arange(ww) + g.x0_core - xo) / xs)[None, :]
        Y = ((np.arange(h) + g.y0_core - yo) / ys)[:, None]
        basis = np.stack([np.broadcast_to(X ** m * Y ** n, (h, ww)).astype(np.float32)
                          for (m, n) in terms], axis=-1)
        plane = np.einsum("hwm,bm->bhw", basis, bg[g.grid_id]).astype(np.float32)
        d[:, pad + g.y0_core:pad + g.y1_core, pad + g.x0_core:pad + g.x1_core] -= \
            torch.from_numpy(plane).to(dev)
    return d, w


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
    sync()
    tm["islands"] = time.perf_counter() - t1
    # ---- island LM ----
    t2 = time.perf_counter()
    fx, fy = x.copy(), y.copy()
    flr, fe1, fe2, fsn = logre.copy(), ee1.copy(), ee2.copy(), sn.copy()
    conv = np.zeros(n, bool)
    stall = np.zeros(n, bool)
    nst = np.zeros(n, np.int32)
    giant = np.zeros(n, bool)
    for gid in giants:
        giant[iset.members[int(gid)]] = True
    n_fitted_islands = 0
    classes = []
    batch = []
    trees = {}
    from .island_solver import PsfTree, run_island_batch
    for plan in plans:
        host = class_host_arrays(plan, pos_x=x, pos_y=y, source_type=stype, logre=logre,
                                 ee1=ee1, ee2=ee2, sersic_n=sn, free_n_all=free_n,
                                 pos_max_shift_pix=cfg.pos_max_shift_pix, pad=pad)
        kern = make_class_kernels(plan, fld.nb, fld.n_terms, reg=cfg.reg,
                                  precision=cfg.precision)
        if plan.tier not in trees:
            coefs, geo = fld.psf(plan.tier)
            trees[plan.tier] = PsfTree(coefs, geo, plan.tier, fld.n_terms)
        batch.append((kern, host, np.arange(plan.n_islands)))
        classes.append((plan, host))
    sersic_dev = None
    if any(p.kind != "star" for p in plans):
        from .sersic_table import device_table
        sersic_dev = device_table(dstack.device)
    results = run_island_batch(batch, trees, cfg.max_steps, dstack, wstack,
                               precision=cfg.precision, timer=timer, sersic_dev=sersic_dev,
                               split=cfg.island_split, prerender=cfg.galaxy_prerender)
    for (plan, host), res in zip(classes, results):
        n_fitted_islands += plan.n_islands
        mm = plan.mmask
        src = plan.midx[mm]
        fx[src] = res["p"][:, :, 0][mm]
        fy[src] = res["p"][:, :, 1][mm]
        ext = (stype[src] != SOURCE_STAR)
        flr[src[ext]] = res["p"][:, :, 2][mm][ext]
        fe1[src[ext]] = res["p"][:, :, 3][mm][ext]
        fe2[src[ext]] = res["p"][:, :, 4][mm][ext]
        rows = np.repeat(np.arange(plan.n_islands), mm.sum(axis=1))
        conv[src] = res["conv"][rows]
        stall[src] = res["stall"][rows]
        nst[src] = res["nst"][rows]
    sync()
    tm["island_fit"] = time.perf_counter() - t2
    # ---- pass 2 ----
    t3 = time.perf_counter()
    flux, var, _ = linear_pass(fld, grids, fx, fy, stype, flr, fe1, fe2, fsn, cfg, gtier, timer)
    sync()
    tm["pass2"] = time.perf_counter() - t3
    tm["total"] = time.perf_counter() - t0
    rec = dict(n_sources=n, n_islands=int(iset.n_islands), n_fitted_islands=n_fitted_islands,
               n_giant_islands=int(giants.size), classes=[
                   (p.bbox_class, p.nsrc_class, p.kind, p.tier, p.n_islands) for p in plans],
               knobs_off=["scene_subtract", "blank_response_correction", "wing_peeling",
                          "param_errors", "source_metrics", "marginal_flux_err", "priors"])
    if keep_artifacts:
        rec["_artifacts"] = dict(classes=classes, dstack=dstack, wstack=wstack, pad=pad,
                                 grids=grids, island_set=iset, gtier=gtier)
    return FitResult(flux, np.sqrt(np.maximum(var, 0.0)), s1_flux, s1_var, fx, fy, flr, fe1,
                     fe2, fsn, iset.labels.astype(np.int64), conv, stall, nst, giant, tm, rec)
