"""Island class planning and per-class host arrays (host, NumPy).

Follows ``tract7dt/fit/grid_gpu/island_plan.py``: LM defaults (:55-74),
``plan_island_classes`` (:356-490: star vs mixed kind, member stamp tier
from the {native, 111, 221} ladder covering 2 * nre(n) * Re, joint nsrc cap
-> giants) and ``class_host_arrays`` (:493-699: centred window origins,
seeds, keep masks, box bounds), plus ``free_n_gate`` (:199-217) and
``giant_seed_rows`` (:1019-1088) and ``unpack_class_results``
(:702-1016).  Catalog priors are not part of this build (every prior
inverse sigma is zero, the reference's exact no-op).
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

from .source_model import SERSIC_N_MAX, SERSIC_N_MIN, SOURCE_SERSIC, SOURCE_STAR, sersic_n_to_u

DEFAULT_MAX_STEPS = 30
DEFAULT_MAX_STEPS_MAIN = 8
DEFAULT_CHI2_FTOL_REL = 1e-4
DEFAULT_STEP_TOL_PIX = 0.1
DEFAULT_POS_STEP_MAX = 0.5
DEFAULT_SHAPE_STEP_MAX = 0.15
DEFAULT_UN_STEP_MAX = 0.5
DEFAULT_SHAPE_STEP_TOL = 0.01
DEFAULT_REG = 1e-9
JOINT_NSRC_CAP = 32


@dataclass
class IslandFitResult:
    """Per-source results of every source in a fitted or giant island
    (island_plan.py:220-319), rows aligned with ``source_idx``; island-level
    quantities replicated onto each member."""
    bandnames: list
    source_idx: np.ndarray
    pos_x: np.ndarray
    pos_y: np.ndarray
    logre: np.ndarray
    ee1: np.ndarray
    ee2: np.ndarray
    sersic_n: np.ndarray
    fluxes: np.ndarray              # (n_fit, n_bands)
    flux_err: np.ndarray
    chi2: np.ndarray
    n_steps: np.ndarray
    converged: np.ndarray
    free_n: np.ndarray
    stalled: np.ndarray = None
    diverged: np.ndarray = None
    stall_dchi2: np.ndarray = None
    railed: np.ndarray = None
    chi2_nu: np.ndarray = None
    island_id: np.ndarray = None
    island_nsrc: np.ndarray = None
    island_giant: np.ndarray = None
    re_upper_pix: np.ndarray = None
    pos_x_err: np.ndarray = None
    pos_y_err: np.ndarray = None
    re_err: np.ndarray = None
    re_prof_lo: np.ndarray = None
    re_prof_hi: np.ndarray = None
    ell_err: np.ndarray = None
    theta_err: np.ndarray = None
    sersic_n_err: np.ndarray = None
    pos_x_prior_frac: np.ndarray = None
    pos_y_prior_frac: np.ndarray = None
    re_prior_frac: np.ndarray = None
    ell_prior_frac: np.ndarray = None
    theta_prior_frac: np.ndarray = None
    sersic_n_prior_frac: np.ndarray = None
    rchisq: np.ndarray = None
    fracflux: np.ndarray = None
    fracin: np.ndarray = None
    fracmasked: np.ndarray = None
    flux_win: np.ndarray = None
    flux_marg_var: np.ndarray = None
    bucket_stats: list = field(default_factory=list)

    @property
    def requeue_idx(self) -> np.ndarray:
        return self.source_idx[~self.converged]

    @classmethod
    def from_rows(cls, bandnames, rows: list) -> "IslandFitResult":
        """Concatenate unpacked / giant row dicts (source order within each)."""
        keys = [f.name for f in dataclasses.fields(cls)
                if f.name not in ("bandnames", "source_idx", "n_steps", "converged",
                                  "stalled", "bucket_stats")]
        cat = lambda k: np.concatenate([r[k] for r in rows]) if rows else np.empty(0)  # noqa
        return cls(bandnames=list(bandnames), source_idx=cat("src").astype(np.int64),
                   n_steps=cat("n_steps").astype(np.int32),
                   converged=cat("converged").astype(bool),
                   stalled=cat("stalled").astype(bool),
                   **{k: cat(k) for k in keys if not rows or k in rows[0]})


@dataclass
class IslandClassPlan:
    class_idx: int
    bbox_class: int
    nsrc_class: int
    kind: str          # "star" | "mixed"
    tier: int
    isl_ids: np.ndarray
    midx: np.ndarray   # (n_isl, N) member source indices (padding -> 0)
    mmask: np.ndarray  # (n_isl, N)
    bbox: np.ndarray   # (n_isl, 4)
    n_src_of: np.ndarray
    n_gal_class: int

    @property
    def n_islands(self) -> int:
        return int(self.isl_ids.size)


def free_n_gate(source_type, s1_flux, s1_flux_var, fit_sersic_n: bool = True,
                n_fit_min_snr: float = 30.0) -> np.ndarray:
    """SERSIC rows whose median-band S1 S/N (|flux| / sqrt(var), the
    variance floored at 1e-30) reaches ``n_fit_min_snr`` fit their n
    (island_plan.py:199-217)."""
    st = np.asarray(source_type)
    out = np.zeros(st.size, bool)
    if not fit_sersic_n or not np.any(st == SOURCE_SERSIC):
        return out
    f = np.abs(np.asarray(s1_flux, np.float64))
    v = np.maximum(np.asarray(s1_flux_var, np.float64), 1e-30)
    with np.errstate(divide="ignore", invalid="ignore"):
        snr = np.nanmedian(f / np.sqrt(v), axis=-1) if f.ndim == 2 else f / np.sqrt(v)
    return (st == SOURCE_SERSIC) & np.isfinite(snr) & (snr >= float(n_fit_min_snr))


def plan_island_classes(island_set, buckets, source_type, logre, sersic_n, native: int,
                        *, joint_nsrc_cap: int | None = JOINT_NSRC_CAP):
    """Split occupied buckets into (B, N, kind, tier) classes; islands with
    more members than ``joint_nsrc_cap`` go to the giant list."""
    todo = dict(buckets.buckets)
    capped = []
    if joint_nsrc_cap is not None:
        for k in [k for k in todo if k[1] > int(joint_nsrc_cap)]:
            capped.extend(todo.pop(k).island_ids.tolist())
    plans = []
    for (bc, nc) in sorted(todo):
        b = todo[(bc, nc)]
        midx = np.maximum(b.member_idx.astype(np.int64), 0)
        mm = b.member_mask.astype(bool)
        is_gal = (source_type[midx] != SOURCE_STAR) & mm
        has_ext = is_gal.any(axis=1)
        n_gal = is_gal.sum(axis=1)
        sn = sersic_n[midx]
        nre = np.where(sn < 1.5, 4.0, np.where(sn < 3.0, 6.0, 8.0))
        req = np.where(is_gal, nre * np.exp(logre[midx].astype(np.float64)), 0.0).max(axis=1)
        ladder = [t for t in (native, 111, 221) if t >= native]
        tier = np.full(b.n_islands, ladder[-1], np.int64)
        for t in reversed(ladder):
            tier = np.where(2 * req <= t, t, tier)
        groups = [("star", native, np.where(~has_ext)[0])]
        groups += [("mixed", int(t), np.where(has_ext & (tier == t))[0]) for t in ladder]
        for kind, T, sel in groups:
            if sel.size == 0:
                continue
            plans.append(IslandClassPlan(
                class_idx=len(plans), bbox_class=int(bc), nsrc_class=int(nc), kind=kind,
                tier=int(T), isl_ids=b.island_ids[sel].astype(np.int64), midx=midx[sel],
                mmask=mm[sel], bbox=island_set.bbox[b.island_ids[sel]],
                n_src_of=island_set.n_src[b.island_ids[sel]].astype(np.int32),
                n_gal_class=int(n_gal[sel].max()) if kind == "mixed" else 0))
    return plans, capped


def class_host_arrays(plan: IslandClassPlan, *, pos_x, pos_y, source_type, logre, ee1, ee2,
                      sersic_n, free_n_all, pos_max_shift_pix: float, pad: int) -> dict:
    """Per-class window origins (padded-stack and image frames), member
    arrays, keep masks, seeds and box bounds."""
    B, N, T = plan.bbox_class, plan.nsrc_class, plan.tier
    midx, mm = plan.midx, plan.mmask
    n_isl = plan.n_islands
    bbox = plan.bbox
    wb = bbox[:, 2] - bbox[:, 0]
    hb = bbox[:, 3] - bbox[:, 1]
    ox = bbox[:, 0] - (B - wb) // 2 + pad
    oy = bbox[:, 1] - (B - hb) // 2 + pad
    sx = pos_x[midx].astype(np.float32)
    sy = pos_y[midx].astype(np.float32)
    st = source_type[midx].astype(np.int32)
    lr = logre[midx].astype(np.float32)
    e1 = ee1[midx].astype(np.float32)
    e2 = ee2[midx].astype(np.float32)
    sn = sersic_n[midx].astype(np.float32)
    u = sersic_n_to_u(sn).astype(np.float32)   # in float32, like the reference
    is_ext = st != SOURCE_STAR
    fn = free_n_all[midx] & mm
    keep = np.zeros((n_isl, N, 6), bool)
    keep[:, :, 0:2] = mm[:, :, None]
    keep[:, :, 2:5] = (is_ext & mm)[:, :, None]
    keep[:, :, 5] = fn
    p0 = np.stack([sx, sy, lr, e1, e2, u], axis=2).astype(np.float32)
    one = np.ones((n_isl, N), np.float32)
    lower = np.stack([sx - pos_max_shift_pix, sy - pos_max_shift_pix,
                      np.full((n_isl, N), np.log(0.1), np.float32), -100 * one, -100 * one,
                      -12 * one], axis=2).astype(np.float32)
    upper = np.stack([sx + pos_max_shift_pix, sy + pos_max_shift_pix,
                      np.full((n_isl, N), np.log(0.75 * T), np.float32), 100 * one, 100 * one,
                      12 * one], axis=2).astype(np.float32)
    # galaxy gather map: per island the slots of its extended members first
    # (slot order kept), the first n_gal_class of them, and which are real
    gal = is_ext & mm
    G = int(plan.n_gal_class)
    first = np.argsort(~gal, axis=1, kind="stable")[:, :G]
    return dict(ox=ox.astype(np.int64), oy=oy.astype(np.int64),
                x0b=(ox - pad).astype(np.float32), y0b=(oy - pad).astype(np.float32),
                st=st, lr=lr, e1=e1, e2=e2, sn=sn, fn_flag=fn, is_ext=is_ext, mmask=mm,
                keep=keep, gsel=first.astype(np.int32),
                gvalid=np.take_along_axis(gal, first, axis=1),
                p0=p0, lower=lower, upper=upper, prior_mean=p0.copy(),
                area=(wb * hb).astype(np.int64),
                prior_inv_sigma=np.zeros_like(p0), has_priors=False, member_s1_flux=None)


def giant_seed_rows(giant_ids, island_set, *, pos_x, pos_y, logre, ee1, ee2, sersic_n,
                    n_bands: int) -> list:
    """Source rows of giant islands, which are not fitted: members keep their
    seed geometry, fitted quantities are NaN, flags say converged and giant;
    one dict per island in ascending island id, with the keys of
    :func:`unpack_class_results` (island_plan.py:1019-1088)."""
    rows = []
    for g in sorted(int(i) for i in np.ravel(giant_ids)):
        src = np.asarray(island_set.members[g], np.int64)
        m = src.size
        nan1 = lambda: np.full(m, np.nan, np.float32)                      # noqa: E731
        nanb = lambda: np.full((m, max(int(n_bands), 0)), np.nan, np.float32)  # noqa: E731
        geo = {k: np.asarray(v)[src].astype(np.float32) for k, v in (
            ("pos_x", pos_x), ("pos_y", pos_y), ("logre", logre), ("ee1", ee1), ("ee2", ee2),
            ("sersic_n", sersic_n))}
        row = dict(src=src, **geo, fluxes=nanb(), flux_err=nanb(), chi2=nan1(), chi2_nu=nan1(),
                   n_steps=np.zeros(m, np.int32), converged=np.ones(m, bool),
                   stalled=np.zeros(m, bool), diverged=np.zeros(m, bool), stall_dchi2=nan1(),
                   railed=np.zeros(m, bool), free_n=np.zeros(m, bool),
                   island_id=np.full(m, g, np.int32), island_nsrc=np.full(m, m, np.int32),
                   island_giant=np.ones(m, bool), re_upper_pix=nan1(),
                   flux_win=nanb(), flux_marg_var=nanb())
        for k in _NAN_ROW_KEYS:
            row[k] = nan1()
        rows.append(row)
    return rows


# per-source error / disclosure keys that are NaN without the post-fit pass
_NAN_ROW_KEYS = ("pos_x_err", "pos_y_err", "re_err", "ell_err", "theta_err", "sersic_n_err",
                 "re_prof_lo", "re_prof_hi", "pos_x_prior_frac", "pos_y_prior_frac",
                 "re_prior_frac", "ell_prior_frac", "theta_prior_frac", "sersic_n_prior_frac",
                 "rchisq", "fracflux", "fracin", "fracmasked")


def _catalog_errors(p, var6, cov, keep, fn_flag, n_final):
    """Catalog-frame 1-sigma errors of fitted members from the marginal
    variances of (x, y, logre, ee1, ee2, u) and cov(ee1, ee2): x and y
    directly; Re = exp(logre) scales by Re; the axis ratio's ELL =
    2e / (1 + e) with e = 1 - exp(-|ee|) and the position angle
    THETA = -atan2(ee2, ee1) / 2 (degrees) through their gradients in
    (ee1, ee2) against the 2x2 covariance; Sersic n through the logit map's
    dn/du (free-n members only).  Variances that are not positive and
    finite, or parameters that were not fitted, give NaN."""
    usable = keep & np.isfinite(var6) & (var6 > 0)
    sig = np.where(usable, np.sqrt(np.where(usable, var6, 1.0)), np.nan)
    re_err = np.exp(p[:, 2].astype(np.float64)) * sig[:, 2]
    a, b = p[:, 3].astype(np.float64), p[:, 4].astype(np.float64)
    r = np.hypot(a, b)
    pair = usable[:, 3] & usable[:, 4] & np.isfinite(cov) & (r > 0)
    rr = np.where(r > 0, r, 1.0)
    radial = (a * a * var6[:, 3] + 2.0 * a * b * cov + b * b * var6[:, 4]) / (rr * rr)
    tangential = (b * b * var6[:, 3] - 2.0 * a * b * cov + a * a * var6[:, 4]) / rr ** 4
    t = np.exp(-r)
    v_ell = (2.0 * t / (2.0 - t) ** 2) ** 2 * radial
    v_th = (90.0 / np.pi) ** 2 * tangential

    def root(v):
        ok = pair & np.isfinite(v) & (v > 0)
        return np.where(ok, np.sqrt(np.where(ok, v, 1.0)), np.nan).astype(np.float32)

    n = n_final.astype(np.float64)
    dndu = (n - SERSIC_N_MIN) * (SERSIC_N_MAX - n) / (SERSIC_N_MAX - SERSIC_N_MIN)
    n_err = np.where(fn_flag, sig[:, 5] * np.abs(dndu), np.nan).astype(np.float32)
    return (sig[:, 0].astype(np.float32), sig[:, 1].astype(np.float32),
            re_err.astype(np.float32), root(v_ell), root(v_th), n_err)


def unpack_class_results(plan: IslandClassPlan, host: dict, island_set, res_p, res_f, res_fe,
                         res_chi2, res_conv, res_stall, res_nst, res_perr=None,
                         res_pdchi2=None, res_reprof=None, reprof_deltas=None,
                         res_smetrics=None, res_swin=None, res_femarg=None, res_div=None,
                         res_railed=None, res_ngood=None, n_bands: int = 0) -> dict:
    """Island rows of one class -> per-source rows of its real members
    (island_plan.py:702-1016): fitted geometry (stars keep their seed shape,
    fixed-n members their seed n), per-band fluxes and errors, per-island
    chi2 and its reduced value chi2 / (good pixels - fitted nonlinear
    parameters - n_bands x members), LM flags, catalog-frame errors from
    ``res_perr`` (n_isl, N, >= 7) and the per-source metrics.  Catalog
    priors are not part of this build, so every prior fraction is NaN;
    ``res_reprof`` (the Re profile scan, off by default in the reference) is
    not produced and its bounds are NaN."""
    mm = np.asarray(host["mmask"], bool)
    isl_row, slot = np.nonzero(mm)
    n_mem = isl_row.size
    p = np.asarray(res_p)[isl_row, slot]                                  # (n_mem, 6)
    ext = host["is_ext"][isl_row, slot]
    fn = host["fn_flag"][isl_row, slot]
    n_fit = SERSIC_N_MIN + (SERSIC_N_MAX - SERSIC_N_MIN) / (1.0 + np.exp(-p[:, 5]))
    n_final = np.where(fn, n_fit, host["sn"][isl_row, slot]).astype(np.float32)
    if res_perr is None:
        errs = tuple(np.full(n_mem, np.nan, np.float32) for _ in range(6))
    else:
        pe = np.asarray(res_perr)[isl_row, slot]
        errs = _catalog_errors(p, pe[:, :6].astype(np.float64), pe[:, 6].astype(np.float64),
                               host["keep"][isl_row, slot], fn, n_final)
    if res_ngood is not None and int(n_bands) > 0:
        n_par = (host["keep"] & mm[:, :, None]).reshape(mm.shape[0], -1).sum(axis=1)
        dof = (np.asarray(res_ngood, np.float64) - n_par - float(n_bands) * mm.sum(axis=1))
        c2 = np.asarray(res_chi2, np.float64)
        with np.errstate(invalid="ignore", divide="ignore"):
            chi2_nu = np.where((dof > 0) & np.isfinite(c2), c2 / np.maximum(dof, 1.0), np.nan)
        chi2_nu = chi2_nu[isl_row].astype(np.float32)
    else:
        chi2_nu = np.full(n_mem, np.nan, np.float32)

    def per_source(a, fill_shape=None):
        if a is None:
            return np.full(fill_shape or n_mem, np.nan, np.float32)
        return np.asarray(a)[isl_row, slot]

    def per_band(a):
        if a is None:
            return np.full((n_mem, max(int(n_bands), 0)), np.nan, np.float32)
        return np.asarray(a)[isl_row, :, slot]

    met = None if res_smetrics is None else np.asarray(res_smetrics)[isl_row, slot]
    isl_ids = np.asarray(plan.isl_ids)
    out = dict(
        src=plan.midx[isl_row, slot],
        pos_x=p[:, 0], pos_y=p[:, 1],
        logre=np.where(ext, p[:, 2], host["lr"][isl_row, slot]),
        ee1=np.where(ext, p[:, 3], host["e1"][isl_row, slot]),
        ee2=np.where(ext, p[:, 4], host["e2"][isl_row, slot]),
        sersic_n=n_final, fluxes=per_band(res_f), flux_err=per_band(res_fe),
        chi2=np.asarray(res_chi2)[isl_row], chi2_nu=chi2_nu,
        n_steps=np.asarray(res_nst)[isl_row], converged=np.asarray(res_conv)[isl_row],
        stalled=np.asarray(res_stall)[isl_row],
        diverged=(np.zeros(n_mem, bool) if res_div is None else np.asarray(res_div)[isl_row]),
        stall_dchi2=(np.full(n_mem, np.nan, np.float32) if res_pdchi2 is None
                     else np.asarray(res_pdchi2)[isl_row].astype(np.float32)),
        railed=(np.zeros(n_mem, bool) if res_railed is None
                else np.asarray(res_railed)[isl_row]),
        free_n=fn, island_id=isl_ids[isl_row].astype(np.int32),
        island_nsrc=np.asarray(island_set.n_src)[isl_ids[isl_row]].astype(np.int32),
        island_giant=np.zeros(n_mem, bool),
        re_upper_pix=np.exp(host["upper"][isl_row, slot, 2].astype(np.float64)
                            ).astype(np.float32),
        flux_win=per_band(res_swin), flux_marg_var=per_band(res_femarg))
    for k, v in zip(("pos_x_err", "pos_y_err", "re_err", "ell_err", "theta_err",
                     "sersic_n_err"), errs):
        out[k] = v
    for q, k in enumerate(("rchisq", "fracflux", "fracin", "fracmasked")):
        out[k] = np.full(n_mem, np.nan, np.float32) if met is None else met[:, q]
    for k in _NAN_ROW_KEYS:
        out.setdefault(k, np.full(n_mem, np.nan, np.float32))
    return out
