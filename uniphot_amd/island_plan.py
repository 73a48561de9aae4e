"""Island class planning and per-class host arrays (host, NumPy).

Follows ``tract7dt/fit/grid_gpu/island_plan.py``: LM defaults (:55-74),
``plan_island_classes`` (:356-490: star vs mixed kind, member stamp tier
from the {native, 111, 221} ladder covering 2 * nre(n) * Re, joint nsrc cap
-> giants) and ``class_host_arrays`` (:493-699: centred window origins,
seeds, keep masks, box bounds), plus ``free_n_gate`` (:199-217) and
``giant_seed_rows`` (:1019-1088).  Catalog priors are not part of this
build (every prior inverse sigma is zero, the reference's exact no-op).
"""
from __future__ import annotations

from dataclasses import dataclass

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


def free_n_gate(source_type, s1_flux, s1_var, fit_sersic_n: bool = True,
                min_snr: float = 30.0) -> np.ndarray:
    """SERSIC rows whose median-band S1 S/N reaches ``min_snr`` fit n."""
    st = np.asarray(source_type)
    if not fit_sersic_n:
        return np.zeros(st.size, bool)
    f = np.asarray(s1_flux, np.float64)
    v = np.asarray(s1_var, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        snr = np.nanmedian(f / np.sqrt(np.maximum(v, 1e-300)), axis=1) if f.ndim == 2 \
            else f / np.sqrt(np.maximum(v, 1e-300))
    return (st == SOURCE_SERSIC) & np.isfinite(snr) & (snr >= float(min_snr))


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
    u = sersic_n_to_u(sn.astype(np.float64)).astype(np.float32)
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
    return dict(ox=ox.astype(np.int64), oy=oy.astype(np.int64),
                x0b=(ox - pad).astype(np.float32), y0b=(oy - pad).astype(np.float32),
                st=st, lr=lr, e1=e1, e2=e2, sn=sn, fn_flag=fn, is_ext=is_ext, mmask=mm,
                keep=keep, p0=p0, lower=lower, upper=upper, prior_mean=p0.copy(),
                area=(wb * hb).astype(np.int64),
                prior_inv_sigma=np.zeros_like(p0), has_priors=False)


def giant_seed_rows(giant_ids, island_set, pos_x, pos_y, logre, ee1, ee2, sersic_n):
    """Giant islands are not fitted: members keep their seed geometry
    (island_plan.py:1019-1088)."""
    rows = [m for g in giant_ids for m in island_set.members[int(g)]]
    idx = np.asarray(rows, np.int64)
    return dict(src=idx, x=pos_x[idx], y=pos_y[idx], logre=logre[idx], ee1=ee1[idx],
                ee2=ee2[idx], sersic_n=sersic_n[idx])
