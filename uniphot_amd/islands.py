"""Island decomposition (host, NumPy/SciPy).

Follows ``tract7dt/fit/grid_gpu/islands.py``: coupling radii (:68-103),
PSF radial profile (:128-157), flux-aware radii (:234-302), connected
components of the coupling graph with labels ordered by minimum member
index and half-open integer bboxes (:423-550), and (bbox, nsrc) class
bucketing (:614-663).  Island classes used by the island stage:
``island_stage.py:133-134``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy import sparse
from scipy.sparse.csgraph import connected_components
from scipy.spatial import cKDTree

from .source_model import SOURCE_STAR

RE_MIN, RE_MAX = 0.3, 200.0
SERSIC_B1 = 1.6783469900166605
ISLAND_BBOX_CLASSES = (32, 48, 64, 96, 128, 192, 256, 384)
ISLAND_NSRC_CLASSES = (1, 2, 4, 8, 16, 32, 64)


def nre_multiplier(sersic_n) -> np.ndarray:
    n = np.asarray(sersic_n, np.float64)
    n = np.where(np.isfinite(n), n, 1.0)
    return np.where(n < 1.5, 4.0, np.where(n < 3.0, 6.0, 8.0))


def coupling_radii(source_type, logre, sersic_n, r_psf: float) -> np.ndarray:
    """Stars r_psf; extended nre(n) * clip(Re) + r_psf."""
    if r_psf <= 0:
        raise ValueError(f"r_psf must be positive, got {r_psf}")
    re = np.exp(np.asarray(logre, np.float64))
    re = np.clip(np.where(np.isfinite(re), re, RE_MIN), RE_MIN, RE_MAX)
    ext = nre_multiplier(sersic_n) * re + float(r_psf)
    return np.where(np.asarray(source_type) == SOURCE_STAR, float(r_psf), ext)


def psf_radial_profile(stamp, bin_width: float = 0.5):
    """Monotone (running-min) annulus-mean unit-flux profile of a centred
    stamp; returns (r, p) with r[0] = 0 anchored at the central value."""
    st = np.asarray(stamp, np.float64)
    tot = st.sum()
    if not np.isfinite(tot) or tot <= 0:
        raise ValueError("psf_radial_profile: stamp must have positive sum")
    st = st / tot
    H, W = st.shape
    yy, xx = np.mgrid[0:H, 0:W]
    rr = np.hypot(yy - (H - 1) / 2.0, xx - (W - 1) / 2.0).ravel()
    nbin = int(np.ceil(rr.max() / bin_width)) + 1
    idx = np.minimum((rr / bin_width).astype(np.int64), nbin - 1)
    sums = np.bincount(idx, weights=st.ravel(), minlength=nbin)
    cnts = np.bincount(idx, minlength=nbin)
    ok = cnts > 0
    r = (np.nonzero(ok)[0] + 0.5) * bin_width
    p = np.minimum.accumulate(np.maximum(sums[ok] / cnts[ok], 1e-12))
    return np.concatenate([[0.0], r]), np.concatenate([[p[0]], p])


def _invert_profile(psf_r, psf_p, thresh):
    lp = np.log(psf_p[::-1])
    rr = psf_r[::-1]
    t = np.asarray(thresh, np.float64)
    out = np.full(t.shape, psf_r[-1])
    pos = np.isfinite(t) & (t > 0)
    tc = np.clip(np.log(np.where(pos, t, 1.0)), lp[0], lp[-1])
    out[pos] = np.interp(tc, lp, rr)[pos]
    out[np.isfinite(t) & (t >= psf_p[0])] = 0.0
    out[np.isinf(t) & (t > 0)] = 0.0
    return out


def _ext_shape(logre, sersic_n):
    from scipy.special import gamma
    re = np.clip(np.where(np.isfinite(logre), np.exp(logre), RE_MIN), RE_MIN, RE_MAX)
    n = np.where(np.isfinite(sersic_n), sersic_n, 1.0)
    n_eff = np.where(n < 1.5, 1.0, n)
    b = np.where(n_eff < 1.5, SERSIC_B1, 2.0 * n_eff - 1.0 / 3.0)
    i0 = b ** (2.0 * n_eff) / (2.0 * np.pi * n_eff * gamma(2.0 * n_eff) * re ** 2)
    return re, n_eff, b, i0


def flux_aware_radii(source_type, logre, sersic_n, flux, sigma_sky: float, kappa: float,
                     psf_r, psf_p, r_psf: float, r_core: float | None = None) -> np.ndarray:
    """Radius where flux * unit profile = kappa * sigma_sky, floored at
    r_core (default PSF FWHM) and capped at the geometric radius."""
    if sigma_sky <= 0 or not np.isfinite(sigma_sky):
        raise ValueError(f"sigma_sky must be positive, got {sigma_sky}")
    if kappa <= 0:
        raise ValueError(f"kappa must be positive, got {kappa}")
    st = np.asarray(source_type)
    logre = np.asarray(logre, np.float64)
    sersic_n = np.asarray(sersic_n, np.float64)
    flux = np.asarray(flux, np.float64)
    if r_core is None:
        r_core = 2.0 * float(_invert_profile(psf_r, psf_p, np.array([0.5 * psf_p[0]]))[0])
    if r_core <= 0:
        raise ValueError(f"r_core must be positive, got {r_core}")
    good = np.isfinite(flux) & (flux > 0)
    thresh = np.where(good, kappa * float(sigma_sky) / np.where(good, flux, 1.0), np.inf)
    r_star = _invert_profile(psf_r, psf_p, thresh)
    re, n_eff, b, i0 = _ext_shape(logre, sersic_n)
    with np.errstate(divide="ignore", invalid="ignore"):
        arg = np.log(i0 / thresh)
    r_ext = np.where(arg > 0, re * (np.maximum(arg, 0.0) / b) ** n_eff, 0.0)
    cap = coupling_radii(st, logre, sersic_n, r_psf)
    return np.clip(np.where(st == SOURCE_STAR, r_star, r_ext), r_core, cap)


@dataclass
class IslandSet:
    n_sources: int
    n_islands: int
    labels: np.ndarray
    radius: np.ndarray
    members: list = field(repr=False)
    n_src: np.ndarray = field(repr=False)
    bbox: np.ndarray = field(repr=False)   # (n_isl, 4) x0, y0, x1, y1 half-open
    r_psf: float = 0.0

    @property
    def bbox_size(self) -> np.ndarray:
        return np.maximum(self.bbox[:, 2] - self.bbox[:, 0], self.bbox[:, 3] - self.bbox[:, 1])


def coupled_pairs(px, py, radius):
    """All i < j with dist(i, j) < r_i + r_j (kd-tree candidates at the
    largest possible separation, exact predicate afterwards)."""
    n = px.size
    if n < 2:
        return np.empty(0, np.int64), np.empty(0, np.int64)
    tree = cKDTree(np.column_stack([px, py]))
    cand = tree.query_pairs(r=2.0 * float(radius.max()), output_type="ndarray")
    if cand.size == 0:
        return np.empty(0, np.int64), np.empty(0, np.int64)
    i, j = cand[:, 0].astype(np.int64), cand[:, 1].astype(np.int64)
    d2 = (px[i] - px[j]) ** 2 + (py[i] - py[j]) ** 2
    rs = radius[i] + radius[j]
    keep = d2 < rs * rs
    return i[keep], j[keep]


def build_islands(pos_x, pos_y, source_type, logre, sersic_n, r_psf: float, *,
                  radius=None) -> IslandSet:
    px = np.asarray(pos_x, np.float64)
    py = np.asarray(pos_y, np.float64)
    n = px.size
    if not (np.all(np.isfinite(px)) and np.all(np.isfinite(py))):
        raise ValueError("build_islands requires finite positions")
    if source_type is None:
        source_type = np.zeros(n, np.int32)
    if logre is None:
        logre = np.zeros(n)
    if sersic_n is None:
        sersic_n = np.ones(n)
    if radius is None:
        radius = coupling_radii(source_type, logre, sersic_n, r_psf)
    radius = np.asarray(radius, np.float64)
    if n and not (np.all(np.isfinite(radius)) and np.all(radius > 0)):
        raise ValueError("radius must be finite and positive")
    if n == 0:
        return IslandSet(0, 0, np.empty(0, np.int32), radius, [], np.empty(0, np.int32),
                         np.empty((0, 4), np.int64), float(r_psf))
    i, j = coupled_pairs(px, py, radius)
    if i.size:
        g = sparse.coo_matrix((np.ones(i.size, np.int8), (i, j)), shape=(n, n))
        _, raw = connected_components(g, directed=False)
    else:
        raw = np.arange(n)
    n_isl = int(raw.max()) + 1
    first = np.full(n_isl, n, np.int64)
    np.minimum.at(first, raw, np.arange(n))
    remap = np.empty(n_isl, np.int64)
    remap[np.argsort(first, kind="stable")] = np.arange(n_isl)
    labels = remap[raw].astype(np.int32)
    order = np.argsort(labels, kind="stable")
    counts = np.bincount(labels, minlength=n_isl).astype(np.int32)
    starts = np.concatenate([[0], np.cumsum(counts[:-1])]).astype(np.int64)
    members = [order[s:s + c].astype(np.int32) for s, c in zip(starts, counts)]
    bbox = np.empty((n_isl, 4), np.int64)
    bbox[:, 0] = np.floor(np.minimum.reduceat((px - radius)[order], starts))
    bbox[:, 1] = np.floor(np.minimum.reduceat((py - radius)[order], starts))
    bbox[:, 2] = np.ceil(np.maximum.reduceat((px + radius)[order], starts))
    bbox[:, 3] = np.ceil(np.maximum.reduceat((py + radius)[order], starts))
    return IslandSet(n, n_isl, labels, radius, members, counts, bbox, float(r_psf))


@dataclass
class IslandBucket:
    bbox_class: int
    nsrc_class: int
    island_ids: np.ndarray
    member_idx: np.ndarray = field(repr=False)   # (n_b, N) int32, -1 padding
    member_mask: np.ndarray = field(repr=False)  # (n_b, N) bool
    bbox: np.ndarray = field(repr=False)

    @property
    def n_islands(self) -> int:
        return int(self.island_ids.size)


@dataclass
class BucketedIslands:
    bbox_classes: tuple
    nsrc_classes: tuple
    buckets: dict
    giant_ids: np.ndarray


def bucket_islands(isl: IslandSet, bbox_classes=ISLAND_BBOX_CLASSES,
                   nsrc_classes=ISLAND_NSRC_CLASSES) -> BucketedIslands:
    """Smallest (bbox, nsrc) class holding each island; oversize -> giant."""
    bc = tuple(sorted(int(c) for c in bbox_classes))
    nc = tuple(sorted(int(c) for c in nsrc_classes))
    bi = np.searchsorted(np.asarray(bc), isl.bbox_size, side="left")
    ni = np.searchsorted(np.asarray(nc), isl.n_src, side="left")
    giant = (bi >= len(bc)) | (ni >= len(nc))
    buckets = {}
    ok = np.where(~giant)[0]
    if ok.size:
        keys = bi[ok] * len(nc) + ni[ok]
        for key in np.unique(keys):
            sel = ok[keys == key]
            B, N = bc[int(key) // len(nc)], nc[int(key) % len(nc)]
            midx = np.full((sel.size, N), -1, np.int32)
            mmask = np.zeros((sel.size, N), bool)
            for row, k in enumerate(sel):
                m = isl.members[k]
                midx[row, :m.size] = m
                mmask[row, :m.size] = True
            buckets[(B, N)] = IslandBucket(B, N, sel.astype(np.int32), midx, mmask,
                                           isl.bbox[sel])
    return BucketedIslands(bc, nc, buckets, np.where(giant)[0].astype(np.int32))
