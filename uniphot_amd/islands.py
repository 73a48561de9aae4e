"""Island decomposition: which catalogue sources are fitted jointly.

Two sources are coupled when their coupling disks overlap (distance below
the sum of their radii); islands are the connected groups of coupled
sources.  The results must equal the reference's bit for bit (membership
decides every fit), so the radius rule, island order and box convention
are the reference's (``tract7dt/fit/grid_gpu/islands.py``):
  * geometric radius (:68-103): r_psf for stars, nre(n) * clip(Re) + r_psf
    for extended sources;
  * flux-aware radius (:234-302): where flux * unit profile falls to
    kappa * sigma_sky -- the PSF's annulus profile (:128-179) for stars,
    an analytic Sersic falloff for extended sources -- floored at the PSF
    FWHM and capped at the geometric radius;
  * islands numbered by their smallest member index, members ascending,
    half-open integer boxes over the member disks (:423-550);
  * the smallest (box side, member count) class from the island stage's
    ladders (``island_stage.py:133-134``), larger islands "giant"
    (:614-663).
The search and grouping below are this module's own: sources split into a
small-radius population (pairs from one kd-tree pair query at twice the
population's largest radius) and the few large-radius sources (a ball query
each); components by min-root union-find with pointer jumping, whose roots
are the smallest member indices.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.spatial import cKDTree

from .source_model import SOURCE_STAR

RE_MIN, RE_MAX = 0.3, 200.0
SERSIC_B1 = 1.6783469900166605          # exact b for n = 1
# the island stage's class ladders (island_stage.py:133-134) and the
# bucketing function's own defaults (islands.py:59-60)
ISLAND_BBOX_CLASSES = (32, 48, 64, 96, 128, 192, 256, 384)
ISLAND_NSRC_CLASSES = (1, 2, 4, 8, 16, 32, 64)
DEFAULT_BBOX_CLASSES = (64, 96, 128, 192, 256, 384)
DEFAULT_NSRC_CLASSES = (1, 2, 4, 8, 16, 32, 64, 128)
# share of sources (by radius) searched by the pair query; the rest by ball
# queries of their own
_SMALL_RADIUS_QUANTILE = 0.95


def nre_multiplier(sersic_n) -> np.ndarray:
    """Radius multiple of Re that bounds an extended source's light: 4 below
    n = 1.5, 6 below n = 3, 8 above (non-finite n counts as 1)."""
    n = np.asarray(sersic_n, np.float64)
    n = np.where(np.isfinite(n), n, 1.0)
    return np.select([n < 1.5, n < 3.0], [4.0, 6.0], 8.0)


def _clean_re(logre) -> np.ndarray:
    re = np.exp(np.asarray(logre, np.float64))
    return np.clip(np.where(np.isfinite(re), re, RE_MIN), RE_MIN, RE_MAX)


def coupling_radii(source_type, logre, sersic_n, r_psf: float) -> np.ndarray:
    """Geometric coupling radius per source."""
    if not r_psf > 0:
        raise ValueError(f"r_psf must be positive, got {r_psf}")
    star = np.asarray(source_type) == SOURCE_STAR
    ext = nre_multiplier(sersic_n) * _clean_re(logre) + float(r_psf)
    return np.where(star, float(r_psf), ext)


@dataclass
class PsfProfile:
    """Unit-flux annulus profile of a centred PSF stamp: p[k] is the mean
    of the normalised stamp over pixels whose centre distance falls in bin
    k of width ``bin_width``, taken at the bin centre, floored at 1e-12 and
    made non-increasing; a first knot at r = 0 repeats the central value."""
    r: np.ndarray
    p: np.ndarray

    @classmethod
    def of_stamp(cls, stamp, bin_width: float = 0.5) -> "PsfProfile":
        img = np.asarray(stamp, np.float64)
        total = img.sum()
        if not (np.isfinite(total) and total > 0):
            raise ValueError("psf_radial_profile: stamp must have positive sum")
        img = img / total
        h, w = img.shape
        dist = np.hypot(*np.meshgrid(np.arange(h) - (h - 1) / 2.0,
                                     np.arange(w) - (w - 1) / 2.0, indexing="ij")).ravel()
        n_bin = int(np.ceil(dist.max() / bin_width)) + 1
        k = np.minimum((dist / bin_width).astype(np.int64), n_bin - 1)
        total_in = np.bincount(k, weights=img.ravel(), minlength=n_bin)
        pix_in = np.bincount(k, minlength=n_bin)
        used = np.flatnonzero(pix_in)
        mean = np.minimum.accumulate(np.maximum(total_in[used] / pix_in[used], 1e-12))
        return cls(np.r_[0.0, (used + 0.5) * bin_width], np.r_[mean[0], mean])

    def radius_at_level(self, level) -> np.ndarray:
        """Radius where the profile falls to ``level`` (linear in log p):
        0 at or above the central value and for +inf, the outermost radius
        for levels below the profile or not positive."""
        lev = np.asarray(level, np.float64)
        out = np.full(lev.shape, self.r[-1])
        ok = np.isfinite(lev) & (lev > 0)
        logp = np.log(self.p[::-1])               # ascending
        x = np.clip(np.log(np.where(ok, lev, 1.0)), logp[0], logp[-1])
        out = np.where(ok, np.interp(x, logp, self.r[::-1]), out)
        inside = (np.isfinite(lev) & (lev >= self.p[0])) | (np.isinf(lev) & (lev > 0))
        return np.where(inside, 0.0, out)


def psf_radial_profile(stamp, bin_width: float = 0.5):
    """(r, p) of :class:`PsfProfile` (the reference's return layout)."""
    prof = PsfProfile.of_stamp(stamp, bin_width)
    return prof.r, prof.p


def _sersic_falloff(logre, sersic_n):
    """(Re, n_eff, b, I0) of the unit-flux profile I0 exp(-b (r/Re)^(1/n_eff))
    that bounds an extended source's light: an exponential below n = 1.5,
    b = 2n - 1/3 above."""
    from scipy.special import gamma
    re = _clean_re(logre)
    n = np.asarray(sersic_n, np.float64)
    n_eff = np.where(np.isfinite(n) & (n >= 1.5), n, 1.0)
    b = np.where(n_eff < 1.5, SERSIC_B1, 2.0 * n_eff - 1.0 / 3.0)
    i0 = b ** (2.0 * n_eff) / (2.0 * np.pi * n_eff * gamma(2.0 * n_eff) * re ** 2)
    return re, n_eff, b, i0


def flux_aware_radii(source_type, logre, sersic_n, flux, sigma_sky: float, kappa: float,
                     psf_r, psf_p, r_psf: float, r_core: float | None = None) -> np.ndarray:
    """Per-source radius where flux x (unit profile) equals kappa x sigma_sky,
    floored at ``r_core`` (default the PSF FWHM) and capped at the
    geometric radius; sources without a positive finite flux get the
    floor."""
    if not (np.isfinite(sigma_sky) and sigma_sky > 0):
        raise ValueError(f"sigma_sky must be positive, got {sigma_sky}")
    if not kappa > 0:
        raise ValueError(f"kappa must be positive, got {kappa}")
    prof = PsfProfile(np.asarray(psf_r, np.float64), np.asarray(psf_p, np.float64))
    if r_core is None:
        r_core = 2.0 * float(prof.radius_at_level(np.array([0.5 * prof.p[0]]))[0])
    if not r_core > 0:
        raise ValueError(f"r_core must be positive, got {r_core}")
    f = np.asarray(flux, np.float64)
    lit = np.isfinite(f) & (f > 0)
    level = np.where(lit, kappa * float(sigma_sky) / np.where(lit, f, 1.0), np.inf)
    re, n_eff, b, i0 = _sersic_falloff(logre, sersic_n)
    with np.errstate(divide="ignore", invalid="ignore"):
        depth = np.log(i0 / level)                   # -inf without flux
    r_ext = np.where(depth > 0, re * (np.maximum(depth, 0.0) / b) ** n_eff, 0.0)
    star = np.asarray(source_type) == SOURCE_STAR
    r = np.where(star, prof.radius_at_level(level), r_ext)
    return np.clip(r, r_core, coupling_radii(source_type, logre, sersic_n, r_psf))


@dataclass
class IslandSet:
    n_sources: int
    n_islands: int
    labels: np.ndarray                     # (n,) island of each source
    radius: np.ndarray                     # (n,) coupling radius used
    members: list = field(repr=False)      # per island, ascending source ids
    n_src: np.ndarray = field(repr=False)  # (n_isl,) member counts
    bbox: np.ndarray = field(repr=False)   # (n_isl, 4) x0, y0, x1, y1 half-open
    r_psf: float = 0.0

    @property
    def bbox_size(self) -> np.ndarray:
        return np.maximum(self.bbox[:, 2] - self.bbox[:, 0], self.bbox[:, 3] - self.bbox[:, 1])


def coupled_pairs(px, py, radius):
    """Index pairs (i, j) with |p_i - p_j| < r_i + r_j, each coupled pair at
    least once (a pair of two large-radius sources may appear twice)."""
    n = px.size
    if n < 2:
        return np.empty(0, np.int64), np.empty(0, np.int64)
    xy = np.column_stack([px, py])
    cut = float(np.quantile(radius, _SMALL_RADIUS_QUANTILE))
    small = np.flatnonzero(radius <= cut)
    large = np.flatnonzero(radius > cut)
    cand_i, cand_j = [], []
    if small.size > 1:
        pr = cKDTree(xy[small]).query_pairs(r=2.0 * float(radius[small].max()),
                                            output_type="ndarray")
        cand_i.append(small[pr[:, 0]])
        cand_j.append(small[pr[:, 1]])
    if large.size:
        near = cKDTree(xy).query_ball_point(xy[large], radius[large] + float(radius.max()))
        cnt = np.fromiter(map(len, near), np.int64, large.size)
        cand_i.append(np.repeat(large, cnt))
        cand_j.append(np.concatenate(near).astype(np.int64) if cnt.sum()
                      else np.empty(0, np.int64))
    i = np.concatenate(cand_i) if cand_i else np.empty(0, np.int64)
    j = np.concatenate(cand_j) if cand_j else np.empty(0, np.int64)
    reach = radius[i] + radius[j]
    hit = (i != j) & ((px[i] - px[j]) ** 2 + (py[i] - py[j]) ** 2 < reach * reach)
    return i[hit], j[hit]


def _min_roots(n: int, i, j) -> np.ndarray:
    """Root of every node after joining the pairs (i, j): union-find where a
    root is always hooked under a smaller root, so each component's root is
    its smallest index; full pointer jumping between rounds."""
    parent = np.arange(n)
    while True:
        while True:                                   # flatten to the roots
            up = parent[parent]
            if np.array_equal(up, parent):
                break
            parent = up
        ri, rj = parent[i], parent[j]
        split = ri != rj
        if not split.any():
            return parent
        lo, hi = np.minimum(ri, rj)[split], np.maximum(ri, rj)[split]
        np.minimum.at(parent, hi, lo)


def build_islands(pos_x, pos_y, source_type, logre, sersic_n, r_psf: float, *,
                  radius=None, forced_singleton=None) -> IslandSet:
    """Islands of a catalogue; ``radius`` overrides the geometric radii and
    sources flagged in ``forced_singleton`` couple to nothing (the
    reference's subtract-only sources, islands.py:423-470)."""
    px = np.asarray(pos_x, np.float64)
    py = np.asarray(pos_y, np.float64)
    n = px.size
    if py.size != n:
        raise ValueError("pos_x and pos_y must have the same length")
    if not (np.isfinite(px).all() and np.isfinite(py).all()):
        raise ValueError("build_islands needs finite positions")
    if radius is None:
        st = np.zeros(n, np.int32) if source_type is None else source_type
        radius = coupling_radii(st, np.zeros(n) if logre is None else logre,
                                np.ones(n) if sersic_n is None else sersic_n, r_psf)
    radius = np.asarray(radius, np.float64)
    if radius.size != n:
        raise ValueError("radius must have one entry per source")
    if n and not (np.isfinite(radius).all() and (radius > 0).all()):
        raise ValueError("radius must be finite and positive")
    if n == 0:
        return IslandSet(0, 0, np.empty(0, np.int32), radius, [], np.empty(0, np.int32),
                         np.empty((0, 4), np.int64), float(r_psf))
    i, j = coupled_pairs(px, py, radius)
    if forced_singleton is not None:
        alone = np.asarray(forced_singleton, bool)
        if alone.size != n:
            raise ValueError("forced_singleton must have one entry per source")
        free = ~(alone[i] | alone[j])
        i, j = i[free], j[free]
    root = _min_roots(n, i, j)
    heads, labels = np.unique(root, return_inverse=True)   # heads ascending
    labels = labels.astype(np.int32)
    n_isl = heads.size
    n_src = np.bincount(labels, minlength=n_isl).astype(np.int32)
    by_island = np.argsort(labels, kind="stable").astype(np.int32)
    members = np.split(by_island, np.cumsum(n_src)[:-1])
    box = np.empty((n_isl, 4), np.float64)
    box[:, :2] = np.inf
    box[:, 2:] = -np.inf
    np.minimum.at(box[:, 0], labels, px - radius)
    np.minimum.at(box[:, 1], labels, py - radius)
    np.maximum.at(box[:, 2], labels, px + radius)
    np.maximum.at(box[:, 3], labels, py + radius)
    bbox = np.concatenate([np.floor(box[:, :2]), np.ceil(box[:, 2:])], axis=1).astype(np.int64)
    return IslandSet(n, n_isl, labels, radius, members, n_src, bbox, float(r_psf))


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


def bucket_islands(isl: IslandSet, bbox_classes=DEFAULT_BBOX_CLASSES,
                   nsrc_classes=DEFAULT_NSRC_CLASSES) -> BucketedIslands:
    """Each island into the smallest (box side, member count) class that
    holds it; islands beyond either ladder are giants."""
    bc = np.sort(np.asarray(bbox_classes, np.int64))
    nc = np.sort(np.asarray(nsrc_classes, np.int64))
    kb = np.searchsorted(bc, isl.bbox_size, side="left")
    kn = np.searchsorted(nc, isl.n_src, side="left")
    giant = (kb >= bc.size) | (kn >= nc.size)
    # each member's slot within its island (members are ascending; the set
    # may hold only some of the catalogue's islands)
    first = np.concatenate([[0], np.cumsum(isl.n_src)[:-1]]).astype(np.int64)
    slot = np.zeros(isl.n_sources, np.int64)
    if isl.n_islands:
        order = np.concatenate(isl.members).astype(np.int64)
        slot[order] = np.arange(order.size) - np.repeat(first, isl.n_src)
    buckets = {}
    cls = np.where(giant, -1, kb * nc.size + kn)
    for key in np.unique(cls[~giant]):
        ids = np.flatnonzero(cls == key)
        B, N = int(bc[key // nc.size]), int(nc[key % nc.size])
        src = np.concatenate([isl.members[k] for k in ids]).astype(np.int64)
        row = np.repeat(np.arange(ids.size), isl.n_src[ids])
        midx = np.full((ids.size, N), -1, np.int32)
        midx[row, slot[src]] = src
        buckets[(B, N)] = IslandBucket(B, N, ids.astype(np.int32), midx, midx >= 0,
                                       isl.bbox[ids])
    return BucketedIslands(tuple(int(c) for c in bc), tuple(int(c) for c in nc), buckets,
                           np.flatnonzero(giant).astype(np.int32))
