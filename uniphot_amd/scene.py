"""Scene subtraction for the island fit (island_stage.py:1762-2117,
island_solver.py:1795-1830 and :2779-2789, island_plan.py:664-680).

Every catalogued STAR with a non-zero S1 flux is rendered at that flux and
subtracted once from the padded field stack (``subtract_stars``).  Each
island's B x B window is then gathered into a mosaic of island cutouts and
the island's own member stars are added back there at their seeds and S1
fluxes (``island_mosaic``), so the joint fit sees its members at data level
while the light of stars outside the island is removed.  Extended sources
are never subtracted (their seed shapes are not trusted for it).

Both steps run on the device: stamps come from ``uph_render_stamps`` (the
stamp renderer of the linear solve), the weighted sums from
``uph_stamp_sum`` over disjoint windows (entries summed in catalog order, so
the result does not depend on the launch shape) and the cutouts from
``uph_gather_windows``.  The island kernels read only their B x B window,
so the mosaic needs no margins.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .forced_phot_sparse import render_unit_stamps
from .source_model import SOURCE_STAR

SCENE_TILE = 64


@dataclass
class SceneStars:
    """What the member add-back needs: catalog seeds, types, the S1 fluxes
    (n, nb) with NaN read as 0 (island_stage.py:1699), and ``psf_of(T)`` ->
    device (coefs, geo) of the band PSFs padded to stamp tier T."""
    x: np.ndarray
    y: np.ndarray
    stype: np.ndarray
    flux: np.ndarray
    psf_of: object
    n_terms: int


def star_flux(s1_flux) -> np.ndarray:
    return np.nan_to_num(np.asarray(s1_flux, np.float32), nan=0.0)


def _rint(v) -> np.ndarray:
    """Round half to even, as numpy and the stamp renderer do."""
    return np.rint(np.asarray(v, np.float32)).astype(np.int64)


def tile_lists(ix, iy, T: int, H: int, W: int, tile: int = SCENE_TILE):
    """Tiles of an (H, W) image touched by T x T stamps at top-left (ix, iy):
    (off (n_tiles + 1) i64, win (n_tiles, 4) i32 = x0, y0, h, w,
    list (entries) i32 stamp index), entries of a tile in stamp order.
    Stamps wholly outside the image get no entry."""
    ix = np.asarray(ix, np.int64)
    iy = np.asarray(iy, np.int64)
    ntx, nty = (W + tile - 1) // tile, (H + tile - 1) // tile
    x0, x1 = np.maximum(ix, 0), np.minimum(ix + T, W)
    y0, y1 = np.maximum(iy, 0), np.minimum(iy + T, H)
    inside = (x1 > x0) & (y1 > y0)
    tx0, tx1 = x0 // tile, (x1 - 1) // tile
    ty0, ty1 = y0 // tile, (y1 - 1) // tile
    span = (T + tile - 1) // tile + 1
    src, tid = [], []
    for dy in range(span):
        for dx in range(span):
            tx, ty = tx0 + dx, ty0 + dy
            ok = inside & (tx <= tx1) & (ty <= ty1)
            s = np.nonzero(ok)[0]
            src.append(s)
            tid.append(ty[s] * ntx + tx[s])
    src = np.concatenate(src)
    tid = np.concatenate(tid)
    order = np.lexsort((src, tid))
    src, tid = src[order], tid[order]
    tiles, start = np.unique(tid, return_index=True)
    off = np.append(start, src.size).astype(np.int64)
    tx, ty = tiles % ntx, tiles // ntx
    win = np.stack([tx * tile, ty * tile, np.minimum(tile, H - ty * tile),
                    np.minimum(tile, W - tx * tile)], axis=1).astype(np.int32)
    return off, win, src.astype(np.int32)


def _band_chunk(n_st: int, T: int, nb: int, budget: float) -> int:
    return int(max(1, min(nb, budget // max(1.0, n_st * T * T * 4.0))))


def _stamp_sum(out, T, n_st, nbc, b0, off, win, lst, ix, iy, stamps, flux, scale, max_pix,
               stream=None):
    import torch
    dev = out.device
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    L, P = _lib.lib(), _lib.ptr
    t = [up(off, np.int64), up(win, np.int32), up(lst, np.int32), up(ix, np.int32),
         up(iy, np.int32)]
    fl = up(flux, np.float32)
    nb, H, W = (int(s) for s in out.shape)
    _lib.check(L.uph_stamp_sum(int(win.shape[0]), int(max_pix), int(T), int(n_st), int(nbc),
                               int(b0), nb, *(P(a) for a in t), P(stamps), P(fl), float(scale),
                               H, W, P(out), _lib.stream_ptr(stream)))
    return t, fl


def subtract_stars(dstack, pad: int, psf_coefs, psf_geo, T: int, n_terms: int, x, y, stype,
                   flux, *, tile: int = SCENE_TILE, stamp_budget: float = 2e9, stream=None):
    """Field-wide STAR subtraction at S1 flux, in place on the padded stack
    dstack (nb, H + 2 pad, W + 2 pad) (island_stage.py:2040-2117): star k's
    native T x T stamp (PSF at (x, y), Lanczos-shifted by x - round(x)) times
    flux[k, b] at top-left round(pos) - T // 2 + pad, clipped to the stack.
    Returns the number of stars subtracted."""
    flux = star_flux(flux)
    stype = np.asarray(stype)
    nb, Hp, Wp = (int(s) for s in dstack.shape)
    stars = np.nonzero((stype == SOURCE_STAR) & np.any(flux != 0.0, axis=1))[0]
    if stars.size == 0:
        return 0
    xs = np.asarray(x, np.float32)[stars]
    ys = np.asarray(y, np.float32)[stars]
    ix = _rint(xs) - T // 2 + pad
    iy = _rint(ys) - T // 2 + pad
    off, win, lst = tile_lists(ix, iy, T, Hp, Wp, tile)   # tiles inside the stack
    if win.shape[0] == 0:
        return 0
    n_st = stars.size
    nbc = _band_chunk(n_st, T, nb, stamp_budget)
    z = np.zeros(n_st, np.float32)
    for b0 in range(0, nb, nbc):
        m = min(nbc, nb - b0)
        bands = np.repeat(np.arange(b0, b0 + m, dtype=np.int32), n_st)
        stamps = render_unit_stamps(psf_coefs, psf_geo, T, n_terms, bands, np.tile(xs, m),
                                    np.tile(ys, m), np.zeros(n_st * m, np.float32),
                                    np.zeros(n_st * m, np.float32),
                                    np.zeros(n_st * m, np.int32), np.tile(z, m), np.tile(z, m),
                                    np.tile(z, m), np.ones(n_st * m, np.float32), stream=stream)
        keep = _stamp_sum(dstack, T, n_st, m, b0, off, win, lst, ix, iy, stamps,
                          flux[stars, b0:b0 + m], -1.0, tile * tile, stream)
        del keep, stamps
    return int(n_st)


def shelf_pack(sides, width: int | None = None):
    """Top-left (mx, my) of square windows in a mosaic (shelves of equal
    height, widest windows first) and the mosaic shape (Hm, Wm)."""
    sides = np.asarray(sides, np.int64)
    n = sides.size
    if width is None:
        width = max(int(sides.max()), int(np.ceil(np.sqrt(float(np.sum(sides * sides))))))
    order = np.argsort(-sides, kind="stable")
    mx = np.zeros(n, np.int64)
    my = np.zeros(n, np.int64)
    cx = cy = shelf = 0
    for i in order:
        s = int(sides[i])
        if cx + s > width:
            cy += shelf
            cx = shelf = 0
        mx[i], my[i] = cx, cy
        cx += s
        shelf = max(shelf, s)
    return mx, my, cy + shelf, width


def mosaic_bytes(nb: int, B: int) -> float:
    """Device bytes one island adds to a launch sequence's mosaic."""
    return 8.0 * nb * B * B


def island_mosaic(chunk, plan_of, dstack, wstack, scene: SceneStars, *,
                  stamp_budget: float = 2e9, stream=None):
    """Cutouts of the islands of a launch sequence [(kernels, host, rows)]:
    (dmos, wmos) (nb, Hm, Wm) holding each island's window of the
    scene-subtracted dstack (and of wstack) with its member stars added back
    at their seeds and S1 fluxes (island_solver.py:2779-2789; the add-back
    renders them as the fit does, at the class's stamp tier), and the chunk
    with hosts whose window origins (ox, oy) point into the mosaic."""
    import torch
    L, P = _lib.lib(), _lib.ptr
    nb = int(dstack.shape[0])
    dev = dstack.device
    sides, sx, sy, own = [], [], [], []
    for j, (kern, host, rows) in enumerate(chunk):
        B = int(plan_of[id(host)].bbox_class)
        rows = np.asarray(rows)
        sides.append(np.full(rows.size, B, np.int64))
        sx.append(np.asarray(host["ox"])[rows])
        sy.append(np.asarray(host["oy"])[rows])
        own.append(np.full(rows.size, j, np.int64))
    sides = np.concatenate(sides)
    sx = np.concatenate(sx)
    sy = np.concatenate(sy)
    own = np.concatenate(own)
    Hs, Ws = int(dstack.shape[1]), int(dstack.shape[2])
    if (sx.min() < 0 or sy.min() < 0 or np.any(sx + sides > Ws) or np.any(sy + sides > Hs)):
        raise ValueError("island_mosaic: an island window is not inside the field stack")
    mx, my, Hm, Wm = shelf_pack(sides)
    dmos = torch.empty((nb, Hm, Wm), dtype=torch.float32, device=dev)
    wmos = torch.empty_like(dmos)
    win = np.stack([sx, sy, mx, my, sides], axis=1).astype(np.int32)
    wd = torch.from_numpy(win).to(dev)
    _lib.check(L.uph_gather_windows(int(win.shape[0]), int(sides.max()), nb, P(wd),
                                    int(dstack.shape[1]), int(dstack.shape[2]), P(dstack),
                                    P(wstack), Hm, Wm, P(dmos), P(wmos),
                                    _lib.stream_ptr(stream)))
    # member stars per tier
    new_chunk, base = [], 0
    per_tier = {}
    for j, (kern, host, rows) in enumerate(chunk):
        plan = plan_of[id(host)]
        rows = np.asarray(rows)
        k = rows.size
        qm = np.arange(base, base + k)
        h2 = dict(host)
        h2["ox"] = np.array(host["ox"], copy=True)
        h2["oy"] = np.array(host["oy"], copy=True)
        h2["ox"][rows] = mx[qm]
        h2["oy"][rows] = my[qm]
        new_chunk.append((kern, h2, rows))
        st = np.asarray(host["st"])[rows]
        mm = np.asarray(host["mmask"])[rows].astype(bool) & (st == SOURCE_STAR)
        src = np.asarray(plan.midx)[rows]
        r_i, m_i = np.nonzero(mm)
        if r_i.size:
            T = int(plan.tier)
            acc = per_tier.setdefault(T, dict(q=[], src=[], x0b=[], y0b=[], B=[]))
            acc["q"].append(qm[r_i])
            acc["src"].append(src[r_i, m_i])
            acc["x0b"].append(np.asarray(host["x0b"])[rows][r_i])
            acc["y0b"].append(np.asarray(host["y0b"])[rows][r_i])
        base += k
    keep = [wd]
    for T, acc in sorted(per_tier.items()):
        q = np.concatenate(acc["q"])
        src = np.concatenate(acc["src"])
        x0b = _rint(np.concatenate(acc["x0b"]))
        y0b = _rint(np.concatenate(acc["y0b"]))
        flux = scene.flux[src]
        live = np.any(flux != 0.0, axis=1)
        q, src, x0b, y0b, flux = q[live], src[live], x0b[live], y0b[live], flux[live]
        if src.size == 0:
            continue
        order = np.lexsort((np.arange(q.size), q))      # windows in order, members in slot order
        q, src, x0b, y0b, flux = q[order], src[order], x0b[order], y0b[order], flux[order]
        xs = scene.x[src].astype(np.float32)
        ys = scene.y[src].astype(np.float32)
        ix = mx[q] + (_rint(xs) - x0b) - T // 2
        iy = my[q] + (_rint(ys) - y0b) - T // 2
        wins, start = np.unique(q, return_index=True)
        off = np.append(start, q.size).astype(np.int64)
        w4 = np.stack([mx[wins], my[wins], sides[wins], sides[wins]], axis=1)
        n_st = src.size
        coefs, geo = scene.psf_of(T)
        nbc = _band_chunk(n_st, T, nb, stamp_budget)
        z = np.zeros(n_st, np.float32)
        lst = np.arange(n_st, dtype=np.int32)
        max_pix = int(sides[wins].max()) ** 2
        for b0 in range(0, nb, nbc):
            m = min(nbc, nb - b0)
            bands = np.repeat(np.arange(b0, b0 + m, dtype=np.int32), n_st)
            stamps = render_unit_stamps(coefs, geo, T, scene.n_terms, bands, np.tile(xs, m),
                                        np.tile(ys, m), np.zeros(n_st * m, np.float32),
                                        np.zeros(n_st * m, np.float32),
                                        np.zeros(n_st * m, np.int32), np.tile(z, m),
                                        np.tile(z, m), np.tile(z, m),
                                        np.ones(n_st * m, np.float32), stream=stream)
            keep.append(_stamp_sum(dmos, T, n_st, m, b0, off, w4, lst, ix, iy, stamps,
                                   flux[:, b0:b0 + m], 1.0, max_pix, stream))
            keep.append(stamps)
    del keep   # stream-ordered frees: later work on the stream reuses the memory
    return dmos, wmos, new_chunk
