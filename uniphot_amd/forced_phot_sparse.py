"""Per grid-band linear flux + background solve (host facade).

Mirrors ``tract7dt/fit/grid_gpu/forced_phot_sparse.py``:
``build_neighbor_list`` (:272-334, host kd-tree, rectangular stamp-overlap
predicate, k_max nearest kept), ``stamp_top_left`` (:255-264), the
per-band inputs ``GridBandInputs`` (:65-105) and the per-band solve
``solve_band_sparse`` with its launch / collect halves (:583-709) and
result type.  The device work is batched: ``solve_problems`` renders every
(source, band) stamp with ``uph_render_stamps`` and solves every (grid,
band) problem in one ``uph_band_solve`` launch; the reference-shaped entry
points are one-problem batches of it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from scipy.spatial import cKDTree

from . import _lib
from .source_model import SOURCE_STAR

DEFAULT_K_MAX = 32


@dataclass
class SparseForcedPhotResult:
    fluxes: np.ndarray
    bg_coefs: np.ndarray
    flux_var: np.ndarray
    bg_var: np.ndarray
    converged: bool


@dataclass
class GridBandInputs:
    """One band of one grid (forced_phot_sparse.py:65-105): ``data``,
    ``invvar``, ``pix_good`` (band_h, band_w) of the grid's window (numpy or
    device tensors), source positions in full-image pixels, the band PSF
    (a :class:`PsfPolynomial`), the window's full-image corner (grid_x0,
    grid_y0) and the background normalisation; optional source types and
    shapes (EXP/DEV/SERSIC render through the Fourier mixture)."""
    data: object
    invvar: object
    pix_good: object
    pos_x: object
    pos_y: object
    psf_jax: object
    band_h: int
    band_w: int
    grid_x0: int
    grid_y0: int
    bg_x_origin: float
    bg_x_scale: float
    bg_y_origin: float
    bg_y_scale: float
    bg_basis_flat: object = None
    source_type: object = None
    logre: object = None
    ee1: object = None
    ee2: object = None
    sersic_n: object = None
    render_backend: str = "direct"


@dataclass
class _PendingBandSolve:
    """A launched one-problem band solve (``launch_band_sparse``)."""
    problem: object
    args: tuple
    kw: dict


def launch_band_sparse(inputs: GridBandInputs, neighbor_idx, neighbor_mask, *,
                       reg: float = 1e-9, bg_poly_order: int = 2) -> _PendingBandSolve:
    """First half of :func:`solve_band_sparse` (forced_phot_sparse.py:
    583-680): validates and stages one (grid, band) problem; the device
    work runs in :func:`collect_band_sparse`."""
    import torch
    from .psf_polynomial import psf_device_arrays
    psf = inputs.psf_jax
    T = int(psf.H_psf)
    if int(psf.W_psf) != T:
        raise ValueError(f"square PSF stamps only, got {psf.H_psf}x{psf.W_psf}")
    h, w = int(inputs.band_h), int(inputs.band_w)
    st = None if inputs.source_type is None else np.asarray(inputs.source_type, np.int32)
    if st is not None and inputs.render_backend == "direct" and np.any(st != SOURCE_STAR):
        raise NotImplementedError(
            "render_backend='direct' for extended sources is not part of this build; "
            "the reference's 'auto' resolves to 'fft' whenever extended sources are present")
    nbr = np.where(np.asarray(neighbor_mask, bool), np.asarray(neighbor_idx, np.int64), -1)
    px = np.asarray(inputs.pos_x, np.float32)
    py = np.asarray(inputs.pos_y, np.float32)

    def dev(a, dt):
        if isinstance(a, torch.Tensor):
            return a.to(device="cuda", dtype=dt).reshape(1, h, w).contiguous()
        return torch.from_numpy(np.ascontiguousarray(np.asarray(a), dt).reshape(1, h, w)).cuda()

    good = dev(inputs.pix_good, np.uint8) if not isinstance(inputs.pix_good, torch.Tensor) \
        else inputs.pix_good.to("cuda").to(torch.uint8).reshape(1, h, w).contiguous()
    iv = dev(inputs.invvar, np.float32)
    iv = torch.where(good.bool(), iv, torch.zeros((), dtype=iv.dtype, device=iv.device))
    c, g = psf_device_arrays([psf], T)
    gx0, gy0 = int(inputs.grid_x0), int(inputs.grid_y0)
    prob = BandProblem(0, 0, 0, h, w, px, py, nbr.astype(np.int32), stype=st,
                       logre=None if inputs.logre is None else np.asarray(inputs.logre, np.float32),
                       ee1=None if inputs.ee1 is None else np.asarray(inputs.ee1, np.float32),
                       ee2=None if inputs.ee2 is None else np.asarray(inputs.ee2, np.float32),
                       sersic_n=None if inputs.sersic_n is None
                       else np.asarray(inputs.sersic_n, np.float32))
    up = lambda a: torch.from_numpy(a).cuda()   # noqa: E731
    args = ([prob], up(c), up(g), T, c.shape[1], dev(inputs.data, np.float32), iv, good)
    kw = dict(bg_order=int(bg_poly_order), img_origin=(gx0, gy0), reg=float(reg),
              k_max=int(nbr.shape[1]) if nbr.ndim == 2 else DEFAULT_K_MAX,
              bg_geo=(float(inputs.bg_x_origin), float(inputs.bg_x_scale),
                      float(inputs.bg_y_origin), float(inputs.bg_y_scale)))
    return _PendingBandSolve(prob, args, kw)


def collect_band_sparse(pending: _PendingBandSolve) -> SparseForcedPhotResult:
    """Second half of :func:`solve_band_sparse`: the device solve and the
    copy back; ``converged`` = every flux and coefficient finite."""
    return solve_problems(*pending.args, **pending.kw)[0]


def solve_band_sparse(inputs: GridBandInputs, neighbor_idx, neighbor_mask, reg: float = 1e-9,
                      bg_poly_order: int = 2) -> SparseForcedPhotResult:
    """Fluxes and background coefficients of one band of one grid
    (forced_phot_sparse.py:683-709): stamps rendered on the device (unit-sum
    renormalised when every source is a star, forced_phot_sparse.py:
    624-637), the normal equations over stamp overlaps and the background
    basis, float64 solve and inverse diagonal."""
    return collect_band_sparse(launch_band_sparse(inputs, neighbor_idx, neighbor_mask, reg=reg,
                                                  bg_poly_order=bg_poly_order))


def build_neighbor_list(pos_x, pos_y, H_psf: int, W_psf: int, k_max: int = DEFAULT_K_MAX):
    """(idx (n, k_max) int32 with -1 padding, mask (n, k_max) bool): sources
    whose stamp boxes overlap (|dx| < W_psf and |dy| < H_psf), keeping the
    k_max nearest when more qualify."""
    px = np.asarray(pos_x, np.float64)
    py = np.asarray(pos_y, np.float64)
    n = px.size
    idx = np.full((n, k_max), -1, np.int32)
    msk = np.zeros((n, k_max), bool)
    fin = np.where(np.isfinite(px) & np.isfinite(py))[0]
    if fin.size == 0:
        return idx, msk
    # every pair within the circumscribed radius once, both directions,
    # filtered to overlapping boxes; per source the candidates in ascending
    # index order, truncated to the k_max nearest where more qualify
    tree = cKDTree(np.column_stack([px[fin], py[fin]]))
    pr = tree.query_pairs(float(np.hypot(W_psf, H_psf)), output_type="ndarray")
    if pr.size == 0:
        return idx, msk
    s = fin[np.concatenate([pr[:, 0], pr[:, 1]])]
    c = fin[np.concatenate([pr[:, 1], pr[:, 0]])]
    keep = (np.abs(px[c] - px[s]) < W_psf) & (np.abs(py[c] - py[s]) < H_psf)
    s, c = s[keep], c[keep]
    if s.size == 0:
        return idx, msk
    o = np.lexsort((c, s))
    s, c = s[o], c[o]
    starts = np.flatnonzero(np.r_[True, s[1:] != s[:-1]])
    counts = np.diff(np.r_[starts, s.size])
    rank = np.arange(s.size) - np.repeat(starts, counts)
    fits = np.repeat(counts <= k_max, counts)
    idx[s[fits], rank[fits]] = c[fits]
    msk[s[fits], rank[fits]] = True
    for g in np.flatnonzero(counts > k_max):
        a, z = starts[g], starts[g] + counts[g]
        src, cand = s[a], c[a:z]
        d2 = (px[cand] - px[src]) ** 2 + (py[cand] - py[src]) ** 2
        cand = cand[np.argsort(d2)[:k_max]]
        idx[src, :cand.size] = cand
        msk[src, :cand.size] = True
    return idx, msk


def stamp_top_left(pos_x, pos_y, grid_x0, grid_y0, H_psf: int, W_psf: int):
    """round(pos - grid origin) - size//2 in float32 arithmetic, half to even."""
    ix = np.rint(np.asarray(pos_x, np.float32) - np.float32(grid_x0)).astype(np.int32) - W_psf // 2
    iy = np.rint(np.asarray(pos_y, np.float32) - np.float32(grid_y0)).astype(np.int32) - H_psf // 2
    return ix, iy


@dataclass
class BandProblem:
    """One (grid, band) problem over a window of the device images."""
    band: int
    x0: int
    y0: int
    h: int
    w: int
    pos_x: np.ndarray      # (n,) float32 full-image coordinates
    pos_y: np.ndarray
    nbr_idx: np.ndarray    # (n, k_max) problem-local, -1 padding
    stype: np.ndarray | None = None
    logre: np.ndarray | None = None
    ee1: np.ndarray | None = None
    ee2: np.ndarray | None = None
    sersic_n: np.ndarray | None = None
    origin: tuple | None = None   # (gx0, gy0) for the sub-pixel remainder


def check_windows(win, n_bands: int, H: int, W: int, what: str) -> None:
    """Raise ValueError unless every window (band, x0, y0, h, w) lies inside
    an (n_bands, H, W) image stack (the kernels index the stack without
    bounds checks)."""
    win = np.asarray(win, np.int64).reshape(-1, 5)
    b, x0, y0, h, w = (win[:, i] for i in range(5))
    bad = ((b < 0) | (b >= n_bands) | (x0 < 0) | (y0 < 0) | (h < 0) | (w < 0)
           | (x0 + w > W) | (y0 + h > H))
    if bad.any():
        q = int(np.nonzero(bad)[0][0])
        raise ValueError(f"{what}: window {q} {win[q].tolist()} (band, x0, y0, h, w) is not "
                         f"inside the ({n_bands}, {H}, {W}) image")


def solve_problems(problems, psf_coefs_dev, psf_geo_dev, T: int, n_terms: int, data_dev,
                   invvar_dev, good_dev, *, bg_order: int = 1, bg_geo=(0.0, 1.0, 0.0, 1.0),
                   img_origin=(0, 0), reg: float = 1e-9, k_max: int = DEFAULT_K_MAX,
                   stream=None, sersic_dev=None, timer=None):
    """Solve a batch of problems sharing one stamp tier T.  Returns a list of
    SparseForcedPhotResult (host numpy)."""
    import torch
    if not problems:
        return []
    dev = data_dev.device
    nb, H, W = (int(s) for s in data_dev.shape)
    M = (bg_order + 1) * (bg_order + 2) // 2
    ns = np.array([p.pos_x.size for p in problems], np.int64)
    off = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    ntot = int(off[-1])
    cat = lambda f, dt: np.concatenate([f(p) for p in problems]).astype(dt) if ntot else \
        np.zeros(0, dt)
    x = cat(lambda p: p.pos_x, np.float32)
    y = cat(lambda p: p.pos_y, np.float32)
    # per-problem scalars spread over the problem's sources; windows are in
    # the arrays' pixels, whose (0, 0) is full-image pixel img_origin, and
    # the sub-pixel remainder is taken against the window's full-image
    # corner unless a problem names another origin
    ox0, oy0 = int(img_origin[0]), int(img_origin[1])
    org = [p.origin if p.origin is not None else (p.x0 + ox0, p.y0 + oy0) for p in problems]
    gx0 = np.repeat(np.array([o[0] for o in org], np.float32), ns)
    gy0 = np.repeat(np.array([o[1] for o in org], np.float32), ns)
    band = np.repeat(np.array([p.band for p in problems], np.int32), ns)
    ixt = np.rint(x - gx0).astype(np.int32) - T // 2
    iyt = np.rint(y - gy0).astype(np.int32) - T // 2
    # window-local top-left (window origin may differ from the rounding origin)
    ixt -= np.repeat(np.array([p.x0 + ox0 - o[0] for p, o in zip(problems, org)], np.int32), ns)
    iyt -= np.repeat(np.array([p.y0 + oy0 - o[1] for p, o in zip(problems, org)], np.int32), ns)
    nbr = np.concatenate([p.nbr_idx[:, :k_max] for p in problems]).astype(np.int32) if ntot \
        else np.zeros((0, k_max), np.int32)
    win = np.array([[p.band, p.x0, p.y0, p.h, p.w] for p in problems], np.int32)
    check_windows(win, nb, H, W, "solve_problems")
    # problems of one grid share their type array: test each array once
    ext_of = {}
    for p in problems:
        if p.stype is not None and id(p.stype) not in ext_of:
            ext_of[id(p.stype)] = bool(np.any(p.stype != SOURCE_STAR))
    mixed_flag = np.array([p.stype is not None and ext_of[id(p.stype)] for p in problems])
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    L = _lib.lib()
    P = _lib.ptr
    st = _lib.stream_ptr(stream)
    stamps = torch.empty((max(ntot, 1), T, T), dtype=torch.float32, device=dev)
    # all-star problems renormalise their stamps, problems with an extended
    # source do not (forced_phot_sparse.py:624-637): one render call per kind
    for mixed in (False, True):
        sel = np.where(mixed_flag == mixed)[0]
        idx = np.flatnonzero(np.repeat(mixed_flag == mixed, ns)).astype(np.int64)
        if idx.size == 0:
            continue
        r_x, r_y, r_gx, r_gy, r_b = (up(v[idx]) for v in (x, y, gx0, gy0, band))
        s_out = torch.empty((idx.size, T, T), dtype=torch.float32, device=dev)
        ext = [None] * 5
        scratch_r, sb = None, 0
        if T > 111:   # stamps beyond one workgroup's LDS use the workspace path
            sb = L.uph_render_scratch_bytes(T, max(1, min(int(idx.size), 256)))
            scratch_r = torch.empty(sb, dtype=torch.uint8, device=dev)
        if mixed:
            def gather(attr, dt, fill):
                return np.concatenate([
                    getattr(problems[q], attr) if getattr(problems[q], attr) is not None
                    else np.full(problems[q].pos_x.size, fill) for q in sel]).astype(dt)
            ext = [up(gather("stype", np.int32, 0)), up(gather("logre", np.float32, 0.0)),
                   up(gather("ee1", np.float32, 0.0)), up(gather("ee2", np.float32, 0.0)),
                   up(gather("sersic_n", np.float32, 1.0))]
            n_gal = int((ext[0] != SOURCE_STAR).sum().item())
            if scratch_r is None:
                sb = L.uph_render_scratch_bytes(T, max(1, min(n_gal, 1024)))
                scratch_r = torch.empty(sb, dtype=torch.uint8, device=dev)
            if sersic_dev is None:
                from .sersic_table import device_table
                sersic_dev = device_table(dev)
        ev = timer.begin("render_stamps_kernel") if timer is not None else None
        _lib.check(L.uph_render_stamps(int(idx.size), T, n_terms, P(psf_coefs_dev),
                                       P(psf_geo_dev), P(sersic_dev) if mixed else None,
                                       P(r_b), P(r_x), P(r_y), P(r_gx), P(r_gy),
                                       *(P(e) for e in ext), 0 if mixed else 1, P(s_out),
                                       P(scratch_r), sb, st))
        if ev is not None:
            timer.end(ev, dict(kind="render", n=int(idx.size), T=T, n_terms=n_terms,
                               mixed=bool(mixed)))
        stamps.index_copy_(0, up(idx), s_out)
    max_src = int(ns.max())
    sbytes = L.uph_band_scratch_bytes(len(problems), max_src, bg_order)
    scratch = torch.empty(max(sbytes, 1), dtype=torch.uint8, device=dev)
    f32 = dict(dtype=torch.float32, device=dev)
    fl = torch.empty(max(ntot, 1), **f32)
    fv = torch.empty(max(ntot, 1), **f32)
    bgc = torch.empty((len(problems), M), **f32)
    bgv = torch.empty((len(problems), M), **f32)
    d_off, d_win, d_ix, d_iy, d_nbr = up(off), up(win), up(ixt), up(iyt), up(nbr)
    ev = timer.begin("band_solve_kernel") if timer is not None else None
    _lib.check(L.uph_band_solve(len(problems), T, bg_order, k_max, P(d_off), P(d_win), H, W,
                                int(img_origin[0]), int(img_origin[1]), P(data_dev),
                                P(invvar_dev), P(good_dev), *map(float, bg_geo), P(stamps),
                                P(d_ix), P(d_iy), P(d_nbr), float(reg), P(fl), P(fv), P(bgc),
                                P(bgv), P(scratch), scratch.numel(), st))
    if ev is not None:
        timer.end(ev, dict(kind="band", n_prob=len(problems), n_src=ntot, T=T,
                           pix=int(sum(p.h * p.w for p in problems))))
    flh, fvh = fl.cpu().numpy(), fv.cpu().numpy()
    bgh, bvh = bgc.cpu().numpy(), bgv.cpu().numpy()
    # converged: every flux and background coefficient of the problem finite
    nbad = np.concatenate([[0], np.cumsum(~np.isfinite(flh[:ntot]))])
    ok = (nbad[off[1:]] == nbad[off[:-1]]) & np.isfinite(bgh).all(axis=1)
    return [SparseForcedPhotResult(flh[off[q]:off[q + 1]], bgh[q], fvh[off[q]:off[q + 1]],
                                   bvh[q], bool(ok[q])) for q in range(len(problems))]


def render_unit_stamps(psf_coefs_dev, psf_geo_dev, T: int, n_terms: int, band, pos_x, pos_y,
                       origin_x, origin_y, stype, logre, ee1, ee2, sersic_n, *, stream=None,
                       sersic_dev=None):
    """(n, T, T) unit-flux stamps on the device (``uph_render_stamps``, no
    unit-sum renormalisation): PSF at the position for stars, the Fourier
    mixture convolution for extended sources, Lanczos-3 shift by the
    remainder of (pos - origin) (source_model.py:331-390)."""
    import torch
    dev = psf_coefs_dev.device
    L = _lib.lib()
    P = _lib.ptr
    n = int(np.asarray(pos_x).size)
    out = torch.empty((max(n, 1), T, T), dtype=torch.float32, device=dev)
    if n == 0:
        return out[:0]
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    if sersic_dev is None:
        from .sersic_table import device_table
        sersic_dev = device_table(dev)
    sb = L.uph_render_scratch_bytes(T, max(1, min(n, 1024)))
    scratch = torch.empty(max(sb, 1), dtype=torch.uint8, device=dev)
    args = [up(band, np.int32), up(pos_x, np.float32), up(pos_y, np.float32),
            up(origin_x, np.float32), up(origin_y, np.float32), up(stype, np.int32),
            up(logre, np.float32), up(ee1, np.float32), up(ee2, np.float32),
            up(sersic_n, np.float32)]
    _lib.check(L.uph_render_stamps(n, T, n_terms, P(psf_coefs_dev), P(psf_geo_dev),
                                   P(sersic_dev), *(P(a) for a in args), 0, P(out), P(scratch),
                                   sb, _lib.stream_ptr(stream)))
    return out
