"""Correlated-noise variance factor F_corr of the post-fit error model.

Mirrors ``tract7dt/fit/grid_gpu/error_model.py``:
``measure_correlated_noise_factor`` (:933-1113, the raw-sky estimator on a
band's central crop of at most 4096 px) and
``measure_correlated_noise_factor_residual`` (:1116-1302, the same estimator
on one (grid, band) residual map, restricted to the cleanest quartile of its
256-px tiles), with the per-band choice of ``fit/grid_gpu/pipeline.py:
2411-2458`` (clean-pixel-weighted mean of the grids' valid residual factors,
else the raw-sky factor, else 1).

F = sum_d rho(d) w(d) / w(0) over |dy|, |dx| <= L = min(T - 1, h // 2 - 1,
w // 2 - 1): rho is the mask-normalised circular autocovariance of the
clean, mean-subtracted map divided by its zero lag, w the autocorrelation of
the unit-sum PSF stamp.  The factor is kept (clamped to >= 1) when the ACF's
lag-2 decay is under half the PSF autocorrelation's (q2 < 0.5, the
resampling shape), else it is 1 (the confusion shape).

The device does every pass over pixels (``uph_fcorr_*``: loading, the
exact median / MAD by radix select, the dilated source mask, the tile
statistics, the masked mean subtraction, the windowed autocovariance of the
map and of the PSF stamp); the host combines per-map scalars and the
(2L + 1)^2 lag windows.  The empty-aperture box-sum cross-check
(:1058-1084) is a logged diagnostic that feeds no output and is not built.
"""
from __future__ import annotations

import math

import numpy as np

CLIP_NSIGMA = 3.0
LOW_NSIGMA = 5.0
MAX_CROP = 4096
MIN_SKY_FRAC = 0.20
RES_TILE = 256
RES_KEEP_Q = 0.25
RES_MIN_TILES = 8
RES_MIN_CLEAN = 20000
MAX_L = 56                       # largest lag window the autocovariance launch stages in LDS
PARTIAL_BUDGET = 1 << 30         # bytes of per-workgroup partial sums per launch


def _ring_index(L):
    yy, xx = np.mgrid[-L:L + 1, -L:L + 1]
    return np.rint(np.sqrt(yy * yy + xx * xx)).astype(np.int64)


def _ring_means(a, ring, L):
    """Ring means of a (2L+1)^2 window (error_model.py:155-171)."""
    out = np.full(L + 1, np.nan)
    for k in range(L + 1):
        v = a[ring == k]
        v = v[np.isfinite(v)]
        if v.size:
            out[k] = v.mean()
    return out


def shape_gate(rho, w, L):
    """(is_resampling, q2, label) of error_model.py:174-244."""
    if L < 2:
        return True, math.nan, "window-too-small"
    ring = _ring_index(L)
    rr, wr = _ring_means(rho, ring, L), _ring_means(w, ring, L)
    r1, w1, r2, w2 = rr[1], wr[1], rr[2], wr[2]
    if not (np.isfinite(r1) and np.isfinite(w1) and w1 > 0 and np.isfinite(r2)
            and np.isfinite(w2) and w2 > 0):
        return True, math.nan, "degenerate"
    if not r1 > 0:
        return True, math.nan, "no-short-lag-corr"
    q2 = float((r2 / r1) / (w2 / w1))
    ok = bool(np.isfinite(q2) and q2 < 0.5)
    return ok, q2, ("resampling" if ok else "confusion")


def _full_window(half, L):
    """(2L+1)^2 lag window from the half-plane layout of uph_fcorr_autocov
    (dy = 0: dx = 0..L; dy = 1..L: dx = -L..L), using A(-d) = A(d)."""
    n = 2 * L + 1
    out = np.empty((n, n))
    out[L, L:] = half[:L + 1]
    out[L + 1:, :] = half[L + 1:].reshape(L, n)
    out[L, :L] = out[L, L + 1:][::-1]
    out[:L, :] = out[L + 1:, :][::-1, ::-1]
    return out


class _Maps:
    """Packed float64 maps on the device (map k: hw[k] at off[k])."""

    def __init__(self, hw, dev):
        import torch
        self.hw = np.asarray(hw, np.int32).reshape(-1, 2)
        npx = self.hw[:, 0].astype(np.int64) * self.hw[:, 1]
        self.off = np.concatenate([[0], np.cumsum(npx)[:-1]]).astype(np.int64)
        self.npx = npx
        self.total = int(npx.sum())
        self.max_pix = int(npx.max()) if npx.size else 0
        self.dev = dev
        up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
        self.d_off, self.d_hw = up(self.off, np.int64), up(self.hw, np.int32)
        self.v = torch.empty(max(self.total, 1), dtype=torch.float64, device=dev)
        self.g0 = torch.empty(max(self.total, 1), dtype=torch.uint8, device=dev)
        self.aux = torch.empty(max(self.total, 1), dtype=torch.uint8, device=dev)


def _tiles(hw, ts):
    """Tile tables of each map: (tile_map, tile_yx, tile_base)."""
    tm, ty, base = [], [], []
    for k, (h, w) in enumerate(hw):
        base.append(len(tm))
        for y0 in range(0, int(h), ts):
            for x0 in range(0, int(w), ts):
                tm.append(k)
                ty.append((y0, x0))
    return (np.asarray(tm, np.int32), np.asarray(ty, np.int32).reshape(-1, 2),
            np.asarray(base, np.int32))


def _estimate(src, src_off, src_stride, hw, phis, *, good_u8=None, good_f32=None,
              residual: bool):
    """F_corr of each map: ``src`` a float32 device tensor, map k at
    src_off[k] with row stride src_stride[k]; ``phis`` (n_maps, T, T) the
    PSF stamps.  Returns a list of (F, diag)."""
    import torch
    from . import _lib
    L_ = _lib.lib()
    P = _lib.ptr
    dev = src.device
    hw = np.asarray(hw, np.int32).reshape(-1, 2)
    n = hw.shape[0]
    diags = [dict(fallback=None, n_clean=0) for _ in range(n)]
    F = np.ones(n)
    if n == 0:
        return []
    stream = _lib.stream_ptr(None)
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    mp = _Maps(hw, dev)
    d_soff, d_sst = up(src_off, np.int64), up(src_stride, np.int32)
    n_good = torch.empty(n, dtype=torch.int64, device=dev)
    _lib.check(L_.uph_fcorr_load(n, mp.max_pix, P(mp.d_off), P(mp.d_hw), P(src), P(d_soff),
                                 P(d_sst), P(good_u8) if good_u8 is not None else None,
                                 P(good_f32) if good_f32 is not None else None, P(mp.v),
                                 P(mp.g0), P(n_good), stream))
    ng = n_good.cpu().numpy()
    size = mp.npx.astype(np.float64)
    live = ng >= (RES_MIN_CLEAN if residual else MIN_SKY_FRAC * size)
    for k in np.nonzero(~live)[0]:
        diags[k]["fallback"] = ("too few good residual pixels" if residual
                                else "too few finite sky pixels")

    def select(center):
        nn = np.maximum(ng, 1)           # np.median: mean of ranks (n-1)//2 and n//2
        qmap = np.repeat(np.arange(n, dtype=np.int32), 2)
        rank = np.stack([(nn - 1) // 2, nn // 2], 1).ravel()
        cen = np.repeat(center, 2)
        nw = int(L_.uph_fcorr_select_work_words(2 * n))
        work = torch.empty(nw, dtype=torch.int32, device=dev)
        out = torch.empty(2 * n, dtype=torch.float64, device=dev)
        keep = [up(qmap, np.int32), up(rank, np.int64), up(cen, np.float64)]
        _lib.check(L_.uph_fcorr_select(2 * n, mp.max_pix, P(keep[0]), P(keep[1]), P(keep[2]),
                                       P(mp.d_off), P(mp.d_hw), P(mp.v), P(mp.g0), P(work), nw,
                                       P(out), stream))
        o = out.cpu().numpy().reshape(n, 2)
        return (o[:, 0] + o[:, 1]) / 2.0

    med = select(np.full(n, np.nan))
    mad = 1.4826 * select(med)
    tmap, tyx, tbase = _tiles(hw, RES_TILE)
    d_tmap, d_tyx = up(tmap, np.int32), up(tyx, np.int32)
    stats = torch.empty((max(tmap.size, 1), 3), dtype=torch.float64, device=dev)

    def tile_stats(mask):
        _lib.check(L_.uph_fcorr_tile_stats(int(tmap.size), P(d_tmap), P(d_tyx), RES_TILE,
                                           P(mp.d_off), P(mp.d_hw), P(mp.v), P(mask), P(stats),
                                           stream))
        return stats.cpu().numpy()[:tmap.size]

    bad_mad = live & ~(np.isfinite(mad) & (mad > 0))
    if bad_mad.any():
        # np.std of the good pixels: tile (count, sum, M2) combined exactly
        st = tile_stats(mp.g0)
        for k in np.nonzero(bad_mad)[0]:
            s = st[tmap == k]
            c, S = s[:, 0].sum(), s[:, 1].sum()
            mu = S / c if c > 0 else 0.0
            tm_ = np.where(s[:, 0] > 0, s[:, 1] / np.maximum(s[:, 0], 1.0), 0.0)
            m2 = (s[:, 2] + s[:, 0] * (tm_ - mu) ** 2).sum()
            mad[k] = math.sqrt(m2 / c) if c > 0 else math.nan
    degenerate = live & ~(np.isfinite(mad) & (mad > 0))
    for k in np.nonzero(degenerate)[0]:
        diags[k]["fallback"] = "degenerate residual" if residual else "degenerate sky"
    live &= ~degenerate
    for k in range(n):
        diags[k].update(med=float(med[k]), mad=float(mad[k]))
    d_med, d_mad = up(np.where(np.isfinite(med), med, 0.0), np.float64), \
        up(np.where(np.isfinite(mad), mad, 0.0), np.float64)
    _lib.check(L_.uph_fcorr_clean(n, mp.max_pix, P(mp.d_off), P(mp.d_hw), P(mp.v), P(mp.g0),
                                  P(d_med), P(d_mad), CLIP_NSIGMA,
                                  CLIP_NSIGMA if residual else LOW_NSIGMA, 1 if residual else 0,
                                  P(mp.aux), stream))
    st = tile_stats(mp.aux)
    cnt, tsum, m2 = st[:, 0], st[:, 1], st[:, 2]
    keep_t = np.zeros(tmap.size, np.uint8)
    tmean = np.zeros(tmap.size)
    for k in np.nonzero(live)[0]:
        sel = np.nonzero(tmap == k)[0]
        if not residual:
            c = cnt[sel].sum()
            diags[k]["n_clean"] = int(c)
            if c < MIN_SKY_FRAC * size[k]:
                diags[k]["fallback"] = "too few clean sky pixels"
                live[k] = False
                continue
            keep_t[sel] = 1
            tmean[sel] = tsum[sel].sum() / c
            continue
        h, w = hw[k]
        th = np.minimum(RES_TILE, h - tyx[sel, 0])
        tw = np.minimum(RES_TILE, w - tyx[sel, 1])
        ok = cnt[sel] >= 0.25 * (th * tw)
        tv = np.where(ok, m2[sel] / np.maximum(cnt[sel], 1.0), np.inf)
        fin = np.isfinite(tv)
        if int(fin.sum()) < RES_MIN_TILES:
            diags[k]["fallback"] = "too few clean tiles"
            live[k] = False
            continue
        thr = float(np.quantile(tv[fin], RES_KEEP_Q))
        kt = fin & (tv <= thr)
        nc = int(cnt[sel][kt].sum())
        diags[k].update(n_clean=nc, n_tiles_kept=int(kt.sum()))
        if nc < RES_MIN_CLEAN:
            diags[k]["fallback"] = "too few clean tile pixels"
            live[k] = False
            continue
        keep_t[sel[kt]] = 1
        tmean[sel[kt]] = tsum[sel[kt]] / cnt[sel[kt]]
    d_keep, d_tmean, d_tbase = up(keep_t, np.uint8), up(tmean, np.float64), up(tbase, np.int32)
    _lib.check(L_.uph_fcorr_mask(n, mp.max_pix, P(mp.d_off), P(mp.d_hw), P(d_tbase), RES_TILE,
                                 P(d_keep), P(d_tmean), P(mp.v), P(mp.aux), stream))
    T = int(np.shape(phis)[1])
    Ls = np.minimum(T - 1, np.minimum(hw[:, 0] // 2 - 1, hw[:, 1] // 2 - 1))
    if (Ls[live] > MAX_L).any():
        raise ValueError(f"F_corr: PSF stamp of {T} px exceeds the {MAX_L + 1}-px lag window "
                         f"of the autocovariance launch")
    for L in np.unique(Ls[live]):
        L = int(L)
        ks = np.nonzero(live & (Ls == L))[0]
        if L < 0:
            for k in ks:
                diags[k]["fallback"] = "map too small"
            continue
        acov = _autocov(mp, ks, L, dev)
        # PSF autocorrelation: each stamp zero-padded to 2T x 2T (circular = linear for |d| < T)
        sm = 2 * T
        stamp_maps = np.zeros((ks.size, sm, sm))
        for j, k in enumerate(ks):
            ph = np.asarray(phis[k], np.float64)
            s = ph.sum()
            stamp_maps[j, :T, :T] = ph / s if (np.isfinite(s) and s > 0) else 0.0
        sp = _Maps(np.tile([sm, sm], (ks.size, 1)), dev)
        sp.v.copy_(torch.from_numpy(stamp_maps.ravel()))
        sp.aux.fill_(1)
        wcov = _autocov(sp, np.arange(ks.size), L, dev)
        nlag = int(L_.uph_fcorr_nlag(L))
        for j, k in enumerate(ks):
            A = _full_window(acov[j, :nlag], L)
            K = _full_window(acov[j, nlag:], L)
            C = A / np.maximum(K, 1.0)
            C0 = float(C[L, L])
            ph = np.asarray(phis[k], np.float64)
            psum = float(ph.sum())
            if not (np.isfinite(C0) and C0 > 0):
                diags[k]["fallback"] = "non-positive variance"
                continue
            if not (np.isfinite(psum) and psum > 0):
                diags[k]["fallback"] = "invalid PSF stamp"
                continue
            rho = C / C0
            wf = _full_window(wcov[j, :nlag], L)
            wn = wf / wf[L, L]
            f = float((rho * wn).sum())
            diags[k].update(F_acf=f, rho_1px=0.5 * (float(rho[L, L + 1]) + float(rho[L + 1, L])),
                            L=L)
            if not np.isfinite(f):
                diags[k]["fallback"] = "non-finite F_corr"
                continue
            ok, q2, label = shape_gate(rho, wn, L)
            diags[k].update(q2=q2, acf_shape=label)
            if not ok:
                diags[k]["fallback"] = "confusion-shaped ACF"
                continue
            F[k] = max(f, 1.0)
    return [(float(F[k]), diags[k]) for k in range(n)]


def _autocov(mp: _Maps, ks, L, dev):
    """(len(ks), 2 nlag) float64 on the host: the half-window A and K of maps ks."""
    import torch
    from . import _lib
    L_ = _lib.lib()
    P = _lib.ptr
    nlag = int(L_.uph_fcorr_nlag(L))
    out = np.empty((len(ks), 2 * nlag))
    rows, cols = 8, 64
    ntiles = ((mp.hw[ks, 0] + rows - 1) // rows) * ((mp.hw[ks, 1] + cols - 1) // cols)
    bpm = int(np.clip(int(ntiles.max()) // 16, 1, 64)) if len(ks) else 1
    per_map = bpm * 2 * nlag * 8
    chunk = max(1, min(len(ks), PARTIAL_BUDGET // per_map, 65535))
    for a in range(0, len(ks), chunk):
        sel = np.asarray(ks[a:a + chunk])
        d_off = torch.from_numpy(mp.off[sel].copy()).to(dev)
        d_hw = torch.from_numpy(mp.hw[sel].copy()).to(dev)
        pd = sel.size * bpm * 2 * nlag
        partial = torch.empty(pd, dtype=torch.float64, device=dev)
        res = torch.empty((sel.size, 2 * nlag), dtype=torch.float64, device=dev)
        _lib.check(L_.uph_fcorr_autocov(int(sel.size), bpm, L, _lib.ptr(d_off), _lib.ptr(d_hw),
                                        P(mp.v), P(mp.aux), P(partial), pd, P(res),
                                        _lib.stream_ptr(None)))
        out[a:a + sel.size] = res.cpu().numpy()
    return out


def fcorr_sky(fld, bands=None):
    """Raw-sky F_corr of each band of the field (error_model.py:933-1113,
    pipeline.py:633-692: the band's PSF at (W / 2, H / 2)); returns a list
    of (F, diag)."""
    from .psf_polynomial import evaluate_psf_at_batch_np
    nb, H, W = fld.nb, fld.H, fld.W
    bands = list(range(nb)) if bands is None else list(bands)
    ch, cw = int(min(H, MAX_CROP)), int(min(W, MAX_CROP))
    y0, x0 = (H - ch) // 2, (W - cw) // 2
    phis = []
    for b in bands:
        phis.append(np.asarray(evaluate_psf_at_batch_np(fld.psfs[b], np.array([W * 0.5]),
                                                        np.array([H * 0.5]))[0], np.float64))
    T = max(p.shape[0] for p in phis)
    ph = np.zeros((len(bands), T, T))
    for j, p in enumerate(phis):
        o = (T - p.shape[0]) // 2
        ph[j, o:o + p.shape[0], o:o + p.shape[0]] = p
    src_off = np.array([b * H * W + y0 * W + x0 for b in bands], np.int64)
    out = []
    per = max(1, (512 << 20) // max(ch * cw, 1))    # at most ~512 Mpx of maps per pass
    for a in range(0, len(bands), per):
        sl = slice(a, a + per)
        k = len(bands[sl])
        out += _estimate(fld.data.reshape(-1), src_off[sl], np.full(k, W, np.int32),
                         np.tile([ch, cw], (k, 1)), ph[sl], good_u8=fld.good.reshape(-1),
                         residual=False)
    return out


def fcorr_residual_windows(work, max_pix, win, phi):
    """Residual F_corr of confusion-pass windows: ``work`` the confusion
    work buffer (residual plane 0, good plane 1 of each problem's six),
    ``win`` (n, 5) band, x0, y0, h, w, ``phi`` (n, T, T) the unit-sum
    window PSFs.  Returns a list of (F, n_clean, fallback)."""
    n = len(win)
    hw = np.asarray(win, np.int64)[:, 3:5]
    off = np.arange(n, dtype=np.int64) * 6 * max_pix
    res = _estimate(work, off, hw[:, 1].astype(np.int32), hw, phi,
                    good_f32=work[max_pix:], residual=True)
    return [(F, int(d.get("n_clean", 0)), d.get("fallback")) for F, d in res]


def choose_band_fcorr(per_grid, raw):
    """Final per-band F_corr (pipeline.py:2411-2458); ``per_grid`` a list of
    per-band lists of (F, n_clean, fallback) or None, ``raw`` the raw-sky
    factors.  Returns (F (nb,), source names, residual combination (nb,))."""
    nb = len(raw)
    acc = np.zeros(nb)
    cnt = np.zeros(nb, np.int64)
    for fr in per_grid:
        for b in range(min(nb, len(fr or []))):
            if fr[b] is None:
                continue
            Fv, nc, fb = fr[b]
            if fb is not None or not np.isfinite(Fv) or Fv <= 1.0 or nc <= 0:
                continue
            acc[b] += float(Fv) * int(nc)
            cnt[b] += int(nc)
    fres = np.where(cnt > 0, acc / np.maximum(cnt, 1), np.nan)
    out, src = np.ones(nb), []
    for b in range(nb):
        if np.isfinite(fres[b]) and fres[b] > 1.0:
            out[b] = fres[b]
            src.append("residual")
        elif np.isfinite(raw[b]) and raw[b] > 1.0:
            out[b] = raw[b]
            src.append("raw-sky")
        else:
            src.append("none")
    return out, src, fres
