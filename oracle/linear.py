"""FP64 restatement of the per grid-band linear flux + background solve
(TEST INFRASTRUCTURE).

Reference: ``tract7dt/fit/grid_gpu/forced_phot_sparse.py`` --
``render_stamps`` (:126-220, unit-sum renorm for all-star grids),
``stamp_top_left`` (:255-264), ``build_neighbor_list`` (:272-334),
``solver_inner`` (:342-554); background basis
``grid_gpu/bg_polynomial.py:125-178``.

Parity pins: tests/test_oracle_linear.py solves the same system as a dense
weighted least-squares problem over the full design matrix with
``numpy.linalg.lstsq`` (exact when no neighbour list truncates), and checks
the zero-coverage known answer of the reference's
``tests/test_output_contract_nodata.py:142-222`` (a source whose stamp has
no valid pixel gets flux 0 while covered sources stay finite).
"""
from __future__ import annotations

import numpy as np

from . import core


def bg_basis(h, w, x0, y0, xo, xs, yo, ys, order=1):
    """(h*w, M) polynomial background basis in normalised full-image
    coordinates (bg_polynomial.py:125-178)."""
    yy, xx = np.mgrid[:h, :w].astype(np.float64)
    X = (xx + x0 - xo) / xs
    Y = (yy + y0 - yo) / ys
    cols = [(X ** m * Y ** n).ravel() for (m, n) in core.terms_for_order(order)]
    return np.stack(cols, axis=-1)


def neighbor_list(px, py, H, W, k_max=32):
    """Sources whose stamp boxes overlap (|dx| < W and |dy| < H), the k_max
    nearest kept (forced_phot_sparse.py:272-334).  Brute force O(n^2)."""
    n = px.size
    idx = np.full((n, k_max), -1, np.int64)
    msk = np.zeros((n, k_max), bool)
    for s in range(n):
        dx = np.abs(px - px[s])
        dy = np.abs(py - py[s])
        cand = np.where((dx < W) & (dy < H))[0]
        cand = cand[cand != s]
        if cand.size > k_max:
            d2 = (px[cand] - px[s]) ** 2 + (py[cand] - py[s]) ** 2
            cand = cand[np.argsort(d2, kind="stable")[:k_max]]
        idx[s, :cand.size] = cand
        msk[s, :cand.size] = True
    return idx, msk


def render_grid_stamps(psf, px, py, gx0, gy0, stype, lr, e1, e2, sn):
    """Unit-flux stamps for the linear solve: all-star grids renormalise the
    raw PSF to unit sum before the shift; grids with any extended source do
    not (forced_phot_sparse.py:624-637)."""
    all_star = bool(np.all(np.asarray(stype) == core.SOURCE_STAR))
    out = []
    for i in range(px.size):
        s, _ = core.member_stamp(psf, float(px[i]), float(py[i]), gx0, gy0,
                                 int(stype[i]), float(lr[i]), float(e1[i]),
                                 float(e2[i]), float(sn[i]), renorm=all_star)
        out.append(s)
    return np.stack(out) if out else np.zeros((0,) + psf["coefs"].shape[:2])


def solve_band(data, invvar, pix_good, stamps, ix_top, iy_top, nbr_idx, nbr_mask,
               basis, reg=1e-9):
    """One grid-band solve (forced_phot_sparse.py:342-554).
    Returns fluxes (n,), bg (M,), flux_var (n,), bg_var (M,)."""
    h, w = data.shape
    n, T = stamps.shape[0], stamps.shape[1]
    M = basis.shape[1]
    wg = np.where(pix_good, invvar, 0.0)
    Bg = basis.reshape(h, w, M)
    ky = np.arange(T)
    diag = np.zeros(n)
    tdv = np.zeros(n)
    tb = np.zeros((n, M))
    eff = np.zeros_like(stamps)
    wu = np.zeros_like(stamps)
    for s in range(n):
        ys = iy_top[s] + ky[:, None]
        xs = ix_top[s] + ky[None, :]
        inb = (ys >= 0) & (ys < h) & (xs >= 0) & (xs < w)
        yc, xc = np.clip(ys, 0, h - 1), np.clip(xs, 0, w - 1)
        wu[s] = np.where(inb, wg[yc, xc], 0.0)
        du = data[yc, xc]
        bu = Bg[yc, xc, :]
        eff[s] = np.where(inb, stamps[s], 0.0)
        ws = wu[s] * eff[s]
        diag[s] = np.sum(ws * eff[s])
        tdv[s] = np.sum(ws * du)
        tb[s] = np.einsum("ij,ijm->m", ws, bu)
    A = np.zeros((n + M, n + M))
    A[np.arange(n), np.arange(n)] = diag
    for s in range(n):
        for kk in range(nbr_idx.shape[1]):
            if not nbr_mask[s, kk]:
                continue
            t = nbr_idx[s, kk]
            di = iy_top[s] - iy_top[t]
            dj = ix_top[s] - ix_top[t]
            tky = ky[:, None] + di
            tkx = ky[None, :] + dj
            in_t = (tky >= 0) & (tky < T) & (tkx >= 0) & (tkx < T)
            tv = np.where(in_t, eff[t][np.clip(tky, 0, T - 1), np.clip(tkx, 0, T - 1)], 0.0)
            A[s, t] += np.sum(wu[s] * eff[s] * tv)
    A[:n, n:] = tb
    A[n:, :n] = tb.T
    bw = basis * wg.reshape(-1)[:, None]
    A[n:, n:] = bw.T @ basis
    rhs = np.concatenate([tdv, bw.T @ data.reshape(-1)])
    dg = np.diag(A)
    A = A + np.diag(reg * np.maximum(np.abs(dg), 1e-30))
    x = np.linalg.solve(A, rhs)
    inv = np.linalg.inv(A)
    var = np.diag(inv)
    return x[:n], x[n:], var[:n], var[n:]


def wing_image(psf, px, py, gx0, gy0, stype, lr, e1, e2, sn, flux, h, w, Tb, Ts):
    """Wing of the big sources on an (h, w) window whose pixel (0, 0) is
    full-image (gx0, gy0) (solver_grid.py:1578-1626): sum over sources of
    flux x (unit render at stamp size Tb - unit render at Ts), each render
    placed at round(pos - origin) - size // 2 and clipped to the window."""
    out = np.zeros((h, w))
    for T, sign in ((Tb, 1.0), (Ts, -1.0)):
        p = dict(psf, coefs=core.pad_psf_coefs(psf["coefs"], T))
        for i in range(px.size):
            s, _ = core.member_stamp(p, float(px[i]), float(py[i]), gx0, gy0, int(stype[i]),
                                     float(lr[i]), float(e1[i]), float(e2[i]), float(sn[i]))
            ix, iy = core.stamp_top_left(float(px[i]), float(py[i]), gx0, gy0, T)
            y0, x0 = max(0, iy), max(0, ix)
            y1, x1 = min(h, iy + T), min(w, ix + T)
            if y1 > y0 and x1 > x0:
                out[y0:y1, x0:x1] += sign * float(flux[i]) * s[y0 - iy:y1 - iy, x0 - ix:x1 - ix]
    return out


def blank_response(data, invvar, pix_good, stamps, ix_top, iy_top, flux, bg_coefs, basis, phi,
                   wing=None):
    """Blank-response median of one grid-band window (solver_grid.py:2867-
    2943): model = sum flux x unit stamps (placed, clipped) + basis @ bg;
    w = invvar on good pixels, r = d - wing - model there; resp =
    corr(w r, phi) / corr(w, phi^2), 'same' correlation through
    scipy.signal.fftconvolve with the flipped template as the reference
    does; NaN off the good pixels or where the denominator is 0;
    b = nanmedian(resp)."""
    from scipy.signal import fftconvolve
    h, w = data.shape
    T = stamps.shape[1]
    model = (basis @ np.asarray(bg_coefs, np.float64)).reshape(h, w)
    for s in range(stamps.shape[0]):
        y0, x0 = max(0, iy_top[s]), max(0, ix_top[s])
        y1, x1 = min(h, iy_top[s] + T), min(w, ix_top[s] + T)
        if y1 > y0 and x1 > x0:
            model[y0:y1, x0:x1] += float(flux[s]) * stamps[s][y0 - iy_top[s]:y1 - iy_top[s],
                                                               x0 - ix_top[s]:x1 - ix_top[s]]
    d = data - (wing if wing is not None else 0.0)
    good = pix_good & (invvar > 0)
    wgt = np.where(good, invvar, 0.0)
    r = np.where(good, d - model, 0.0)
    num = fftconvolve(wgt * r, phi[::-1, ::-1], mode="same")
    den = fftconvolve(wgt, (phi * phi)[::-1, ::-1], mode="same")
    with np.errstate(divide="ignore", invalid="ignore"):
        resp = np.where(den > 0, num / np.maximum(den, 1e-30), np.nan)
    resp = np.where(good, resp, np.nan)
    return float(np.nanmedian(resp)) if np.isfinite(resp).any() else float("nan")
