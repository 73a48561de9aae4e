"""Pins of the float64 grid-band linear solve oracle: the sparse
neighbour-list normal equations against a dense weighted least-squares
solve over the full design matrix (exact while no neighbour list
truncates), the inverse-diagonal variances against the dense covariance,
the zero-coverage known answer of the reference's
tests/test_output_contract_nodata.py:142-222, and the neighbour list
against a brute-force pair scan."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import linear as olin
from uniphot_amd import forced_phot_sparse as fps


def _problem(seed, n=7, h=48, w=52, T=15, order=1):
    rng = np.random.default_rng(seed)
    px = rng.uniform(2, w - 3, n)
    py = rng.uniform(2, h - 3, n)
    ix, iy = fps.stamp_top_left(px, py, 0, 0, T, T)
    yy, xx = np.mgrid[:T, :T] - T // 2
    stamps = np.stack([np.exp(-(xx ** 2 + yy ** 2) / (2 * s * s)) for s in rng.uniform(1, 3, n)])
    stamps /= stamps.sum(axis=(1, 2), keepdims=True)
    # window-normalised plane: well conditioned, so the 1e-9 ridge stays
    # far below the comparison tolerance
    basis = olin.bg_basis(h, w, 100, 200, 100 + (w - 1) / 2, w / 2, 200 + (h - 1) / 2, h / 2,
                          order=order)
    flux = rng.uniform(1e3, 1e4, n)
    model = basis @ rng.uniform(-5, 5, basis.shape[1])
    img = model.reshape(h, w).copy()
    for s in range(n):
        for a in range(T):
            for b in range(T):
                y, x = iy[s] + a, ix[s] + b
                if 0 <= y < h and 0 <= x < w:
                    img[y, x] += flux[s] * stamps[s, a, b]
    data = img + rng.normal(0, 3, img.shape)
    invvar = rng.uniform(0.05, 0.2, img.shape)
    good = rng.uniform(size=img.shape) > 0.05
    return dict(px=px, py=py, ix=ix, iy=iy, stamps=stamps, basis=basis, data=data,
                invvar=invvar, good=good, h=h, w=w, T=T)


def _dense(P):
    h, w, T = P["h"], P["w"], P["T"]
    n = P["stamps"].shape[0]
    cols = []
    for s in range(n):
        c = np.zeros((h, w))
        for a in range(T):
            for b in range(T):
                y, x = P["iy"][s] + a, P["ix"][s] + b
                if 0 <= y < h and 0 <= x < w:
                    c[y, x] = P["stamps"][s, a, b]
        cols.append(c.ravel())
    X = np.column_stack(cols + [P["basis"]])
    wt = np.where(P["good"], P["invvar"], 0.0).ravel()
    return X, wt


@pytest.mark.parametrize("order", [0, 1, 2])
def test_sparse_solve_matches_dense_least_squares(order):
    P = _problem(1 + order, order=order)
    idx, msk = olin.neighbor_list(P["px"], P["py"], P["T"], P["T"], k_max=32)
    fl, bg, fv, bv = olin.solve_band(P["data"], P["invvar"], P["good"], P["stamps"], P["ix"],
                                     P["iy"], idx, msk, P["basis"])
    X, wt = _dense(P)
    sw = np.sqrt(wt)
    ref, *_ = np.linalg.lstsq(X * sw[:, None], P["data"].ravel() * sw, rcond=None)
    n = P["stamps"].shape[0]
    np.testing.assert_allclose(fl, ref[:n], rtol=1e-6)
    np.testing.assert_allclose(bg, ref[n:], rtol=1e-6, atol=1e-6)
    C = np.linalg.inv((X * wt[:, None]).T @ X)
    np.testing.assert_allclose(np.concatenate([fv, bv]), np.diag(C), rtol=1e-6)


def test_zero_coverage_source_gets_zero_flux_and_others_are_unchanged():
    P = _problem(7)
    n = P["stamps"].shape[0]
    # source 0 loses every pixel of its stamp
    y0, x0 = P["iy"][0], P["ix"][0]
    T = P["T"]
    cover = np.zeros((P["h"], P["w"]), bool)
    cover[max(y0, 0):y0 + T, max(x0, 0):x0 + T] = True
    good = P["good"] & ~cover
    idx, msk = olin.neighbor_list(P["px"], P["py"], T, T, k_max=32)
    fl, bg, fv, bv = olin.solve_band(P["data"], P["invvar"], good, P["stamps"], P["ix"],
                                     P["iy"], idx, msk, P["basis"])
    assert fl[0] == 0.0
    assert np.all(np.isfinite(fl[1:])) and np.all(np.isfinite(bg))
    # the covered sources solve the same system as without source 0
    X, _ = _dense(P)
    wt = np.where(good, P["invvar"], 0.0).ravel()
    keep = np.r_[False, np.ones(n - 1, bool), np.ones(P["basis"].shape[1], bool)]
    sw = np.sqrt(wt)
    ref, *_ = np.linalg.lstsq((X * sw[:, None])[:, keep], P["data"].ravel() * sw, rcond=None)
    np.testing.assert_allclose(fl[1:], ref[:n - 1], rtol=1e-6)


def test_neighbor_list_matches_brute_force_and_truncates_to_nearest():
    rng = np.random.default_rng(0)
    px = rng.uniform(0, 120, 90)
    py = rng.uniform(0, 120, 90)
    for k_max in (32, 5):
        idx, msk = fps.build_neighbor_list(px, py, 21, 25, k_max=k_max)
        for s in range(px.size):
            dx, dy = np.abs(px - px[s]), np.abs(py - py[s])
            cand = [t for t in range(px.size) if t != s and dx[t] < 25 and dy[t] < 21]
            got = sorted(idx[s][msk[s]].tolist())
            if len(cand) <= k_max:
                assert got == sorted(cand)
            else:
                d2 = np.array([(px[t] - px[s]) ** 2 + (py[t] - py[s]) ** 2 for t in cand])
                assert len(got) == k_max
                assert np.max(d2[np.isin(cand, got)]) <= np.min(d2[~np.isin(cand, got)])
        ref_idx, ref_msk = olin.neighbor_list(px, py, 21, 25, k_max=k_max)
        for s in range(px.size):
            assert sorted(idx[s][msk[s]]) == sorted(ref_idx[s][ref_msk[s]])


def test_blank_response_closed_forms():
    """A constant residual c on a 200 x 200 window (uniform weight): more
    than half of the pixels see the whole template, where the response is
    c / sum(phi^2) -- the median; data equal to the model: 0."""
    rng = np.random.default_rng(3)
    h = w = 200
    T = 55
    yy, xx = np.mgrid[:T, :T] - T // 2
    phi = np.exp(-0.5 * (xx ** 2 + yy ** 2) / 2.0 ** 2)
    phi /= phi.sum()
    stamps = np.stack([phi, phi])
    ix = np.array([30, 110])
    iy = np.array([60, 20])
    flux = np.array([500.0, 800.0])
    basis = olin.bg_basis(h, w, 0, 0, 99.5, 100.0, 99.5, 100.0, order=1)
    bg = np.array([3.0, 0.5, -0.2])
    model = (basis @ bg).reshape(h, w)
    for s in range(2):
        model[iy[s]:iy[s] + T, ix[s]:ix[s] + T] += flux[s] * stamps[s]
    iv = np.full((h, w), 0.04)
    good = np.ones((h, w), bool)
    assert abs(olin.blank_response(model, iv, good, stamps, ix, iy, flux, bg, basis, phi)) < 1e-9
    c = 1.7
    b = olin.blank_response(model + c, iv, good, stamps, ix, iy, flux, bg, basis, phi)
    assert b == pytest.approx(c / np.sum(phi * phi), rel=1e-9)
    # a wing is taken off the data first
    wing = rng.normal(0, 1, (h, w))
    b2 = olin.blank_response(model + c + wing, iv, good, stamps, ix, iy, flux, bg, basis, phi,
                             wing=wing)
    assert b2 == pytest.approx(b, rel=1e-9)
