"""The CPU restatement (oracle/cpu_fit.cpp) against the float64 NumPy
oracle: its float64 build reproduces the oracle's island fits (step counts,
flags, parameters, fluxes) and grid-band solves; its float32 build -- the
CPU baseline the benchmark times -- holds the flux bar on well-conditioned
islands."""
from __future__ import annotations

import numpy as np
import pytest

from _scenes import mixed_classes, oracle_island, small_mixed_scene, small_star_scene, \
    star_classes
from oracle import cpu as ocpu
from oracle import island as oisl
from oracle import linear as olin
from uniphot_amd import forced_phot_sparse as fps


@pytest.fixture(scope="module", autouse=True)
def built():
    from oracle.build_cpu import build
    build()


def _items(sc, classes):
    plans, hosts, pad, d, w = classes
    items, isls = [], []
    for plan, host in zip(plans, hosts):
        B, T = plan.bbox_class, plan.tier
        for r in range(plan.n_islands):
            ox, oy = int(host["ox"][r]), int(host["oy"][r])
            items.append(dict(data=d[:, oy:oy + B, ox:ox + B], w=w[:, oy:oy + B, ox:ox + B],
                              x0b=host["x0b"][r], y0b=host["y0b"][r], st=host["st"][r],
                              sn=host["sn"][r], fn=host["fn_flag"][r], mmask=host["mmask"][r],
                              keep=host["keep"][r], p0=host["p0"][r].ravel(),
                              lower=host["lower"][r].ravel(), upper=host["upper"][r].ravel(),
                              T=T))
            isls.append((oracle_island(plan, host, r, d, w, sc.psfs), host["p0"][r],
                         host["mmask"][r]))
    psfs = [dict(coefs=p.coefs, x_origin=p.x_origin, x_scale=p.x_scale, y_origin=p.y_origin,
                 y_scale=p.y_scale) for p in sc.psfs]
    return items, isls, psfs


def _scenes():
    s1 = small_star_scene(n_src=14, n_bands=2)
    s2 = small_mixed_scene()
    s3 = small_mixed_scene(seed=5, types=(0, 3, 3), n_range=(0.5, 3.0))
    return [(s1, star_classes(s1)), (s2, mixed_classes(s2)),
            (s3, mixed_classes(s3, free_n=s3.stype == 3))]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_cpu_restatement_reproduces_the_oracle_fits(k):
    sc, classes = _scenes()[k]
    items, isls, psfs = _items(sc, classes)
    o = oisl.LMOptions()
    r64 = ocpu.fit_islands(items, psfs, precision="fp64")
    r32 = ocpu.fit_islands(items, psfs, precision="fp32")
    for a, b, (isl, p0, mm) in zip(r64, r32, isls):
        st = oisl.run_island(p0, isl, o)
        f, e = oisl.final_fluxes(st["p"], isl, o)
        assert a["nst"] == st["nst"] and a["conv"] == st["conv"] and a["stall"] == st["stall"]
        dp = np.abs(a["p"] - st["p"]).reshape(-1, 6)[mm]
        # the float32 parameter state agrees to a few ulps (positions near
        # 150 px have a 1.5e-5 px ulp)
        assert dp.max() < 1e-4, dp.max()
        rel = np.abs(a["fluxes"][:, mm] - f[:, mm]) / np.maximum(np.abs(f[:, mm]), e[:, mm])
        assert rel.max() < 1e-5, rel.max()
        np.testing.assert_allclose(a["ferr"][:, mm], e[:, mm], rtol=1e-5)
        H, _, _, _ = oisl.hg(st["p"], isl, o)
        if oisl.kappa_scaled(H) < 50:
            rel32 = np.abs(b["fluxes"][:, mm] - f[:, mm]) / np.maximum(np.abs(f[:, mm]),
                                                                     e[:, mm])
            assert b["nst"] == st["nst"] and rel32.max() < 2e-5, rel32.max()


def test_cpu_band_solve_matches_the_oracle():
    rng = np.random.default_rng(4)
    probs, refs = [], []
    T = 21
    for q in range(3):
        n, h, w = 6 + q, 50, 56
        px, py = rng.uniform(3, w - 4, n), rng.uniform(3, h - 4, n)
        ix, iy = fps.stamp_top_left(px, py, 0, 0, T, T)
        yy, xx = np.mgrid[:T, :T] - T // 2
        stamps = np.stack([np.exp(-(xx ** 2 + yy ** 2) / (2 * s * s))
                           for s in rng.uniform(1, 3, n)])
        stamps /= stamps.sum(axis=(1, 2), keepdims=True)
        data = rng.normal(0, 3, (h, w)) + 20.0
        for s in range(n):
            y0, x0 = iy[s], ix[s]
            ys, xs = slice(max(y0, 0), min(y0 + T, h)), slice(max(x0, 0), min(x0 + T, w))
            data[ys, xs] += 1e3 * stamps[s][ys.start - y0:ys.stop - y0, xs.start - x0:xs.stop - x0]
        invvar = rng.uniform(0.05, 0.2, (h, w))
        good = rng.uniform(size=(h, w)) > 0.05
        idx, msk = olin.neighbor_list(px, py, T, T, k_max=32)
        gx0, gy0 = 100.0 + 10 * q, 200.0
        geo = (150.0, 100.0, 230.0, 80.0)
        basis = olin.bg_basis(h, w, gx0, gy0, *geo, order=1)
        refs.append(olin.solve_band(data, invvar, good, stamps, ix, iy, idx, msk, basis))
        probs.append(dict(data=data, invvar=invvar, good=good, stamps=stamps, ix=ix, iy=iy,
                          nbr=np.where(msk, idx, -1), gx0=gx0, gy0=gy0))
    for prec, tol in (("fp64", 1e-6), ("fp32", 1e-4)):
        out = ocpu.band_solve(probs, geo, precision=prec)
        for (fl, fv, bg, bv), (rfl, rbg, rfv, rbv) in zip(out, refs):
            np.testing.assert_allclose(fl, rfl, rtol=tol)
            np.testing.assert_allclose(fv, rfv, rtol=tol)
            np.testing.assert_allclose(bg, rbg, rtol=tol, atol=tol)
            np.testing.assert_allclose(bv, rbv, rtol=tol)


def test_cpu_stamps_match_the_oracle():
    from oracle import core as ocore
    sc = small_mixed_scene()
    psfs = [dict(coefs=p.coefs, x_origin=p.x_origin, x_scale=p.x_scale, y_origin=p.y_origin,
                 y_scale=p.y_scale) for p in sc.psfs]
    n = sc.x.size
    band = np.arange(n) % 2
    for T in (55, 111):
        got = ocpu.render_stamps(psfs, T, band, sc.x, sc.y, 5.0, 9.0, sc.stype, sc.logre,
                                 sc.ee1, sc.ee2, sc.sersic_n, precision="fp64")
        for i in range(n):
            ps = dict(psfs[band[i]], coefs=ocore.pad_psf_coefs(
                psfs[band[i]]["coefs"].astype(np.float64), T))
            ref, _ = ocore.member_stamp(ps, float(sc.x[i]), float(sc.y[i]), 5.0, 9.0,
                                        int(sc.stype[i]), float(sc.logre[i]), float(sc.ee1[i]),
                                        float(sc.ee2[i]), float(sc.sersic_n[i]))
            assert np.abs(got[i] - ref).max() < 1e-6 * np.abs(ref).max(), (T, i)
    star = np.zeros(n, np.int32)
    got = ocpu.render_stamps(psfs, 55, band, sc.x, sc.y, 5.0, 9.0, star, sc.logre, sc.ee1,
                             sc.ee2, sc.sersic_n, renorm=True, precision="fp64")
    for i in range(3):
        ps = dict(psfs[band[i]], coefs=psfs[band[i]]["coefs"].astype(np.float64))
        ref, _ = ocore.member_stamp(ps, float(sc.x[i]), float(sc.y[i]), 5.0, 9.0, 0,
                                    renorm=True)
        assert np.abs(got[i] - ref).max() < 1e-6 * np.abs(ref).max()
