"""End-to-end island-mode fit on the GPU (pass 1 -> islands -> LM -> pass 2)
against the float64 oracle stage by stage on a small seeded star field."""
import numpy as np
import pytest

from _scenes import mixed_classes, small_mixed_scene, small_star_scene
from oracle import core as ocore
from oracle import island as oisl
from oracle import linear as olin

pytestmark = pytest.mark.gpu


def test_fit_field_star_scene(cuda):
    import torch
    from uniphot_amd import forced_phot_sparse as fps
    from uniphot_amd.grid import assign_sources_to_grids, divide_image_into_grids
    from uniphot_amd.pipeline import DeviceField, FitConfig, fit_field
    sc = small_star_scene(seed=11, H=240, W=240, n_bands=3, n_src=40)
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    cfg = FitConfig(n_rows=2, n_cols=2)
    res = fit_field(fld, sc.x, sc.y, sc.stype, sc.logre, sc.ee1, sc.ee2, sc.sersic_n, cfg,
                    sync=torch.cuda.synchronize, keep_artifacts=True)
    n = sc.x.size
    assert res.record["n_islands"] >= 1 and res.flux.shape == (n, 3)
    assert np.all(res.island_conv | res.island_giant)
    # every fitted island: GPU vs oracle LM on the same stacks
    art = res.record["_artifacts"]
    w = art["wstack"].cpu().numpy()
    psfs = [dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                 y_origin=p.y_origin, y_scale=p.y_scale) for p in sc.psfs]
    # the island cutouts of the scene subtraction, restated by the oracle
    s1 = np.nan_to_num(res.s1_flux.astype(np.float64), nan=0.0)
    draw = art["dstack_raw"].cpu().numpy().astype(np.float64)
    d = draw - oisl.scene_star_image(draw.shape, art["pad"], psfs, sc.x, sc.y, sc.stype, s1)
    o = oisl.LMOptions()
    checked = 0
    for plan, host in art["classes"]:
        for r in range(plan.n_islands):
            B, T = plan.bbox_class, plan.tier
            ox, oy = int(host["ox"][r]), int(host["oy"][r])
            mem = plan.midx[r][host["mmask"][r]]
            add = oisl.member_addback([dict(q, coefs=ocore.pad_psf_coefs(q["coefs"], T))
                                       for q in psfs], host["x0b"][r], host["y0b"][r], B,
                                      sc.x[mem], sc.y[mem], sc.stype[mem], s1[mem])
            isl = oisl.Island(d[:, oy:oy + B, ox:ox + B] + add, w[:, oy:oy + B, ox:ox + B],
                              host["x0b"][r], host["y0b"][r], host["st"][r], host["sn"][r],
                              host["fn_flag"][r], host["mmask"][r], host["keep"][r],
                              host["lower"][r].ravel(), host["upper"][r].ravel(), psfs, T)
            st = oisl.run_island(host["p0"][r].ravel().astype(np.float64), isl, o)
            src = plan.midx[r][host["mmask"][r]]
            po = st["p"].reshape(-1, 6)[host["mmask"][r]]
            assert np.max(np.abs(res.x[src] - po[:, 0])) < 1e-4
            assert np.max(np.abs(res.y[src] - po[:, 1])) < 1e-4
            assert np.all(res.island_nst[src] == st["nst"])
            checked += 1
    assert checked == res.record["n_fitted_islands"]
    # pass 2 at the GPU-fitted positions: oracle linear solve per grid-band,
    # then the blank-response correction of the STAR fluxes (whose device
    # median agrees with the oracle's to 1e-3 sigma, test_gpu_linear.py)
    from uniphot_amd.pipeline import _pad_center
    from uniphot_amd.psf_polynomial import evaluate_psf_at_np
    halo = cfg.halo
    grids = divide_image_into_grids(240, 240, 2, 2, halo)
    assign_sources_to_grids(grids, sc.x, sc.y)   # pass-2 keeps the pass-1 grid fit sets
    bg_geo = (0.5 * 239, 120.0, 0.5 * 239, 120.0)
    for g in grids:
        fi = g.fit_indices
        xs, ys = res.x[fi], res.y[fi]
        nbr, msk = fps.build_neighbor_list(xs, ys, 55, 55)
        ix, iy = fps.stamp_top_left(xs, ys, g.x0_pad, g.y0_pad, 55, 55)
        basis = olin.bg_basis(g.pad_h, g.pad_w, g.x0_pad, g.y0_pad, *bg_geo, order=1)
        for b in range(3):
            stamps = np.stack([ocore.member_stamp(psfs[b], float(a), float(c), g.x0_pad,
                                                  g.y0_pad, 0, renorm=True)[0]
                               for a, c in zip(xs, ys)])
            sl = (b, slice(g.y0_pad, g.y1_pad), slice(g.x0_pad, g.x1_pad))
            f, bgc, fv, _ = olin.solve_band(sc.data[sl].astype(np.float64),
                                            sc.invvar[sl].astype(np.float64), sc.good[sl],
                                            stamps, ix, iy, nbr, msk, basis)
            raw = np.stack([ocore.member_stamp(psfs[b], float(a), float(c), g.x0_pad,
                                               g.y0_pad, 0)[0] for a, c in zip(xs, ys)])
            phi = np.clip(evaluate_psf_at_np(sc.psfs[b], g.x0_pad + 0.5 * g.pad_w,
                                             g.y0_pad + 0.5 * g.pad_h).astype(np.float64),
                          0.0, None)
            phi = _pad_center(phi, 55)
            phi /= phi.sum()
            bh = olin.blank_response(sc.data[sl].astype(np.float64),
                                     sc.invvar[sl].astype(np.float64), sc.good[sl], raw, ix, iy,
                                     f, bgc, basis, phi)
            own = g.source_indices
            fo = f[:own.size] - (bh if np.isfinite(bh) else 0.0)
            sig = np.sqrt(fv[:own.size])
            err = np.abs(res.flux[own, b] - fo)
            assert np.all(err < 1e-5 * np.maximum(np.abs(fo), sig) + 1e-3 * sig), \
                (g.grid_id, b, np.max(err / np.maximum(np.abs(fo), sig)))
    # sanity against the truth (the order-2 PSF model and the Lanczos shift of
    # an undersampled core differ from the pixel-integrated truth at the
    # percent level, so this is a loose check)
    rel_truth = np.abs(res.flux - sc.true_flux) / sc.true_flux
    assert np.median(rel_truth) < 0.05
    # catalogue writeback of the same fit
    import pandas as pd
    from uniphot_amd.output import write_fit_columns
    out = write_fit_columns(pd.DataFrame(dict(ID=np.arange(n))), res, ["a", "b", "c"],
                            x_init=sc.x, y_init=sc.y, source_type=sc.stype)
    for b, bn in enumerate("abc"):
        np.testing.assert_array_equal(out[f"FLUX_{bn}_fit"], res.flux[:, b])
        np.testing.assert_allclose(out[f"FLUXERR_{bn}_fit"], np.sqrt(res.flux_var[:, b]),
                                   rtol=1e-6)
    assert out["opt_converged"].all() and (out["grid_id_fit"] >= 0).all()
    fitted = (res.island_id >= 0) & ~res.island_giant
    assert np.all(np.isfinite(out["x_pix_white_err"][fitted]))


def test_scene_subtract_and_island_cutouts_match_oracle(cuda):
    """Field-wide STAR subtraction at S1 flux and the island cutout mosaic
    with member add-back, against the oracle on a mixed scene (stars and
    galaxies, a zero and a NaN S1 flux)."""
    import torch
    from uniphot_amd.pipeline import psf_bank
    from uniphot_amd.scene import SceneStars, island_mosaic, star_flux, subtract_stars
    sc = small_mixed_scene(seed=5, H=200, W=200, n_bands=2, n_src=24)
    plans, hosts, pad, d, w = mixed_classes(sc)
    rng = np.random.default_rng(1)
    flux = (sc.true_flux * rng.uniform(0.7, 1.3, sc.true_flux.shape)).astype(np.float32)
    stars = np.nonzero(sc.stype == 0)[0]
    flux[stars[0], 1] = 0.0
    flux[stars[1], 0] = np.nan
    nt = int(sc.psfs[0].n_terms)
    psf_of = psf_bank(sc.psfs)
    coefs, geo = psf_of(55)
    dd = torch.from_numpy(d).cuda()
    wd = torch.from_numpy(w).cuda()
    n_sub = subtract_stars(dd, pad, coefs, geo, 55, nt, sc.x, sc.y, sc.stype, flux)
    assert n_sub == stars.size
    psfs = [dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                 y_origin=p.y_origin, y_scale=p.y_scale) for p in sc.psfs]
    f64 = np.nan_to_num(flux.astype(np.float64), nan=0.0)
    scene = oisl.scene_star_image(d.shape, pad, psfs, sc.x, sc.y, sc.stype, f64)
    want = d.astype(np.float64) - scene
    got = dd.cpu().numpy()
    tol = 1e-5 * np.abs(scene).max()
    assert np.max(np.abs(got - want)) < tol, np.max(np.abs(got - want)) / tol
    # cutouts
    chunk = [(None, h, np.arange(p.n_islands)) for p, h in zip(plans, hosts)]
    plan_of = {id(h): p for p, h in zip(plans, hosts)}
    ss = SceneStars(sc.x, sc.y, sc.stype, star_flux(flux), psf_of, nt)
    dmos, wmos, moved = island_mosaic(chunk, plan_of, dd, wd, ss)
    dm, wm = dmos.cpu().numpy(), wmos.cpu().numpy()
    n_checked = 0
    for (plan, host), (_, h2, rows) in zip(zip(plans, hosts), moved):
        B, T = plan.bbox_class, plan.tier
        pt = [dict(q, coefs=ocore.pad_psf_coefs(q["coefs"], T)) for q in psfs]
        for r in rows:
            ox, oy = int(host["ox"][r]), int(host["oy"][r])
            mx, my = int(h2["ox"][r]), int(h2["oy"][r])
            mem = plan.midx[r][host["mmask"][r]]
            add = oisl.member_addback(pt, host["x0b"][r], host["y0b"][r], B, sc.x[mem],
                                      sc.y[mem], sc.stype[mem], f64[mem])
            cw = want[:, oy:oy + B, ox:ox + B] + add
            cg = dm[:, my:my + B, mx:mx + B]
            assert np.max(np.abs(cg - cw)) < tol, (plan.kind, r)
            assert np.array_equal(wm[:, my:my + B, mx:mx + B], w[:, oy:oy + B, ox:ox + B])
            n_checked += 1
    assert n_checked == sum(p.n_islands for p in plans)
    # the windows do not overlap in the mosaic
    cover = np.zeros(dm.shape[1:], np.int32)
    for (plan, _), (_, h2, rows) in zip(zip(plans, hosts), moved):
        B = plan.bbox_class
        for r in rows:
            cover[int(h2["oy"][r]):int(h2["oy"][r]) + B, int(h2["ox"][r]):int(h2["ox"][r]) + B] += 1
    assert cover.max() == 1


def test_autotuned_launch_shape_leaves_the_fit_unchanged(cuda, tmp_path, monkeypatch):
    """The autotuner times the island launch sequence at each candidate grid
    size, caches its choice under the workload key, and a second fit reuses
    it; the grid size changes no result (each (island, band) item runs on
    one workgroup whatever the grid)."""
    import dataclasses
    import json
    import torch
    from uniphot_amd import autotune
    from uniphot_amd.pipeline import DeviceField, FitConfig, fit_field
    monkeypatch.setenv("UPH_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    sc = small_mixed_scene(seed=3, n_src=24)
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    args = (sc.x, sc.y, sc.stype, sc.logre, sc.ee1, sc.ee2, sc.sersic_n)
    cfg = FitConfig(n_rows=1, n_cols=1)
    sync = torch.cuda.synchronize
    n_log = len(autotune.log())
    r0 = fit_field(fld, *args, cfg, sync=sync)
    tuned = autotune.log()[n_log:]
    assert tuned and not any(t["cached"] for t in tuned)
    cache = json.loads((tmp_path / "tune.json").read_text())
    for t in tuned:
        assert cache[t["key"]]["choice"] == t["choice"] in autotune.BLOCK_CANDIDATES
        assert sorted(cache[t["key"]]["ms"]) == sorted(map(str, autotune.BLOCK_CANDIDATES))
    fit_field(fld, *args, cfg, sync=sync)
    assert all(t["cached"] for t in autotune.log()[n_log + len(tuned):])
    outs = [r0] + [fit_field(fld, *args, dataclasses.replace(cfg, island_blocks=dict(fp32=b, fp64=b)),
                             sync=sync) for b in (256, 2048)]
    for r in outs[1:]:
        for k in ("flux", "x", "y", "logre", "ee1", "ee2", "island_nst", "island_flux"):
            assert np.array_equal(getattr(r, k), getattr(r0, k), equal_nan=True), k
