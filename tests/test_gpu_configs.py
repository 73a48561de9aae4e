"""The benchmark workloads against the float64 oracle on the GPU:
  * the 1-band 256^2 field of 40 stars and 10 EXP galaxies in full (every
    island's LM fit and every grid-band pass-2 solve);
  * island samples of the 20-band clustered field and of its 40-band
    scaling (sub-fields of the same generator at the same source density,
    1.19e-3 px^-2, so the island statistics are those of the full fields);
  * an island sample of the 10-band power-law cutout batch;
  * the serving stream's tiles at its own sizes, alone and in the mix.
Bars: per source max over bands of |f - f_o| / max(|f_o|, sigma_o) <= 1e-5,
positions within 1e-4 px, the oracle's step count and flags (islands that
stall: the flux bar at the device's own stopping point, see _assert_islands).
The field samples are fitted by the oracle on the cutouts the device built
(its float32 scene subtraction and member add-back); the 1-band field runs
the whole preparation in float64 as well."""
import time

import numpy as np
import pytest

from oracle import core as ocore
from oracle import island as oisl
from oracle import linear as olin

pytestmark = pytest.mark.gpu


def _psf_dicts(psfs):
    return [dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                 y_origin=p.y_origin, y_scale=p.y_scale) for p in psfs]


def _field_view(res):
    """Per-source island results of a field fit; p in the fit's parameter
    layout (x, y, logre, ee1, ee2, u of the Sersic index)."""
    from uniphot_amd.source_model import sersic_n_to_u
    p = np.stack([res.x, res.y, res.logre, res.ee1, res.ee2,
                  sersic_n_to_u(np.asarray(res.sersic_n, np.float64))], axis=1)
    return dict(flux=res.island_flux, p=p.astype(np.float32), nst=res.island_nst,
                conv=res.island_conv, stall=res.island_stall, fp64=res.island_fp64)


def _set_view(fit):
    """Per-source island results of an island-set fit."""
    return dict(flux=fit.flux, p=fit.p.astype(np.float32), nst=fit.nst, conv=fit.conv,
                stall=fit.stall, fp64=fit.fp64_refit)


def _oracle_islands(v, classes, psfs, cut_of, items, budget_s=None, reg=1e-9):
    """Oracle LM fit of class rows ``items`` [(class index, row)] on the
    cutouts ``cut_of(class index, row)`` -> (data, weights); compared with the
    device fit ``v`` (a per-source view).  Returns per island the max flux
    rel-err, max position error (px), same steps and flags, whether it
    stalled, and the flux rel-err of the oracle's flux solve at the device's
    parameters."""
    o = oisl.LMOptions(reg=reg)
    out = []
    t0 = time.perf_counter()
    for pi, r in items:
        if budget_s is not None and out and time.perf_counter() - t0 > budget_s:
            break
        plan, host = classes[pi]
        T = plan.tier
        ps = [dict(q, coefs=ocore.pad_psf_coefs(q["coefs"], T)) for q in psfs]
        mm = host["mmask"][r]
        mem = plan.midx[r][mm]
        dcut, wcut = cut_of(pi, r)
        isl = oisl.Island(dcut, wcut, host["x0b"][r], host["y0b"][r], host["st"][r],
                          host["sn"][r], host["fn_flag"][r], mm, host["keep"][r],
                          host["lower"][r].ravel(), host["upper"][r].ravel(), ps, T)
        st = oisl.run_island(host["p0"][r], isl, o)
        f, e = oisl.final_fluxes(st["p"], isl, o)
        fo, eo = f[:, mm].T, e[:, mm].T
        po = st["p"].reshape(-1, 6)[mm]
        fd = v["flux"][mem].astype(np.float64)
        rel = np.abs(fd - fo) / np.maximum(np.abs(fo), eo)
        pos = np.hypot(v["p"][mem, 0] - po[:, 0], v["p"][mem, 1] - po[:, 1])
        same = (np.all(v["nst"][mem] == st["nst"])
                and np.all(v["conv"][mem] == bool(st["conv"]))
                and np.all(v["stall"][mem] == bool(st["stall"])))
        # the flux solve at the device's parameters (the device's own stopping point)
        pd = st["p"].reshape(-1, 6).copy()
        pd[mm] = v["p"][mem]
        fa, ea = oisl.final_fluxes(pd.ravel().astype(np.float32), isl, o)
        rel_at = np.abs(fd - fa[:, mm].T) / np.maximum(np.abs(fa[:, mm].T), ea[:, mm].T)
        out.append(dict(cls=(plan.bbox_class, plan.nsrc_class, T), row=int(r), n=int(mm.sum()),
                        rel=float(np.max(rel)), rel_at=float(np.max(rel_at)),
                        pos=float(np.max(pos)), same=bool(same), stalled=bool(st["stall"]),
                        fp64=bool(np.any(v["fp64"][mem]))))
    return out


def _assert_islands(rows, what):
    """Every island: the oracle's step count and flags, positions within
    1e-4 px, and fluxes within 1e-5 of the oracle's flux solve at the
    device's parameters.  Islands that converged: fluxes within 1e-5 of the
    oracle's own fit.  A stalled island stops on a plateau of chi2 where
    positions that agree to 1e-4 px can still move faint members' fluxes by
    more than 1e-5; its end-to-end flux difference is reported."""
    bad = [r for r in rows if not (r["pos"] <= 1e-4 and r["same"] and r["rel_at"] <= 1e-5
                                   and (r["stalled"] or r["rel"] <= 1e-5))]
    st = [r for r in rows if r["stalled"]]
    print(what, len(rows), "islands,", sum(r["n"] for r in rows), "sources,",
          sum(r["fp64"] for r in rows), "on the float64 kernels,", len(st), "stalled; worst rel",
          max(r["rel"] for r in rows if not r["stalled"]) if len(st) < len(rows) else None,
          "at device parameters", max(r["rel_at"] for r in rows), "pos",
          max(r["pos"] for r in rows), "stalled end-to-end",
          [round(r["rel"], 8) for r in st], flush=True)
    assert not bad, (what, bad)


def _device_cutouts(art, items):
    """The island cutouts the device fitted, rebuilt by the fit's own
    mosaic code (window of the scene-subtracted stack plus the member
    add-back at the class tier), for class rows ``items``."""
    from uniphot_amd.scene import island_mosaic
    by_cls = {}
    for pi, r in items:
        by_cls.setdefault(pi, []).append(int(r))
    keys = sorted(by_cls)
    chunk = [(None, art["classes"][pi][1], np.asarray(by_cls[pi])) for pi in keys]
    plan_of = {id(h): p for p, h in art["classes"]}
    dmos, wmos, moved = island_mosaic(chunk, plan_of, art["dstack"], art["wstack"],
                                      art["scene"])
    dm = dmos.cpu().numpy().astype(np.float64)
    wm = wmos.cpu().numpy()
    out = {}
    for pi, (_, h2, rows) in zip(keys, moved):
        B = art["classes"][pi][0].bbox_class
        for r in rows:
            mx, my = int(h2["ox"][r]), int(h2["oy"][r])
            out[(pi, int(r))] = (dm[:, my:my + B, mx:mx + B], wm[:, my:my + B, mx:mx + B])
    return out


def _sample_items(classes, rng, max_members=8):
    """Class rows of at most ``max_members`` members (the oracle's cost
    grows with members and tier) in a random order, one row of every class
    first."""
    first, rest = [], []
    for pi, (plan, host) in enumerate(classes):
        rows = [r for r in rng.permutation(plan.n_islands)
                if int(host["mmask"][r].sum()) <= max_members]
        if rows:
            first.append((pi, rows[0]))
            rest += [(pi, r) for r in rows[1:]]
    rng.shuffle(first)
    rng.shuffle(rest)
    return first + rest


def _fit_field(sc, **cfg_kw):
    import torch
    from uniphot_amd.pipeline import DeviceField, FitConfig, fit_field
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    cfg = FitConfig(**cfg_kw)
    res = fit_field(fld, sc.x, sc.y, sc.stype, sc.logre, sc.ee1, sc.ee2, sc.sersic_n, cfg,
                    sync=torch.cuda.synchronize, keep_artifacts=True)
    return fld, cfg, res


def test_c1_field_in_full_against_the_oracle(cuda):
    """1 band, 256^2, 40 STAR + 10 EXP, 8 px linking, 128^2 background
    tiles: every island's fit and every pass-2 grid solve."""
    from uniphot_amd import forced_phot_sparse as fps
    from scenes import field as synth
    from uniphot_amd.pipeline import _pad_center, wing_sources
    from uniphot_amd.psf_polynomial import evaluate_psf_at_np
    sc = synth.scene_c1()
    print("c1 scene", sc.x.size, "sources", flush=True)
    fld, cfg, res = _fit_field(sc, n_rows=2, n_cols=2, link_radius=4.0)
    art = res.record["_artifacts"]
    psfs = _psf_dicts(sc.psfs)
    s1 = np.nan_to_num(res.s1_flux.astype(np.float64), nan=0.0)
    # the scene subtraction restated by the oracle on the raw stack
    draw = art["dstack_raw"].cpu().numpy().astype(np.float64)
    d = draw - oisl.scene_star_image(draw.shape, art["pad"], psfs, sc.x, sc.y, sc.stype, s1)
    w = art["wstack"].cpu().numpy()
    items = [(pi, r) for pi, (plan, _) in enumerate(art["classes"])
             for r in range(plan.n_islands)]

    def cut_of(pi, r):
        plan, host = art["classes"][pi]
        B, T = plan.bbox_class, plan.tier
        ox, oy = int(host["ox"][r]), int(host["oy"][r])
        mem = plan.midx[r][host["mmask"][r]]
        ps = [dict(q, coefs=ocore.pad_psf_coefs(q["coefs"], T)) for q in psfs]
        add = oisl.member_addback(ps, host["x0b"][r], host["y0b"][r], B, sc.x[mem], sc.y[mem],
                                  sc.stype[mem], s1[mem])
        return d[:, oy:oy + B, ox:ox + B] + add, w[:, oy:oy + B, ox:ox + B]
    rows = _oracle_islands(_field_view(res), art["classes"], psfs, cut_of, items, reg=cfg.reg)
    assert len(rows) == res.record["n_fitted_islands"]
    _assert_islands(rows, "c1")
    # pass 2 at the fitted parameters, every grid and band
    assert wing_sources(sc.stype, res.logre, res.sersic_n, fld.native, 10**9).size == 0
    bg_geo = (0.5 * 255, 128.0, 0.5 * 255, 128.0)
    worst = 0.0
    for g in art["grids"]:
        fi = g.fit_indices
        T = art["gtier"][g.grid_id]
        ps = dict(psfs[0], coefs=ocore.pad_psf_coefs(psfs[0]["coefs"], T))
        xs, ys = res.x[fi], res.y[fi]
        st = sc.stype[fi]
        lr, e1, e2, sn = res.logre[fi], res.ee1[fi], res.ee2[fi], res.sersic_n[fi]
        nbr, msk = fps.build_neighbor_list(xs, ys, T, T)
        ix, iy = fps.stamp_top_left(xs, ys, g.x0_pad, g.y0_pad, T, T)
        basis = olin.bg_basis(g.pad_h, g.pad_w, g.x0_pad, g.y0_pad, *bg_geo, order=1)
        stamps = olin.render_grid_stamps(ps, xs, ys, g.x0_pad, g.y0_pad, st, lr, e1, e2, sn)
        sl = (0, slice(g.y0_pad, g.y1_pad), slice(g.x0_pad, g.x1_pad))
        dat = sc.data[sl].astype(np.float64)
        iv = sc.invvar[sl].astype(np.float64)
        f, bgc, fv, _ = olin.solve_band(dat, iv, sc.good[sl], stamps, ix, iy, nbr, msk, basis)
        own = g.source_indices.size
        fo = f[:own].copy()
        if T == fld.native:
            raw = np.stack([ocore.member_stamp(ps, float(xs[i]), float(ys[i]), g.x0_pad,
                                               g.y0_pad, int(st[i]), float(lr[i]),
                                               float(e1[i]), float(e2[i]), float(sn[i]))[0]
                            for i in range(fi.size)])
            phi = np.clip(evaluate_psf_at_np(sc.psfs[0], g.x0_pad + 0.5 * g.pad_w,
                                             g.y0_pad + 0.5 * g.pad_h).astype(np.float64),
                          0.0, None)
            phi = _pad_center(phi, T)
            phi /= phi.sum()
            bh = olin.blank_response(dat, iv, sc.good[sl], raw, ix, iy, f, bgc, basis, phi)
            if np.isfinite(bh):
                fo[st[:own] == 0] -= bh
        sig = np.sqrt(fv[:own])
        err = np.abs(res.flux[g.source_indices, 0] - fo)
        worst = max(worst, float(np.max(err / np.maximum(np.abs(fo), sig))))
        assert np.all(err < 1e-5 * np.maximum(np.abs(fo), sig) + 1e-3 * sig), g.grid_id
    print("c1 pass 2 worst rel", worst, flush=True)


@pytest.mark.parametrize("n_bands,side,n_src,budget", [(20, 1024, 1250, 45.0),
                                                       (40, 768, 703, 45.0)],
                         ids=["c2-20band", "c4-40band"])
def test_field_island_sample_against_the_oracle(cuda, n_bands, side, n_src, budget):
    """Clustered field (70/15/10/5 % STAR/EXP/DEV/SERSIC, n in [0.5, 6],
    fluxes 1e2-1e6, quadratic PSF variation, 256^2 background tiles) at the
    workloads' density; one island of every class first, then random ones,
    within a time budget of the oracle."""
    from scenes import field as synth
    sc = synth.scene_c2(H=side, W=side, n_bands=n_bands, n_src=n_src)
    print(f"field {side}^2 x {n_bands} bands, {sc.x.size} sources", flush=True)
    fld, cfg, res = _fit_field(sc, n_rows=side // 256, n_cols=side // 256)
    art = res.record["_artifacts"]
    print("islands", res.record["n_islands"], "float64", res.record["n_fp64_refit_islands"],
          flush=True)
    rng = np.random.default_rng(n_bands)
    items = _sample_items(art["classes"], rng)[:80]
    cuts = _device_cutouts(art, items)
    rows = _oracle_islands(_field_view(res), art["classes"], _psf_dicts(sc.psfs),
                           lambda pi, r: cuts[(pi, r)], items, budget_s=budget, reg=cfg.reg)
    assert len(rows) >= 3
    _assert_islands(rows, f"{n_bands}-band field")


def test_cutout_island_sample_against_the_oracle(cuda):
    """Power-law island cutouts (k = 1-64, 15-95 px, 10 bands) fitted in one
    batch; a sample of the islands against the oracle."""
    import torch
    from uniphot_amd.pipeline import FitConfig, fit_cutouts
    from scenes.cutouts import scene_cutouts
    cut = scene_cutouts(seed=9, n_isl=400, n_bands=10, device="cuda", mosaic_w=4096)
    res = fit_cutouts(cut, FitConfig(), sync=torch.cuda.synchronize)
    d = cut.dstack.cpu().numpy().astype(np.float64)
    w = cut.wstack.cpu().numpy()
    rng = np.random.default_rng(3)

    def cut_of(pi, r):
        plan, host = res.classes[pi]
        B = plan.bbox_class
        ox, oy = int(host["ox"][r]), int(host["oy"][r])
        return d[:, oy:oy + B, ox:ox + B], w[:, oy:oy + B, ox:ox + B]
    rows = _oracle_islands(_set_view(res), res.classes, _psf_dicts(cut.psfs), cut_of,
                           _sample_items(res.classes, rng), budget_s=45.0)
    assert len(rows) >= 3
    _assert_islands(rows, "cutouts")


def test_serving_stream_tiles_alone_equal_in_the_mix(cuda):
    """The serving stream's own request draw (sides 256-4096, 1-40 bands,
    1e-3 sources/px^2), first eight requests: each tile's result alone equals
    its result among the others in flight."""
    from uniphot_amd.serving import TileServer, latency_summary
    from scenes.tiles import tile_stream
    reqs, bank = tile_stream(n_req=8, rate=50.0)
    print("tiles", [(r.side, r.n_bands, r.n_sources) for r in reqs], flush=True)
    alone = [TileServer(bank, max_inflight=1).serve([r], realtime=False)[0] for r in reqs]
    mix = TileServer(bank, hbm_budget=200e9, max_inflight=4).serve(reqs, realtime=False)
    print(latency_summary(mix), flush=True)
    for a, m in zip(alone, mix):
        assert a.rid == m.rid
        for k in ("flux", "flux_err", "x", "y", "logre", "island_nst", "island_conv"):
            assert np.array_equal(getattr(a.result, k), getattr(m.result, k), equal_nan=True), \
                (a.rid, k)
