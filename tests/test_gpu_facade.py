"""The reference-shaped entry points on the device: solve_band_sparse with
GridBandInputs, render_source_stamps / render_star_stamps, run_one_grid,
run_class_chunk (compute flags, init_state resume) and island_fit, each
called with reference-shaped inputs and checked against the float64
oracle or against the batched path it wraps."""
import dataclasses

import numpy as np
import pytest

from _scenes import oracle_island, small_mixed_scene, small_star_scene, star_classes
from oracle import core as ocore
from oracle import island as oisl
from oracle import linear as olin

pytestmark = pytest.mark.gpu


def _psf_dict(p):
    return dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                y_origin=p.y_origin, y_scale=p.y_scale)


def test_solve_band_sparse_matches_oracle(cuda):
    from uniphot_amd import forced_phot_sparse as fps
    sc = small_star_scene(n_src=20)
    nb, H, W = sc.data.shape
    x0, y0, h, w = 10, 5, 140, 150
    inside = (sc.x >= x0) & (sc.x < x0 + w) & (sc.y >= y0) & (sc.y < y0 + h)
    xs, ys = sc.x[inside], sc.y[inside]
    nbr, msk = fps.build_neighbor_list(xs, ys, 55, 55)
    geo = (0.5 * (W - 1), 0.5 * W, 0.5 * (H - 1), 0.5 * H)
    for b in range(nb):
        sl = (b, slice(y0, y0 + h), slice(x0, x0 + w))
        inp = fps.GridBandInputs(sc.data[sl], np.where(sc.good[sl], sc.invvar[sl], 0),
                                 sc.good[sl], xs, ys, sc.psfs[b], h, w, x0, y0, *geo)
        r = fps.solve_band_sparse(inp, nbr, msk, reg=1e-9, bg_poly_order=1)
        psf = _psf_dict(sc.psfs[b])
        stamps = np.stack([ocore.member_stamp(psf, float(a), float(c), x0, y0, 0,
                                              renorm=True)[0] for a, c in zip(xs, ys)])
        ix, iy = fps.stamp_top_left(xs, ys, x0, y0, 55, 55)
        f, bgc, fv, bv = olin.solve_band(sc.data[sl].astype(np.float64),
                                         sc.invvar[sl].astype(np.float64), sc.good[sl], stamps,
                                         ix, iy, nbr, msk, olin.bg_basis(h, w, x0, y0, *geo,
                                                                         order=1))
        assert r.converged
        assert np.max(np.abs(r.fluxes - f) / np.maximum(np.abs(f), np.sqrt(fv))) < 1e-5
        assert np.allclose(r.flux_var, fv, rtol=1e-5)
        assert np.allclose(r.bg_var, bv, rtol=1e-4)


def test_render_source_stamps_match_oracle(cuda):
    from uniphot_amd.source_model import render_source_stamps, render_star_stamps
    sc = small_mixed_scene(types=(0, 1, 2, 3), n_range=(0.5, 4.0))
    psf = sc.psfs[0]
    got = render_source_stamps(psf, sc.x, sc.y, 7, 11, sc.stype, sc.logre, sc.ee1, sc.ee2,
                               sc.sersic_n, render_backend="fft").cpu().numpy()
    for i in range(sc.x.size):
        ref, _ = ocore.member_stamp(_psf_dict(psf), float(sc.x[i]), float(sc.y[i]), 7.0, 11.0,
                                    int(sc.stype[i]), float(sc.logre[i]), float(sc.ee1[i]),
                                    float(sc.ee2[i]), float(sc.sersic_n[i]))
        assert np.abs(got[i] - ref).max() < 2e-6 * np.abs(ref).max(), i
    star = sc.stype == 0
    only = render_star_stamps(psf, sc.x[star], sc.y[star], 7, 11).cpu().numpy()
    assert np.array_equal(only, got[star])      # the STAR rows, bit for bit
    # the default backend is "direct", which takes STAR and EXP only
    with pytest.raises(ValueError, match="DEV/SERSIC"):
        render_source_stamps(psf, sc.x, sc.y, 7, 11, sc.stype, sc.logre, sc.ee1, sc.ee2)


def test_run_one_grid_equals_the_field_pass_and_the_oracle(cuda):
    """One grid as a PerGridTask (window arrays, full-image positions) gives
    the fluxes of the same grid inside the field's linear pass, and the
    oracle's solve."""
    from uniphot_amd import forced_phot_sparse as fps
    from uniphot_amd.grid import assign_sources_to_grids, divide_image_into_grids
    from uniphot_amd.pipeline import DeviceField, FitConfig, linear_pass
    from uniphot_amd.runner import PerGridTask, run_one_grid
    sc = small_star_scene(n_src=20)
    nb, H, W = sc.data.shape
    grids = divide_image_into_grids(H, W, 2, 2, 30)
    assign_sources_to_grids(grids, sc.x, sc.y)
    g = grids[3]
    cfg = FitConfig(n_rows=2, n_cols=2, bg_order=1, wing_peeling=False, blank_response=False)
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    z = np.zeros(sc.x.size, np.float32)
    flux, var, bg = linear_pass(fld, grids, sc.x, sc.y, sc.stype, z, z, z, sc.sersic_n, cfg,
                                {gg.grid_id: 55 for gg in grids})
    fi = g.fit_indices
    own = np.isin(fi, g.source_indices)
    sl = (slice(g.y0_pad, g.y1_pad), slice(g.x0_pad, g.x1_pad))
    geo = (0.5 * (W - 1), 0.5 * W, 0.5 * (H - 1), 0.5 * H)
    task = PerGridTask(
        grid_id=g.grid_id, grid_row=g.grid_row, grid_col=g.grid_col, bandnames=["a", "b"],
        data_per_band=[sc.data[b][sl] for b in range(nb)],
        invvar_per_band=[sc.invvar[b][sl] for b in range(nb)],
        pix_good_per_band=[sc.good[b][sl] for b in range(nb)], psfs_per_band=list(sc.psfs),
        psf_shape=(55, 55), grid_x0=g.x0_pad, grid_y0=g.y0_pad, source_indices=fi,
        pos_x_init=sc.x[fi].astype(np.float64), pos_y_init=sc.y[fi].astype(np.float64),
        pos_err_pix=np.full(fi.size, 0.1), source_type=sc.stype[fi],
        logre_init=np.zeros(fi.size), ee1_init=np.zeros(fi.size), ee2_init=np.zeros(fi.size),
        pos_max_shift_pix=5.0, bg_x_origin=geo[0], bg_x_scale=geo[1], bg_y_origin=geo[2],
        bg_y_scale=geo[3], bg_poly_order=1, position_solver="island",
        owned_source_mask=own)
    r = run_one_grid(task)
    assert r.error is None, r.error
    assert np.array_equal(r.source_indices, g.source_indices)
    ref, refv = flux[g.source_indices], var[g.source_indices]
    sig = np.sqrt(refv)
    assert np.max(np.abs(r.fluxes - ref) / np.maximum(np.abs(ref), sig)) < 1e-6
    np.testing.assert_allclose(r.bg_coefs, bg[g.grid_id], rtol=1e-5, atol=1e-6)
    # and the oracle's solve of band 0 over the grid window
    xs, ys = sc.x[fi], sc.y[fi]
    nbr, msk = fps.build_neighbor_list(xs, ys, 55, 55)
    stamps = np.stack([ocore.member_stamp(_psf_dict(sc.psfs[0]), float(a), float(c),
                                          g.x0_pad, g.y0_pad, 0, renorm=True)[0]
                       for a, c in zip(xs, ys)])
    ix, iy = fps.stamp_top_left(xs, ys, g.x0_pad, g.y0_pad, 55, 55)
    f, _, fv, _ = olin.solve_band(sc.data[0][sl].astype(np.float64),
                                  sc.invvar[0][sl].astype(np.float64), sc.good[0][sl], stamps,
                                  ix, iy, nbr, msk, olin.bg_basis(g.pad_h, g.pad_w, g.x0_pad,
                                                                  g.y0_pad, *geo, order=1))
    assert np.max(np.abs(r.fluxes[:, 0] - f[own]) / np.maximum(np.abs(f[own]),
                                                                np.sqrt(fv[own]))) < 1e-5


def test_run_class_chunk_outputs_and_resume(cuda):
    """The reference's output keys and shapes with every compute flag; a fit
    stopped after 2 steps and resumed from its LM state (init_state) ends
    where the 30-step fit ends, bit for bit, on every island the 2-step run
    left unconverged."""
    import torch
    from uniphot_amd.island_solver import make_class_kernels, psf_tree_device, run_class_chunk
    sc = small_star_scene(seed=3, n_src=24)
    plans, hosts, pad, d, w = star_classes(sc)
    dd, ww = torch.from_numpy(d).to(cuda), torch.from_numpy(w).to(cuda)
    tree = psf_tree_device(sc.psfs, 55, cuda)
    plan, host = plans[0], hosts[0]
    kern = make_class_kernels(plan, 2, sc.psfs[0].n_terms)
    rows = np.arange(plan.n_islands)
    full = run_class_chunk(kern, tree, host, rows, None, 30, dd, ww, pad,
                           compute_param_errors=True, compute_source_metrics=True,
                           compute_flux_marg=True)
    n_c, N = rows.size, plan.nsrc_class
    assert full["p"].shape == (n_c, N, 6) and full["fluxes"].shape == (n_c, 2, N)
    assert full["perr"].shape == (n_c, N, 13) and np.isnan(full["perr"][..., 7:]).all()
    assert full["femarg"].shape == (n_c, 2, N) and full["smetrics"].shape == (n_c, N, 4)
    assert full["swin"].shape == (n_c, 2, N) and full["prej"].shape == (n_c, 6 * N)
    for k in ("diverged", "stall_dchi2", "railed", "n_same", "n_good", "chi2", "lam"):
        assert full[k].shape == (n_c,), k
    first = run_class_chunk(kern, tree, host, rows, None, 2, dd, ww, pad)
    todo = np.flatnonzero(~first["conv"])
    assert todo.size >= 1
    init = {k: first[k][todo] for k in ("p", "lam", "chi2", "consec", "nrej", "nst")}
    rest = run_class_chunk(kern, tree, host, rows[todo], None, 28, dd, ww, pad, init_state=init)
    for k in ("p", "fluxes", "ferr", "chi2", "lam", "nst", "conv", "stall"):
        assert np.array_equal(rest[k], full[k][todo]), k


def test_island_fit_rows_match_oracle(cuda):
    """island_fit on a star field: one row per source of every fitted
    island (plus giants), fluxes at the oracle's within 1e-5."""
    from uniphot_amd.island_solver import island_fit
    from uniphot_amd.islands import (ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES, bucket_islands,
                                     build_islands)
    sc = small_star_scene()
    n = sc.x.size
    iset = build_islands(sc.x, sc.y, sc.stype, sc.logre, sc.sersic_n, 27.5,
                         radius=np.full(n, 6.0))
    bk = bucket_islands(iset, ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES)
    s1 = np.full((n, 2), 1e4)
    res = island_fit(bandnames=["a", "b"], data_minus_bg=sc.data,
                     invvar=np.where(sc.good, sc.invvar, 0), psfs=sc.psfs, island_set=iset,
                     buckets=bk, pos_x=sc.x, pos_y=sc.y, source_type=sc.stype, logre=sc.logre,
                     ee1=sc.ee1, ee2=sc.ee2, sersic_n=sc.sersic_n, s1_flux=s1,
                     s1_flux_var=np.ones((n, 2)), compute_param_errors=True,
                     compute_source_metrics=True)
    assert np.array_equal(np.sort(res.source_idx), np.arange(n))
    assert np.isfinite(res.pos_x_err).all() and np.isfinite(res.rchisq).all()
    # oracle on the same classes (same padded stacks as star_classes)
    plans, hosts, pad, d, w = star_classes(sc)
    o = oisl.LMOptions()
    row_of = {int(s): i for i, s in enumerate(res.source_idx)}
    for plan, host in zip(plans, hosts):
        for r in range(plan.n_islands):
            isl = oracle_island(plan, host, r, d, w, sc.psfs)
            st = oisl.run_island(host["p0"][r], isl, o)
            f, e = oisl.final_fluxes(st["p"], isl, o)
            mm = host["mmask"][r]
            for k, s in zip(np.flatnonzero(mm), plan.midx[r][mm]):
                i = row_of[int(s)]
                assert res.n_steps[i] == st["nst"]
                rel = np.abs(res.fluxes[i] - f[:, k]) / np.maximum(np.abs(f[:, k]), e[:, k])
                assert rel.max() < 1e-5, (s, rel)
