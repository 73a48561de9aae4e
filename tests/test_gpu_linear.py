"""GPU parity of the stamp renderer and the grid-band linear solve against
the float64 oracle (oracle/linear.py)."""
import numpy as np
import pytest

from _scenes import small_star_scene
from oracle import core as ocore
from oracle import linear as olin

pytestmark = pytest.mark.gpu


def _psf_dict(p):
    return dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                y_origin=p.y_origin, y_scale=p.y_scale)


def test_render_stamps_match_oracle(cuda):
    import torch
    from uniphot_amd import _lib
    from uniphot_amd.psf_polynomial import psf_device_arrays
    sc = small_star_scene()
    c, g = psf_device_arrays(sc.psfs, 55)
    dev = cuda
    n = sc.x.size
    band = (np.arange(n) % 2).astype(np.int32)
    gx0 = np.full(n, 17.0, np.float32)
    gy0 = np.full(n, 3.0, np.float32)
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out = torch.empty((n, 55, 55), dtype=torch.float32, device=dev)
    L = _lib.lib()
    P = _lib.ptr
    tc, tg, tb, tx, ty, tgx, tgy = map(up, (c, g, band, sc.x, sc.y, gx0, gy0))
    _lib.check(L.uph_render_stamps(n, 55, c.shape[1], P(tc), P(tg), None, P(tb), P(tx), P(ty),
                                   P(tgx), P(tgy), None, None, None, None, None, 1, P(out),
                                   None, 0, _lib.stream_ptr()))
    got = out.cpu().numpy()
    for i in range(n):
        ref, _ = ocore.member_stamp(_psf_dict(sc.psfs[band[i]]), float(sc.x[i]), float(sc.y[i]),
                                    17.0, 3.0, 0, renorm=True)
        assert np.max(np.abs(got[i] - ref)) < 2e-7 * np.max(np.abs(ref)) * 10


def _check_band_solve(sc, x0, y0, h, w, dev, min_src=0):
    import torch
    from uniphot_amd import forced_phot_sparse as fps
    from uniphot_amd.psf_polynomial import psf_device_arrays
    nb, H, W = sc.data.shape
    c, g = psf_device_arrays(sc.psfs, 55)
    # one grid = a sub-window of the image with all sources whose centre is inside
    inside = (sc.x >= x0) & (sc.x < x0 + w) & (sc.y >= y0) & (sc.y < y0 + h)
    xs, ys = sc.x[inside], sc.y[inside]
    assert xs.size >= min_src
    nbr, msk = fps.build_neighbor_list(xs, ys, 55, 55)
    probs = [fps.BandProblem(b, x0, y0, h, w, xs, ys, nbr) for b in range(nb)]
    bg_geo = (0.5 * (W - 1), 0.5 * W, 0.5 * (H - 1), 0.5 * H)
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    res = fps.solve_problems(probs, up(c), up(g), 55, c.shape[1], up(sc.data),
                             up(np.where(sc.good, sc.invvar, 0).astype(np.float32)),
                             up(sc.good.astype(np.uint8)), bg_order=1, bg_geo=bg_geo)
    for b in range(nb):
        psf = _psf_dict(sc.psfs[b])
        stamps = np.stack([ocore.member_stamp(psf, float(xx), float(yy), x0, y0, 0, renorm=True)[0]
                           for xx, yy in zip(xs, ys)])
        ix, iy = fps.stamp_top_left(xs, ys, x0, y0, 55, 55)
        basis = olin.bg_basis(h, w, x0, y0, *bg_geo, order=1)
        f, bgc, fv, bv = olin.solve_band(sc.data[b, y0:y0 + h, x0:x0 + w].astype(np.float64),
                                         sc.invvar[b, y0:y0 + h, x0:x0 + w].astype(np.float64),
                                         sc.good[b, y0:y0 + h, x0:x0 + w], stamps, ix, iy, nbr,
                                         msk, basis)
        r = res[b]
        sig = np.sqrt(fv)
        assert np.max(np.abs(r.fluxes - f) / np.maximum(np.abs(f), sig)) < 1e-5
        assert np.allclose(r.flux_var, fv, rtol=1e-5)
        assert np.allclose(r.bg_coefs, bgc, rtol=1e-4, atol=1e-4 * np.sqrt(bv).max())
        assert r.converged


def test_band_solve_matches_oracle(cuda):
    _check_band_solve(small_star_scene(n_src=20), 10, 5, 140, 150, cuda)


def test_band_solve_large_system_matches_oracle(cuda):
    """170 sources and a plane background in one window: a 173 x 173 system,
    whose LU runs its first 36 steps in global memory and the rest with the
    trailing block in LDS."""
    sc = small_star_scene(seed=11, H=300, W=300, n_bands=1, n_src=170, clustered=False)
    _check_band_solve(sc, 0, 0, 300, 300, cuda, min_src=138)


def test_wing_peeling_matches_oracle(cuda):
    """Wing peeling of one grid at the native stamp: the device path (S1
    solve, wing images of the big sources from the S1 fluxes, solve of the
    wing subtracted by linearity) against the oracle's S1 solve and its
    direct solve of data - wing."""
    import torch
    from _scenes import small_mixed_scene
    from uniphot_amd import forced_phot_sparse as fps
    from uniphot_amd.grid import GridDef
    from uniphot_amd.pipeline import DeviceField, FitConfig, wing_peel, wing_sources
    sc = small_mixed_scene(seed=9, n_bands=2, n_src=12, order=1, types=(0, 1, 2))
    st, lr, sn = sc.stype.copy(), sc.logre.copy(), sc.sersic_n.copy()
    dev_i = int(np.where(st == 2)[0][0])          # a de Vaucouleurs galaxy made big
    lr[dev_i] = np.float32(np.log(6.0))
    nb, H, W = sc.data.shape
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    g = GridDef(0, 0, 0, 20, 180, 20, 180, 5, 195, 5, 195, 15)
    inside = (sc.x >= g.x0_pad) & (sc.x < g.x1_pad) & (sc.y >= g.y0_pad) & (sc.y < g.y1_pad)
    fi = np.where(inside)[0]
    g.source_indices = fi
    g.halo_source_indices = np.zeros(0, np.int64)
    assert dev_i in fi
    big = wing_sources(st[fi], lr[fi], sn[fi], 55, 64)
    assert big.size >= 1
    xs, ys = sc.x[fi], sc.y[fi]
    nbr, msk = fps.build_neighbor_list(xs, ys, 55, 55)
    probs = [fps.BandProblem(b, g.x0_pad, g.y0_pad, g.pad_h, g.pad_w, xs, ys, nbr, stype=st[fi],
                             logre=lr[fi], ee1=sc.ee1[fi], ee2=sc.ee2[fi], sersic_n=sn[fi])
             for b in range(nb)]
    bg_geo = (0.5 * (W - 1), 0.5 * W, 0.5 * (H - 1), 0.5 * H)
    c55, g55 = fld.psf(55)
    s1 = fps.solve_problems(probs, c55, g55, 55, fld.n_terms, fld.data, fld.invvar, fld.good,
                            bg_order=1, bg_geo=bg_geo)
    cfg = FitConfig()
    fin, _ = wing_peel(fld, 55, probs, [(g, b) for b in range(nb)], s1, sc.x, sc.y, st, lr,
                       sc.ee1, sc.ee2, sn, cfg, bg_geo)
    h, w = g.pad_h, g.pad_w
    basis = olin.bg_basis(h, w, g.x0_pad, g.y0_pad, *bg_geo, order=1)
    ix, iy = fps.stamp_top_left(xs, ys, g.x0_pad, g.y0_pad, 55, 55)
    for b in range(nb):
        psf = _psf_dict(sc.psfs[b])
        stamps = olin.render_grid_stamps(psf, xs, ys, g.x0_pad, g.y0_pad, st[fi], lr[fi],
                                         sc.ee1[fi], sc.ee2[fi], sn[fi])
        sl = (b, slice(g.y0_pad, g.y1_pad), slice(g.x0_pad, g.x1_pad))
        d, iv, gd = (sc.data[sl].astype(np.float64), sc.invvar[sl].astype(np.float64),
                     sc.good[sl])
        f1, _, _, _ = olin.solve_band(d, iv, gd, stamps, ix, iy, nbr, msk, basis)
        bi = fi[big]
        wing = olin.wing_image(psf, sc.x[bi], sc.y[bi], g.x0_pad, g.y0_pad, st[bi], lr[bi],
                               sc.ee1[bi], sc.ee2[bi], sn[bi], f1[big], h, w, 221, 55)
        assert np.abs(wing).max() > 0
        f, bgc, fv, bv = olin.solve_band(d - wing, iv, gd, stamps, ix, iy, nbr, msk, basis)
        r = fin[b]
        sig = np.sqrt(fv)
        err = np.max(np.abs(r.fluxes - f) / np.maximum(np.abs(f), sig))
        moved = np.max(np.abs(f - f1) / np.maximum(np.abs(f), sig))
        print("band", b, "wing moved fluxes by", moved, "device vs oracle", err, flush=True)
        assert moved > 1e-4
        assert err < 1e-4
        assert np.allclose(r.flux_var, fv, rtol=1e-5)


def test_blank_response_matches_oracle(cuda):
    """Pass-2 blank-response median of each band of one grid: device
    (uph_blank_response) against the oracle (scipy fftconvolve +
    nanmedian) on the same final fluxes and background; the STAR fluxes
    move by exactly b_hat."""
    from uniphot_amd import forced_phot_sparse as fps
    from uniphot_amd.grid import GridDef
    from uniphot_amd.pipeline import DeviceField, FitConfig, blank_response_correct, _pad_center
    from uniphot_amd.psf_polynomial import evaluate_psf_at_np
    sc = small_star_scene(n_src=20)
    nb, H, W = sc.data.shape
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    g = GridDef(0, 0, 0, 15, 145, 10, 150, 5, 155, 2, 158, 10)
    inside = (sc.x >= g.x0_pad) & (sc.x < g.x1_pad) & (sc.y >= g.y0_pad) & (sc.y < g.y1_pad)
    fi = np.where(inside)[0]
    g.source_indices = fi
    g.halo_source_indices = np.zeros(0, np.int64)
    xs, ys = sc.x[fi], sc.y[fi]
    nbr, msk = fps.build_neighbor_list(xs, ys, 55, 55)
    probs = [fps.BandProblem(b, g.x0_pad, g.y0_pad, g.pad_h, g.pad_w, xs, ys, nbr)
             for b in range(nb)]
    bg_geo = (0.5 * (W - 1), 0.5 * W, 0.5 * (H - 1), 0.5 * H)
    c55, g55 = fld.psf(55)
    res = fps.solve_problems(probs, c55, g55, 55, fld.n_terms, fld.data, fld.invvar, fld.good,
                             bg_order=1, bg_geo=bg_geo)
    st = sc.stype.astype(np.int32)
    zero = np.zeros_like(sc.x)
    corr, b_hat = blank_response_correct(fld, 55, probs, [(g, b) for b in range(nb)], res,
                                         sc.x, sc.y, st, zero, zero + 1, zero, zero + 1,
                                         FitConfig(), bg_geo)
    h, w = g.pad_h, g.pad_w
    basis = olin.bg_basis(h, w, g.x0_pad, g.y0_pad, *bg_geo, order=1)
    ix, iy = fps.stamp_top_left(xs, ys, g.x0_pad, g.y0_pad, 55, 55)
    for b in range(nb):
        psf = _psf_dict(sc.psfs[b])
        stamps = np.stack([ocore.member_stamp(psf, float(xx), float(yy), g.x0_pad, g.y0_pad, 0,
                                              renorm=False)[0] for xx, yy in zip(xs, ys)])
        phi = np.clip(evaluate_psf_at_np(sc.psfs[b], g.x0_pad + 0.5 * w,
                                         g.y0_pad + 0.5 * h).astype(np.float64), 0.0, None)
        phi = _pad_center(phi, 55)
        phi /= phi.sum()
        sl = (b, slice(g.y0_pad, g.y1_pad), slice(g.x0_pad, g.x1_pad))
        ref = olin.blank_response(sc.data[sl].astype(np.float64), sc.invvar[sl].astype(np.float64),
                                  sc.good[sl], stamps, ix, iy, res[b].fluxes.astype(np.float64),
                                  res[b].bg_coefs, basis, phi)
        sig = np.sqrt(np.median(res[b].flux_var))
        print("band", b, "b_hat device", b_hat[b], "oracle", ref, "flux sigma", sig, flush=True)
        assert abs(b_hat[b] - ref) < 1e-3 * sig + 1e-5 * abs(ref)
        np.testing.assert_allclose(res[b].fluxes - corr[b].fluxes, np.float32(b_hat[b]),
                                   rtol=1e-6, atol=2e-7 * np.abs(res[b].fluxes).max())


def test_background_planes_subtracted_as_the_host_formula(cuda):
    """uph_subtract_bg_planes against the plane each grid core gets in the
    reference's stack build (island_stage.py:1731-1751): basis X^m Y^n in
    float64 normalised coordinates rounded to float32, times the grid's
    coefficients, subtracted on the core only (halo and border untouched)."""
    import torch
    from uniphot_amd.grid import divide_image_into_grids
    from uniphot_amd.pipeline import DeviceField, padded_stacks
    from uniphot_amd.psf_polynomial import terms_for_order
    sc = small_star_scene(n_src=6, H=150, W=170, n_bands=3)
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    nb, H, W = sc.data.shape
    grids = divide_image_into_grids(H, W, 2, 3, 10)
    rng = np.random.default_rng(2)
    for order in (0, 1, 2):
        M = (order + 1) * (order + 2) // 2
        bg = rng.normal(0, 5, (len(grids), nb, M)).astype(np.float32)
        pad = 12
        d, w = padded_stacks(fld, grids, bg, order, pad)
        want = np.zeros((nb, H + 2 * pad, W + 2 * pad), np.float64)
        want[:, pad:pad + H, pad:pad + W] = sc.data
        xo, xs, yo, ys = 0.5 * (W - 1), 0.5 * W, 0.5 * (H - 1), 0.5 * H
        for g in grids:
            h, ww = g.y1_core - g.y0_core, g.x1_core - g.x0_core
            X = ((np.arange(ww) + g.x0_core - xo) / xs)[None, :]
            Y = ((np.arange(h) + g.y0_core - yo) / ys)[:, None]
            basis = np.stack([np.broadcast_to(X ** m * Y ** n, (h, ww)).astype(np.float32)
                              for (m, n) in terms_for_order(order)], axis=-1)
            plane = np.einsum("hwm,bm->bhw", basis.astype(np.float64),
                              bg[g.grid_id].astype(np.float64))
            want[:, pad + g.y0_core:pad + g.y1_core, pad + g.x0_core:pad + g.x1_core] -= plane
        got = d.cpu().numpy()
        assert np.max(np.abs(got - want)) < 1e-5 * (np.abs(want).max() + 1.0), order
        assert np.array_equal(w.cpu().numpy()[:, pad:pad + H, pad:pad + W], sc.invvar)
