"""Confusion flux error on the device (uph_confusion through
error_model.confusion_sigma) against the float64 oracle: windows of a mixed
scene with part of the sources left out of the model (unmodelled light), a
masked patch and a background plane; and the pass-2 wiring of fit_field."""
import types

import numpy as np
import pytest

from _scenes import small_mixed_scene
from oracle import core as ocore
from oracle import error_model as oem

pytestmark = pytest.mark.gpu


def test_confusion_sigma_matches_oracle(cuda):
    from uniphot_amd import forced_phot_sparse as fps
    from uniphot_amd.error_model import confusion_sigma
    from uniphot_amd.pipeline import DeviceField, FitConfig
    sc = small_mixed_scene(seed=4, n_src=30, H=220, W=220)
    good = sc.good.copy()
    good[:, 150:170, 30:60] = False                 # a masked patch
    fld = DeviceField(sc.data, sc.invvar, good, sc.psfs)
    T = fld.native
    n = sc.x.size
    modelled = np.arange(n)[np.arange(n) % 3 != 0]  # a third of the light unmodelled
    cfg = FitConfig()
    windows = [(0, 0, 220, 220), (40, 20, 150, 180)]
    geo = (0.5 * 219, 110.0, 0.5 * 219, 110.0)
    problems, owners, res = [], [], []
    rng = np.random.default_rng(1)
    for (x0, y0, h, w) in windows:
        inside = modelled[(sc.x[modelled] > x0 - 5) & (sc.x[modelled] < x0 + w + 5)
                          & (sc.y[modelled] > y0 - 5) & (sc.y[modelled] < y0 + h + 5)]
        g = types.SimpleNamespace(fit_indices=inside, source_indices=inside)
        for b in range(fld.nb):
            problems.append(fps.BandProblem(band=b, x0=x0, y0=y0, h=h, w=w,
                                            pos_x=sc.x[inside], pos_y=sc.y[inside],
                                            nbr_idx=np.zeros((inside.size, 1), np.int32)))
            owners.append((g, b))
            flux = (sc.true_flux[inside, b] * rng.uniform(0.97, 1.03, inside.size)).astype(np.float32)
            bg = np.array([0.3, -0.2, 0.1], np.float32)
            res.append(fps.SparseForcedPhotResult(flux, bg, None, None, True))
    x = sc.x.astype(np.float32).copy()
    x[modelled[5]] = np.nan                         # a source with no fitted position
    sig = confusion_sigma(fld, T, problems, owners, res, x, sc.y, sc.stype, sc.logre, sc.ee1,
                          sc.ee2, sc.sersic_n, cfg, geo)
    checked = 0
    for pr, (g, b), r, s in zip(problems, owners, res, sig):
        src = g.fit_indices
        ps = dict(coefs=sc.psfs[b].coefs.astype(np.float64), x_origin=sc.psfs[b].x_origin,
                  x_scale=sc.psfs[b].x_scale, y_origin=sc.psfs[b].y_origin,
                  y_scale=sc.psfs[b].y_scale)
        model = np.zeros((pr.h, pr.w))
        for k, i in enumerate(src):
            if not np.isfinite(x[i]):
                continue
            st, _ = ocore.member_stamp(ps, float(x[i]), float(sc.y[i]), pr.x0, pr.y0,
                                       int(sc.stype[i]), float(sc.logre[i]), float(sc.ee1[i]),
                                       float(sc.ee2[i]), float(sc.sersic_n[i]))
            ix, iy = ocore.stamp_top_left(float(x[i]), float(sc.y[i]), pr.x0, pr.y0, T)
            model += float(r.fluxes[k]) * ocore.place_window(st, ix, iy, pr.h, pr.w)
        yy, xx = np.mgrid[:pr.h, :pr.w]
        X = (xx + pr.x0 - geo[0]) / geo[1]
        Y = (yy + pr.y0 - geo[2]) / geo[3]
        model += r.bg_coefs[0] + r.bg_coefs[1] * X + r.bg_coefs[2] * Y
        win = (b, slice(pr.y0, pr.y0 + pr.h), slice(pr.x0, pr.x0 + pr.w))
        phi = ocore.eval_psf(ps, pr.x0 + 0.5 * pr.w, pr.y0 + 0.5 * pr.h).astype(np.float32)
        phi = phi / np.sum(phi, dtype=np.float32)
        xi = np.clip(np.rint(x[src] - np.float32(pr.x0)), 0, pr.w - 1)
        yi = np.clip(np.rint(sc.y[src].astype(np.float32) - np.float32(pr.y0)), 0, pr.h - 1)
        nf = ~np.isfinite(xi)
        ref = oem.confusion_sigma(sc.data[win], model, good[win], np.where(good, sc.invvar, 0)[win],
                                  phi, np.where(nf, 0, xi).astype(int),
                                  np.where(nf, 0, yi).astype(int),
                                  oem.template_variance(sc.stype[src], sc.logre[src]), nf)
        assert np.array_equal(np.isnan(s), np.isnan(ref))
        ok = np.isfinite(ref)
        np.testing.assert_allclose(s[ok], ref[ok], rtol=2e-3, atol=1e-3 * np.nanmax(ref))
        checked += int(ok.sum())
    assert checked > 40


def test_fit_field_carries_the_confusion_error(cuda):
    import torch
    from uniphot_amd.pipeline import DeviceField, FitConfig, fit_field
    sc = small_mixed_scene(seed=8, n_src=20)
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    res = fit_field(fld, sc.x, sc.y, sc.stype, sc.logre, sc.ee1, sc.ee2, sc.sersic_n,
                    FitConfig(n_rows=1, n_cols=1), sync=torch.cuda.synchronize)
    c = res.flux_err_conf
    assert c.shape == (sc.x.size, fld.nb)
    assert np.isfinite(c).mean() > 0.5 and np.all(c[np.isfinite(c)] >= 0)
    off = fit_field(fld, sc.x, sc.y, sc.stype, sc.logre, sc.ee1, sc.ee2, sc.sersic_n,
                    FitConfig(n_rows=1, n_cols=1, confusion=False), sync=torch.cuda.synchronize)
    assert off.flux_err_conf is None
    assert np.array_equal(off.flux, res.flux, equal_nan=True)
