"""Correlated-noise factor on the device (uph_fcorr_*) against the oracle
restatement (oracle/fcorr.py): raw-sky maps and residual windows."""
import numpy as np
import pytest

from oracle import fcorr as of

pytestmark = pytest.mark.gpu


def _smoothed(H, W, h, rng):
    z = rng.standard_normal((H, W))
    k = np.zeros((H, W))
    k[:h.shape[0], :h.shape[1]] = h
    k = np.roll(k, (-(h.shape[0] // 2), -(h.shape[1] // 2)), (0, 1))
    return np.real(np.fft.ifft2(np.fft.fft2(z) * np.fft.fft2(k)))


def _psf(T, s):
    y, x = np.mgrid[:T, :T] - T // 2
    p = np.exp(-0.5 * (x * x + y * y) / s ** 2)
    return (p / p.sum()).astype(np.float32)


def _sky_maps(rng):
    """Four 1024 x 960 float32 sky maps: kernel-correlated noise, PSF-shaped
    noise, white noise with sources and a masked patch, and one mostly
    masked (fallback)."""
    H, W = 1024, 960
    b = np.array([0.2, 0.6, 0.2])
    maps, good = [], []
    maps.append(50.0 + 3.0 * _smoothed(H, W, np.outer(b, b), rng))
    maps.append(-2.0 + _smoothed(H, W, _psf(21, 2.2).astype(np.float64), rng))
    m = 10.0 + rng.standard_normal((H, W))
    for _ in range(40):
        y, x = rng.integers(20, H - 20), rng.integers(20, W - 20)
        m[y - 2:y + 3, x - 2:x + 3] += rng.uniform(10, 1000)
    m[100:140, 300:420] = np.nan
    maps.append(m)
    maps.append(rng.standard_normal((H, W)))
    g = [np.ones((H, W), bool) for _ in range(4)]
    g[2][600:650, 10:200] = False
    g[3][:, 150:] = False
    return [x.astype(np.float32) for x in maps], g


def test_raw_sky_factor_matches_oracle():
    import torch
    from uniphot_amd import fcorr as uf
    rng = np.random.default_rng(11)
    maps, good = _sky_maps(rng)
    H, W = maps[0].shape
    T = 55
    phis = np.stack([_psf(T, 1.3), _psf(T, 2.2), _psf(T, 1.6), _psf(T, 1.6)]).astype(np.float64)
    src = torch.from_numpy(np.stack(maps).ravel()).cuda()
    gd = torch.from_numpy(np.stack(good).astype(np.uint8).ravel()).cuda()
    n = len(maps)
    res = uf._estimate(src, np.arange(n, dtype=np.int64) * H * W, np.full(n, W, np.int32),
                       np.tile([H, W], (n, 1)), phis, good_u8=gd, residual=False)
    for k in range(n):
        print(f"raw-sky map {k}", flush=True)
        F, dg = res[k]
        Fo, do = of.fcorr_sky(maps[k].astype(np.float64), ~good[k], phis[k])
        assert dg["fallback"] == do["fallback"], (k, dg, do)
        if do["fallback"] == "too few finite sky pixels":
            continue
        assert dg["med"] == do["med"] and dg["mad"] == do["mad"]        # exact selection
        assert dg["n_clean"] == do["n_clean"]
        if "F_acf" in do:
            assert dg["F_acf"] == pytest.approx(do["F_acf"], rel=1e-10)
            assert dg["rho_1px"] == pytest.approx(do["rho_1px"], rel=1e-9, abs=1e-12)
            assert dg.get("acf_shape") == do.get("acf_shape")
        assert F == pytest.approx(Fo, rel=1e-10)
    assert res[0][0] > 1.5 and res[1][1]["acf_shape"] == "confusion"


def test_residual_window_factor_matches_oracle():
    import torch
    from uniphot_amd import fcorr as uf
    rng = np.random.default_rng(12)
    b = np.array([0.25, 0.5, 0.25])
    shapes = [(1100, 1030), (300, 300), (1030, 1100)]
    max_pix = max(h * w for h, w in shapes)
    work = np.zeros(6 * max_pix * len(shapes), np.float32)
    rs, gs = [], []
    for k, (h, w) in enumerate(shapes):
        r = _smoothed(h, w, np.outer(b, b), rng) * np.linspace(1.0, 1.3, w)[None, :]
        g = np.ones((h, w), bool)
        g[50:90, 400:500] = False
        for _ in range(30):
            y, x = rng.integers(10, h - 10), rng.integers(10, w - 10)
            r[y - 3:y + 4, x - 3:x + 4] += 40.0
        r = np.where(g, r, 0.0).astype(np.float32)
        work[k * 6 * max_pix:k * 6 * max_pix + h * w] = r.ravel()
        work[k * 6 * max_pix + max_pix:k * 6 * max_pix + max_pix + h * w] = g.ravel()
        rs.append(r)
        gs.append(g)
    T = 55
    phi = np.stack([_psf(T, 1.5)] * len(shapes))
    win = np.array([[0, 0, 0, h, w] for h, w in shapes], np.int32)
    out = uf.fcorr_residual_windows(torch.from_numpy(work).cuda(), max_pix, win, phi)
    for k in range(len(shapes)):
        print(f"residual window {k} {shapes[k]}", flush=True)
        Fo, do = of.fcorr_residual(rs[k].astype(np.float64), gs[k], phi[k])
        F, nc, fb = out[k]
        assert fb == do["fallback"], (k, out[k], do)
        assert nc == do["n_clean"]
        assert F == pytest.approx(Fo, rel=1e-10)
    assert out[1][2] == "too few clean tiles" and out[0][0] > 1.5


def test_fit_field_carries_the_band_factors():
    import torch
    from tests._scenes import small_star_scene
    from uniphot_amd.fcorr import choose_band_fcorr
    from uniphot_amd.pipeline import DeviceField, FitConfig, fit_field
    from uniphot_amd.psf_polynomial import evaluate_psf_at_batch_np
    sc = small_star_scene(seed=13, H=600, W=600, n_bands=2, n_src=30)
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    res = fit_field(fld, sc.x, sc.y, sc.stype, sc.logre, sc.ee1, sc.ee2, sc.sersic_n,
                    FitConfig(n_rows=1, n_cols=1), sync=torch.cuda.synchronize)
    rec = res.record["fcorr"]
    assert res.fcorr.shape == (2,) and np.all(res.fcorr >= 1.0)
    for b in range(2):
        print(f"band {b}: {rec['source'][b]} F={res.fcorr[b]}", flush=True)
        phi = evaluate_psf_at_batch_np(sc.psfs[b], np.array([300.0]), np.array([300.0]))[0]
        Fo, do = of.fcorr_sky(np.asarray(sc.data[b], np.float64), ~np.asarray(sc.good[b], bool),
                              phi.astype(np.float64))
        assert rec["raw_sky"][b] == pytest.approx(Fo, rel=1e-10)
        assert rec["raw_sky_fallback"][b] == do["fallback"]
    # one 600-px grid: 9 residual tiles, so the residual estimator ran on each band
    assert all(np.isfinite(rec["residual"][b]) or rec["source"][b] != "residual"
               for b in range(2))
    F, src, _ = choose_band_fcorr([[(rec["residual"][b], 1, None) if np.isfinite(
        rec["residual"][b]) else None for b in range(2)]], np.array(rec["raw_sky"]))
    np.testing.assert_allclose(F, res.fcorr)
    fitted = np.isfinite(res.x)
    assert np.isfinite(res.chi2_nu[fitted & (res.island_id >= 0)]).any()
