"""Pins of the float64 confusion-noise restatement (oracle/error_model.py)
against closed forms and brute force, and the error-column rules."""
from __future__ import annotations

import numpy as np
import pandas as pd

from oracle import error_model as oem


def _phi(S=21, sig=1.6):
    yy, xx = np.mgrid[:S, :S] - S // 2
    g = np.exp(-0.5 * (xx ** 2 + yy ** 2) / sig ** 2)
    return (g / g.sum()).astype(np.float32)


def test_same_correlation_equals_direct_sum():
    from scipy.signal import correlate2d
    rng = np.random.default_rng(0)
    a, k = rng.normal(size=(40, 33)), rng.normal(size=(9, 9))
    np.testing.assert_allclose(oem.corr_same(a, k), correlate2d(a, k, mode="same"),
                               rtol=0, atol=1e-11)


def test_matched_filter_recovers_an_unmodelled_point_source():
    phi = _phi()
    H = W = 120
    data = np.zeros((H, W))
    c = 10
    for (y, x, f) in ((40, 50, 250.0), (80, 70, -40.0)):
        data[y - c:y + c + 1, x - c:x + c + 1] += f * phi
    good = np.ones((H, W), bool)
    mf, vs, valid = oem.confusion_maps(data, 0.0 * data, good, np.ones((H, W)), phi)
    assert valid[40, 50] and valid[80, 70]
    np.testing.assert_allclose([mf[40, 50], mf[80, 70]], [250.0, -40.0], rtol=1e-7)
    # uniform unit variance: Vstat = 1 / sum(phi^2) in the interior
    np.testing.assert_allclose(vs[60, 60], 1.0 / np.sum(phi.astype(np.float64) ** 2), rtol=1e-9)


def test_zero_and_constant_residuals_have_no_excess_and_annulus_geometry():
    phi = _phi()
    _, sp2, r_in, r_out = oem.psf_scale(phi)
    H = W = 160
    good = np.ones((H, W), bool)
    xi = np.array([80, 79, 5])
    yi = np.array([80, 90, 5])
    for resid in (0.0, 3.0):
        data = np.full((H, W), resid)
        sig = oem.confusion_sigma(data, np.zeros((H, W)), good, np.ones((H, W)), phi, xi[:2],
                                  yi[:2], np.zeros(2), np.zeros(2, bool))
        assert np.all(sig[np.isfinite(sig)] < 1e-6) and np.isfinite(sig).all()
    vf = np.ones((H, W))
    _, n = oem.annulus_mean(vf, vf, r_in, r_out)
    assert n[80, 80] == (2 * r_out + 1) ** 2 - (2 * r_in + 1) ** 2
    # a star's template factor is 1; no fit gives NaN
    sg = sp2 + 0.0
    assert 2 * sg * sg / ((sg + sp2) * sp2) == 1.0
    sig = oem.confusion_sigma(np.zeros((H, W)), np.zeros((H, W)), good, np.ones((H, W)), phi,
                              xi, yi, np.zeros(3), np.array([False, True, False]))
    assert np.isnan(sig[1]) and np.isfinite(sig[0])


def test_white_noise_excess_is_small_against_vstat():
    rng = np.random.default_rng(2)
    phi = _phi()
    H = W = 400
    sigma = 3.0
    data = rng.normal(0, sigma, (H, W))
    good = np.ones((H, W), bool)
    iv = np.full((H, W), 1.0 / sigma ** 2)
    ys, xs = np.mgrid[60:340:40, 60:340:40]
    sig = oem.confusion_sigma(data, np.zeros((H, W)), good, iv, phi, xs.ravel(), ys.ravel(),
                              np.zeros(xs.size), np.zeros(xs.size, bool))
    vstat = sigma ** 2 / np.sum(phi.astype(np.float64) ** 2)
    # the confusion variance of pure white noise is a small fraction of
    # the statistical one (annulus sampling noise only)
    assert np.median(sig ** 2) < 0.1 * vstat
    # an unmodelled population of PSF-shaped sources raises it above that
    for _ in range(300):
        y, x = rng.integers(15, H - 15, 2)
        data[y - 10:y + 11, x - 10:x + 11] += rng.uniform(50, 500) * phi
    sig2 = oem.confusion_sigma(data, np.zeros((H, W)), good, iv, phi, xs.ravel(), ys.ravel(),
                               np.zeros(xs.size), np.zeros(xs.size, bool))
    assert np.median(sig2 ** 2) > vstat


def test_error_columns_rules():
    from uniphot_amd.error_model import add_error_columns
    out = pd.DataFrame({"FLUX_g_fit": [10.0, 5.0, 7.0], "FLUXERR_g_fit": [1.0, 2.0, np.nan],
                        "FLUXERR_MARG_g_fit": [1.0, 3.0, np.nan], "rchisq": [0.5, 4.0, 1.0]})
    conf = np.array([[2.0], [np.nan], [1.0]])
    dvar = np.array([[0.0], [5.0], [np.nan]])
    add_error_columns(out, ["g"], conf, flux_marg_var=dvar)
    np.testing.assert_allclose(out["FLUXERR_CONF_g_fit"], [2.0, np.nan, 1.0])
    np.testing.assert_allclose(out["FLUXERR_TOT_g_fit"], [np.sqrt(5.0), 2.0, np.nan])
    # MARG_TOT: star (dvar 0) -> TOT; dvar -> sqrt(fe^2 + dvar + conf^2) x sqrt(max(1, rchisq))
    np.testing.assert_allclose(out["FLUXERR_MARG_TOT_g_fit"],
                               [np.sqrt(5.0), np.sqrt(4.0 + 5.0) * 2.0, np.nan], rtol=1e-6)
