"""Pins of the float64 oracle's rendering primitives (oracle/core.py)
against independent checks: scipy, closed forms, finite differences and
the reference's own table data (golden fixture)."""
from pathlib import Path

import numpy as np
import pytest
from scipy.ndimage import correlate1d
from scipy.special import erf

from oracle import core

GOLD = Path(__file__).parent / "golden"


@pytest.mark.parametrize("d", [-0.5, -0.31, 0.0, 0.127, 0.4999])
def test_lanczos_taps_normalised_and_scipy_shift(d):
    t = core.lanczos_taps(d)
    assert abs(t.sum() - 1.0) < 1e-14
    rng = np.random.default_rng(1)
    img = rng.normal(size=(23, 31))
    ref = correlate1d(correlate1d(img, core.lanczos_taps(d), axis=1, mode="constant"),
                      core.lanczos_taps(-d), axis=0, mode="constant")
    got = core.correlate_y(core.correlate_x(img, core.lanczos_taps(d)), core.lanczos_taps(-d))
    assert np.allclose(got, ref, rtol=0, atol=1e-13)


def test_lanczos_zero_shift_is_identity():
    # sin(pi k) is not exactly zero in floating point: off-centre taps ~1e-17
    t = core.lanczos_taps(0.0)
    assert np.allclose(t, np.array([0, 0, 0, 1.0, 0, 0, 0]), rtol=0, atol=1e-15)
    img = np.random.default_rng(2).normal(size=(9, 9))
    assert np.allclose(core.lanczos_shift(img, 0.0, 0.0), img, rtol=0, atol=1e-14)


def test_lanczos_tap_derivative_fd():
    for d in (-0.45, -0.1, 0.05, 0.33):
        _, dt = core.lanczos_taps(d, grad=True)
        h = 1e-6
        fd = (core.lanczos_taps(d + h) - core.lanczos_taps(d - h)) / (2 * h)
        assert np.allclose(dt, fd, atol=1e-8)


def test_lanczos_half_pixel_moves_a_delta():
    # a shift by +0.5 splits a delta symmetrically between it and its right
    # neighbour (content moves towards +x)
    img = np.zeros((1, 11))
    img[0, 5] = 1.0
    out = core.correlate_x(img, core.lanczos_taps(0.5))
    assert abs(out[0, 5] - out[0, 6]) < 1e-15 and out[0, 5] > 0.5


def test_psf_eval_constant_and_polynomial():
    rng = np.random.default_rng(3)
    c = rng.normal(size=(5, 5, 6))
    psf = dict(coefs=c, x_origin=10.0, x_scale=20.0, y_origin=-4.0, y_scale=8.0)
    x, y = 17.3, 1.9
    X, Y = (x - 10.0) / 20.0, (y + 4.0) / 8.0
    want = (c[..., 0] + c[..., 1] * X + c[..., 2] * Y + c[..., 3] * X * X + c[..., 4] * X * Y
            + c[..., 5] * Y * Y)
    assert np.allclose(core.eval_psf(psf, x, y), want, atol=1e-13)
    s, sx, sy = core.eval_psf(psf, x, y, grad=True)
    h = 1e-6
    assert np.allclose(sx, (core.eval_psf(psf, x + h, y) - core.eval_psf(psf, x - h, y)) / (2 * h),
                       atol=1e-7)
    assert np.allclose(sy, (core.eval_psf(psf, x, y + h) - core.eval_psf(psf, x, y - h)) / (2 * h),
                       atol=1e-7)
    const = dict(psf, coefs=c[..., :1])
    assert np.array_equal(core.eval_psf(const, 1.0, 2.0), core.eval_psf(const, 99.0, -5.0))


def test_term_order_matches_reference_listing():
    # quartic listing of bg_polynomial.py:54-61
    assert core.terms_for_order(4) == (
        (0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3),
        (4, 0), (3, 1), (2, 2), (1, 3), (0, 4))


def test_sersic_table_matches_reference_rows():
    z = np.load(GOLD / "sersic_rows.npz", allow_pickle=False)
    tb = core.sersic_table()
    for n, a, lv in zip(z["n"], z["amps"], z["log_vars"]):
        # the Shajib row of a knot listed twice (n = 0.8) is the second one
        i = int(np.searchsorted(tb["n_knots"], n + 1e-12, side="right")) - 1
        assert abs(tb["n_knots"][i] - n) < 1e-12
        assert np.allclose(tb["amps"][i], a, rtol=0, atol=1e-12)
        assert np.allclose(tb["log_vars"][i], lv, rtol=0, atol=1e-12)
    assert np.all(z["core"] == 0.0)
    # the Tractor family: the fixture's rows 0, 40, 80 of the reference table
    for j, i in enumerate(range(0, 108, 40)):
        assert tb["n_knots"][i] == z["low_n"][j]
        assert np.array_equal(tb["amps"][i], z["low_amps"][j])
        assert np.array_equal(tb["log_vars"][i], z["low_log_vars"][j])
        assert tb["core_amp"][i] == z["low_core"][j]
    assert tb["n_knots"].size == 1209 and tb["n_min"] == 0.29


def test_sersic_tractor_rows_are_normalised_and_segmented():
    """Tractor rows sum to 1 with the core (their own normalisation), the
    knots rise with each segment boundary (0.42, 0.51, 0.57, 0.62, 0.71,
    0.8) listed twice, and the product's table is the oracle's."""
    from uniphot_amd import sersic_table
    tb = core.sersic_table()
    low = slice(0, 108)  # Tractor rows, through the first of the two n = 0.8 knots
    assert tb["n_knots"][107] == 0.8 == tb["n_knots"][108]
    tot = tb["amps"][low].sum(axis=1) + tb["core_amp"][low]
    assert np.allclose(tot, 1.0, atol=1e-12)
    kn = tb["n_knots"]
    dup = kn[1:][np.diff(kn) == 0]
    np.testing.assert_allclose(dup, [0.42, 0.51, 0.57, 0.62, 0.71, 0.8])
    assert np.all(np.diff(kn) >= 0)
    pt = sersic_table.table()   # the package's own generation of the same rows
    for k in ("n_knots", "amps", "log_vars", "core_amp"):
        np.testing.assert_allclose(pt[k], tb[k], rtol=0, atol=1e-13, err_msg=k)


def test_mog_lookup_interpolates_and_clamps():
    a0, v0, c0 = core.mog_params_at(2.0)
    a1, v1, c1 = core.mog_params_at(2.005)
    am, vm, cm = core.mog_params_at(2.0025)
    assert np.allclose(am, 0.5 * (a0 + a1), atol=1e-14)
    assert np.allclose(np.log(vm), 0.5 * (np.log(v0) + np.log(v1)), atol=1e-12)
    # below the table the lookup clamps to n_min, with zero derivatives
    lo = core.mog_params_at(0.2)
    at = core.mog_params_at(0.29)
    assert np.array_equal(lo[0], at[0])
    _, _, _, da, dv, dc = core.mog_params_at(0.2, grad=True)
    assert np.all(da == 0) and np.all(dv == 0)
    # a duplicated boundary knot resolves to the right segment's first row
    tb = core.sersic_table()
    i = int(np.searchsorted(tb["n_knots"], 0.51, side="left")) + 1
    a51 = core.mog_params_at(0.51)
    assert np.array_equal(a51[0], tb["amps"][i])
    # inside a Tractor segment: linear in amplitude and log variance
    b0, b1, bm = core.mog_params_at(0.6), core.mog_params_at(0.605), core.mog_params_at(0.6025)
    assert np.allclose(bm[0], 0.5 * (b0[0] + b1[0]), atol=1e-14)


def _pix_gauss(T, var, cx, cy):
    edges = np.arange(T + 1) - 0.5
    s = np.sqrt(2 * var)
    px = np.diff(0.5 * erf((edges - cx) / s))
    py = np.diff(0.5 * erf((edges - cy) / s))
    return np.outer(py, px)


def test_fft_galaxy_render_gaussian_closed_form():
    """One round Gaussian component (var v) convolved with a pixel-integrated
    Gaussian PSF (var s2) equals the pixel-integrated Gaussian of var s2 + v
    (no wrap: the stamp is far inside the 128 pad)."""
    T, pad = 55, 128
    s2, v = 1.5 ** 2, 2.0 ** 2
    psf = _pix_gauss(T, s2, T // 2, T // 2)
    logre = 0.0  # Re = 1 px -> component variance v * cxx with cxx = cyy = 1
    terms = core.ellipse_cov_terms(logre, 1e-12, 0.0)
    assert np.allclose(terms, [1.0, 1.0, 0.0], atol=1e-9)
    F, _ = core.galaxy_fourier(np.array([1.0]), np.array([v]), terms, pad)
    got = core.fft_convolve_crop(psf, F, pad)
    want = _pix_gauss(T, s2 + v, T // 2, T // 2)
    assert np.max(np.abs(got - want)) < 1e-6 * want.max()


def test_ellipse_terms_closed_form():
    # axis ratio q, angle theta: cxx = re^2 (q^2 cos^2 + sin^2), etc.
    for (lr, e1, e2) in [(0.3, 0.2, -0.1), (1.1, -0.4, 0.25)]:
        cxx, cyy, cxy = core.ellipse_cov_terms(lr, e1, e2)
        ee = np.hypot(e1, e2)
        q = np.exp(-ee) / (2 - np.exp(-ee))
        th = 0.5 * np.arctan2(e2, e1)
        re2 = np.exp(2 * lr)
        assert np.isclose(cxx, re2 * (q * q * np.cos(th) ** 2 + np.sin(th) ** 2))
        assert np.isclose(cyy, re2 * (q * q * np.sin(th) ** 2 + np.cos(th) ** 2))
        assert np.isclose(cxy, re2 * np.cos(th) * np.sin(th) * (1 - q * q))


def _psf_dict(T=55, nt=3, seed=5):
    rng = np.random.default_rng(seed)
    base = _pix_gauss(T, 2.0, T // 2, T // 2)
    c = np.zeros((T, T, nt))
    c[..., 0] = base
    for t in range(1, nt):
        c[..., t] = 0.05 * base * rng.normal(size=(T, T))
    return dict(coefs=c, x_origin=100.0, x_scale=100.0, y_origin=100.0, y_scale=100.0)


@pytest.mark.parametrize("stype,names", [(0, ("x", "y")), (1, ("x", "y", "lr", "e1", "e2")),
                                         (3, ("x", "y", "n"))])
def test_member_stamp_derivatives_fd(stype, names):
    psf = _psf_dict()
    base = dict(x=120.23, y=87.61, lr=0.6, e1=0.15, e2=-0.2, n=2.1313)  # n off a knot

    def render(v):
        s, _ = core.member_stamp(psf, v["x"], v["y"], 3.0, 7.0, stype, v["lr"], v["e1"],
                                 v["e2"], v["n"])
        return s

    _, g = core.member_stamp(psf, base["x"], base["y"], 3.0, 7.0, stype, base["lr"],
                             base["e1"], base["e2"], base["n"], want=names)
    for nm in names:
        h = 1e-5
        vp, vm = dict(base), dict(base)
        vp[nm] += h
        vm[nm] -= h
        fd = (render(vp) - render(vm)) / (2 * h)
        scale = np.abs(fd).max()
        assert np.max(np.abs(g[nm] - fd)) < 1e-5 * scale + 1e-12, nm


def test_direct_backend_oracle_pins():
    """Direct-backend restatement: the shift-and-add convolution equals
    scipy's FFT convolution; a one-component round profile convolved with a
    Gaussian PSF is the sampled Gaussian of the summed variance; the EXP
    stamp has unit sum before the shift; a STAR equals the FFT backend's."""
    from scipy.signal import fftconvolve
    from oracle import core as oc
    rng = np.random.default_rng(3)
    a, k = rng.normal(size=(15, 15)), rng.normal(size=(15, 15))
    np.testing.assert_allclose(oc.convolve_same(a, k), fftconvolve(a, k, mode="same"),
                               rtol=0, atol=1e-12)
    T = 41
    yy, xx = np.mgrid[:T, :T] - T // 2
    sp, re, v = 1.8, 2.0, 0.9
    psf = np.exp(-0.5 * (xx ** 2 + yy ** 2) / sp ** 2)
    psf /= psf.sum()
    prof = oc.exp_profile(T, np.log(re), 0.0, 0.0, amps=[1.0], vars_=[v])
    got = oc.convolve_same(prof, psf)
    s2 = v * re * re + sp * sp
    want = np.exp(-0.5 * (xx ** 2 + yy ** 2) / s2)
    want /= want.sum()
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9 * want.max())
    pd = dict(coefs=psf[..., None], x_origin=0.0, x_scale=1.0, y_origin=0.0, y_scale=1.0)
    raw = oc.direct_stamp(pd, 20.0, 30.0, 0.0, 0.0, oc.SOURCE_EXP, 0.7, 0.2, -0.1)
    np.testing.assert_allclose(raw.sum(), 1.0, atol=1e-12)  # integer position: no shift
    st_d = oc.direct_stamp(pd, 20.3, 30.6, 0.0, 0.0, oc.SOURCE_STAR)
    st_f, _ = oc.member_stamp(pd, 20.3, 30.6, 0.0, 0.0, oc.SOURCE_STAR)
    np.testing.assert_array_equal(st_d, st_f)
