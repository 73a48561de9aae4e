"""Pins of the PSF-build restatement (oracle/psf_build.py) on the CPU."""
import numpy as np
import pytest

from oracle import psf_build as ob


def _dirichlet_shift(img, sx, sy, pad=ob.FOURIER_PAD):
    """The shift as two real 1-D Dirichlet operators (the device's form)."""
    n0, n1 = img.shape

    def op(n, s):
        m = n + 2 * pad
        t = np.arange(n)[:, None] - np.arange(n)[None, :] - s
        with np.errstate(invalid="ignore", divide="ignore"):
            d = np.sin(np.pi * t) / (m * np.sin(np.pi * t / m))
        return np.where(np.abs(t) < 1e-15, 1.0, d)

    return op(n0, sy) @ img @ op(n1, sx).T


def test_integer_shift_is_an_exact_move():
    rng = np.random.default_rng(1)
    a = rng.standard_normal((21, 21))
    out = ob.fourier_shift(a, 2.0, -1.0)
    ref = np.zeros_like(a)
    ref[:, 2:] = a[:, :-2]
    ref = np.roll(ref, -1, 0)
    ref[-1] = 0.0
    np.testing.assert_allclose(out, ref, atol=1e-12)


@pytest.mark.parametrize("sx,sy", [(0.3, -0.45), (-0.5, 0.5), (0.0, 0.21)])
def test_dirichlet_form_equals_the_fft_shift(sx, sy):
    rng = np.random.default_rng(2)
    a = rng.standard_normal((61, 61))
    np.testing.assert_allclose(_dirichlet_shift(a, sx, sy), ob.fourier_shift(a, sx, sy),
                               atol=1e-12)


def _gauss_star(C, x0, y0, s, flux=1.0):
    yy, xx = np.mgrid[:C, :C]
    return flux * np.exp(-0.5 * ((xx - x0) ** 2 + (yy - y0) ** 2) / s ** 2) / (2 * np.pi * s * s)


def test_centroid_of_a_gaussian_star_lands_on_its_centre():
    cut = np.stack([_gauss_star(61, 30.3, 29.6, 1.7), _gauss_star(61, 29.8, 30.45, 2.2)])
    cx, cy, it = ob.gw_centroid(cut, np.full((2, 2), 30.0), fwhm=4.5)
    np.testing.assert_allclose(cx, [30.3, 29.8], atol=1e-2)   # pixel sampling and the 1e-3 stop
    np.testing.assert_allclose(cy, [29.6, 30.45], atol=1e-2)
    assert 2 <= it <= 8


def test_annulus_median_of_constant_sky_with_outliers():
    img = np.full((100, 100), 7.5, np.float32)
    img[50, 75] = 1e4
    img[25, 50] = -1e4
    bad = np.zeros_like(img, bool)
    assert ob.annulus_median(img, 50, 50, 20.0, 30.0, 3.0, 30, bad) == pytest.approx(7.5)
    bad[:, :] = True
    assert np.isnan(ob.annulus_median(img, 50, 50, 20.0, 30.0, 3.0, 30, bad))


def test_polynomial_fit_recovers_a_polynomial_psf_and_rejects_a_bad_star():
    rng = np.random.default_rng(3)
    H = W = 1000
    order = 2
    n = 40
    x, y = rng.uniform(50, 950, n), rng.uniform(50, 950, n)
    nt = len(ob.terms_for_order(order))
    truth = rng.standard_normal((15, 15, nt)) * 0.01
    truth[:, :, 0] += _gauss_star(15, 7, 7, 1.5)
    A = ob._design(x, y, (H, W), order)
    stamps = (A @ truth.reshape(-1, nt).T).reshape(n, 15, 15)
    stamps[7] = np.roll(stamps[7], 3, 1)              # a contaminated star
    coefs, acc, it, chi2 = ob.fit_polynomial(x, y, stamps.astype(np.float32), (H, W), order)
    assert not acc[7] and acc.sum() == n - 1
    np.testing.assert_allclose(coefs, truth, atol=2e-6)
    assert chi2[7] > 100 * np.median(chi2)


def test_wing_blend_keeps_the_core_and_the_constant_sum():
    rng = np.random.default_rng(4)
    coefs = rng.random((55, 55, 3)).astype(np.float32)
    stamps = rng.random((25, 55, 55)).astype(np.float32)
    out = ob.wing_blend(coefs, stamps, None, fwhm=3.0)
    yy, xx = np.mgrid[:55, :55]
    r = np.hypot(yy - 27, xx - 27)
    assert out[:, :, 0].sum() == pytest.approx(coefs[:, :, 0].sum(), rel=1e-6)
    core = r < 7.5
    ratio = out[:, :, 0][core] / coefs[:, :, 0][core]
    np.testing.assert_allclose(ratio, ratio[0], rtol=1e-6)    # one renormalisation factor
    np.testing.assert_allclose(out[:, :, 1:][r >= 11.5], 0.0)
    assert ob.wing_blend(coefs, stamps[:19], None, 3.0) is coefs   # below 20 stars: unchanged


def test_psf_polynomial_npz_round_trip(tmp_path):
    from uniphot_amd.psf_build import (PsfBuildResult, load_psf_polynomial_for_band,
                                       save_psf_polynomial)
    from uniphot_amd.psf_polynomial import PsfPolynomial
    rng = np.random.default_rng(5)
    psf = PsfPolynomial(rng.random((55, 55, 15)).astype(np.float32), 600.0, 600.0, 550.0,
                        550.0, 55, 55)
    acc = np.array([True, False, True])
    res = PsfBuildResult(psf, np.zeros((3, 55, 55), np.float32), np.array([1.0, 2.0, 3.0]),
                         np.array([4.0, 5.0, 6.0]), np.arange(3), acc, 2, np.zeros(3), 4)
    path = save_psf_polynomial(tmp_path, "m400", res, 4)
    with np.load(path) as z:
        assert sorted(z.files) == sorted(["coefs", "x_origin", "x_scale", "y_origin", "y_scale",
                                          "H_psf", "W_psf", "polynomial_order", "star_x",
                                          "star_y"])
        assert z["x_origin"].dtype == np.float32 and z["H_psf"].dtype == np.int32
        np.testing.assert_array_equal(z["star_x"], [1.0, 3.0])
    back = load_psf_polynomial_for_band(tmp_path, "m400")
    np.testing.assert_array_equal(back.coefs, psf.coefs)
    assert (back.x_origin, back.y_scale, back.H_psf) == (600.0, 550.0, 55)
    with pytest.raises(FileNotFoundError):
        load_psf_polynomial_for_band(tmp_path, "m425")
