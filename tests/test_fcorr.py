"""Pins of the correlated-noise factor restatement (oracle/fcorr.py) and of
the host-side pieces of uniphot_amd.fcorr, on the CPU."""
import numpy as np
import pytest
from scipy.ndimage import binary_dilation
from scipy.signal import fftconvolve

from oracle import fcorr as of
from uniphot_amd import fcorr as uf


def _gauss_psf(T=25, s=1.4):
    y, x = np.mgrid[:T, :T] - T // 2
    p = np.exp(-0.5 * (x * x + y * y) / s ** 2)
    return p / p.sum()


def _smoothed_noise(H, W, h, rng):
    """White noise circularly convolved with the kernel h (exact stationary
    correlation rho_h)."""
    z = rng.standard_normal((H, W))
    k = np.zeros((H, W))
    kh, kw = h.shape
    k[:kh, :kw] = h
    k = np.roll(k, (-(kh // 2), -(kw // 2)), (0, 1))
    return np.real(np.fft.ifft2(np.fft.fft2(z) * np.fft.fft2(k)))


def _rho_of_kernel(h, L):
    a = fftconvolve(h, h[::-1, ::-1], mode="full")
    cy, cx = a.shape[0] // 2, a.shape[1] // 2
    big = np.zeros((2 * L + 1, 2 * L + 1))
    r = min(L, cy)
    big[L - r:L + r + 1, L - r:L + r + 1] = a[cy - r:cy + r + 1, cx - r:cx + r + 1]
    return big / a[cy, cx]


def test_fft_autocovariance_equals_direct_circular_sum():
    rng = np.random.default_rng(1)
    H, W, L = 12, 10, 3
    m = rng.random((H, W)) > 0.3
    x = np.where(m, rng.standard_normal((H, W)), 0.0)
    rho, C0 = of.autocov_window(x, m, L)
    A = np.zeros((2 * L + 1, 2 * L + 1))
    K = np.zeros_like(A)
    for dy in range(-L, L + 1):
        for dx in range(-L, L + 1):
            for yy in range(H):
                for xx in range(W):
                    A[dy + L, dx + L] += x[yy, xx] * x[(yy + dy) % H, (xx + dx) % W]
                    K[dy + L, dx + L] += m[yy, xx] * m[(yy + dy) % H, (xx + dx) % W]
    C = A / np.maximum(K, 1.0)
    np.testing.assert_allclose(C0, C[L, L], rtol=1e-12)
    np.testing.assert_allclose(rho, C / C[L, L], rtol=1e-10, atol=1e-12)


def test_dilation_is_the_radius_two_diamond():
    rng = np.random.default_rng(2)
    s = rng.random((40, 33)) > 0.97
    d = binary_dilation(s, iterations=2)
    ref = np.zeros_like(s)
    for y, x in zip(*np.nonzero(s)):
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                if abs(dy) + abs(dx) <= 2 and 0 <= y + dy < 40 and 0 <= x + dx < 33:
                    ref[y + dy, x + dx] = True
    np.testing.assert_array_equal(d, ref)


def test_white_noise_gives_unit_factor():
    rng = np.random.default_rng(3)
    img = 100.0 + 5.0 * rng.standard_normal((768, 768))
    F, dg = of.fcorr_sky(img, None, _gauss_psf())
    assert dg["fallback"] is None or dg["fallback"] == "confusion-shaped ACF"
    assert abs(dg["F_acf"] - 1.0) < 0.03
    assert abs(F - 1.0) < 0.03


def test_kernel_smoothed_noise_gives_the_closed_form_factor():
    rng = np.random.default_rng(4)
    b = np.array([0.25, 0.5, 0.25])
    h = np.outer(b, b)
    img = 7.0 + _smoothed_noise(1024, 1024, h, rng)
    phi = _gauss_psf()
    F, dg = of.fcorr_sky(img, None, phi)
    L = dg["L"]
    expect = float((_rho_of_kernel(h, L) * of.psf_autocorr(phi, L)).sum())
    assert dg["acf_shape"] == "resampling" and dg["q2"] < 0.5
    assert expect > 1.5
    assert abs(F - expect) < 0.03 * expect
    # the same field through the residual estimator (16 tiles, all clean): the
    # four quietest tiles, each detrended by its own mean, read the same
    # correlation less its lowest frequencies (a few per cent low here)
    F2, dg2 = of.fcorr_residual(img - 7.0, np.ones_like(img, bool), phi)
    assert dg2["n_tiles_kept"] == 4
    assert 0.9 * expect < F2 < 1.01 * expect


def test_psf_shaped_correlation_is_rejected_as_confusion():
    rng = np.random.default_rng(5)
    phi = _gauss_psf(s=2.0)
    img = _smoothed_noise(768, 768, phi, rng)
    F, dg = of.fcorr_sky(img, None, phi)
    assert dg["acf_shape"] == "confusion" and dg["q2"] > 0.8
    assert F == 1.0 and dg["fallback"] == "confusion-shaped ACF"


def test_residual_tile_selection_keeps_the_quiet_quartile():
    rng = np.random.default_rng(6)
    r = rng.standard_normal((1024, 1024))
    scale = 1.0 + np.arange(16).reshape(4, 4)[::-1] * 0.01   # tiles 12..15 the quietest
    r *= np.repeat(np.repeat(scale, 256, 0), 256, 1)
    keep, tv, nfin = of.tile_selection(r, np.ones_like(r, bool))
    assert nfin == 16
    np.testing.assert_array_equal(np.nonzero(keep.ravel())[0], [12, 13, 14, 15])


def test_residual_with_too_few_tiles_falls_back():
    rng = np.random.default_rng(7)
    r = rng.standard_normal((300, 300))
    F, dg = of.fcorr_residual(r, None, _gauss_psf())
    assert F == 1.0 and dg["fallback"] == "too few clean tiles"


def test_band_choice_prefers_residual_then_raw_sky():
    per_grid = [[(2.0, 100, None), (0.9, 100, None), None],
                [(4.0, 300, None), (3.0, 50, "too few clean tiles"), None]]
    F, src = of.choose_band_fcorr(per_grid, [1.5, 1.7, 0.8])
    np.testing.assert_allclose(F, [3.5, 1.7, 1.0])
    assert src == ["residual", "raw-sky", "none"]
    F2, src2, _ = uf.choose_band_fcorr(per_grid, np.array([1.5, 1.7, 0.8]))
    np.testing.assert_allclose(F2, F)
    assert src2 == src


@pytest.mark.parametrize("L", [0, 1, 4])
def test_half_window_layout_round_trips(L):
    rng = np.random.default_rng(8)
    H, W = 17, 19
    x = rng.standard_normal((H, W))
    full = np.zeros((2 * L + 1, 2 * L + 1))
    for dy in range(-L, L + 1):
        for dx in range(-L, L + 1):
            full[dy + L, dx + L] = (x * np.roll(np.roll(x, -dy, 0), -dx, 1)).sum()
    half = np.concatenate([full[L, L:], full[L + 1:, :].ravel()])
    np.testing.assert_allclose(uf._full_window(half, L), full, rtol=1e-12)


def test_shape_gate_matches_the_oracle():
    rng = np.random.default_rng(9)
    L = 6
    rho = rng.standard_normal((2 * L + 1, 2 * L + 1)) * 0.1
    rho[L, L] = 1.0
    rho[L, L + 1] = rho[L, L - 1] = rho[L + 1, L] = rho[L - 1, L] = 0.4
    w = of.psf_autocorr(_gauss_psf(), L)
    assert uf.shape_gate(rho, w, L)[:2] == pytest.approx(of.shape_gate(rho, w, L)[:2],
                                                         nan_ok=True)


def test_geometric_error_columns_known_answer():
    import pandas as pd
    from uniphot_amd.error_model import add_error_columns, add_geometric_error_columns
    out = pd.DataFrame({
        "FLUX_g_fit": [10.0, 0.0, np.nan], "FLUXERR_g_fit": [1.0, 1.0, 1.0],
        "FLUX_r_fit": [30.0, 0.0, 5.0], "FLUXERR_r_fit": [1.0, 1.0, 0.0],
        "x_pix_white_err": [0.1, 0.2, 0.3], "Re_err": [1.0, 2.0, np.nan]})
    fcorr = np.array([2.0, 4.0])
    chi2_nu = np.array([4.0, 0.5, np.nan])
    add_geometric_error_columns(out, ["g", "r"], fcorr, chi2_nu)
    # row 0: SNR^2 100 and 900 -> F_eff = (200 + 3600) / 1000 = 3.8; row 1: no
    # band constrains it -> median 3.0; row 2: neither band valid -> 3.0
    feff = np.array([3.8, 3.0, 3.0])
    dfac = np.sqrt([4.0, 1.0, 1.0])
    np.testing.assert_allclose(out["x_pix_white_err_corr"], np.sqrt(feff) * [0.1, 0.2, 0.3],
                               rtol=1e-6)
    np.testing.assert_allclose(out["x_pix_white_err_tot"],
                               np.sqrt(feff) * dfac * [0.1, 0.2, 0.3], rtol=1e-6)
    assert np.isnan(out["Re_err_tot"][2]) and "ELL_err_tot" not in out.columns
    np.testing.assert_allclose(out["chi2_nu_nl"], chi2_nu.astype(np.float32))
    # FLUXERR_MARG_TOT: the increment scaled by G = max(1, F_eff) max(1, chi2_nu)
    conf = np.array([[0.5, np.nan], [0.0, 0.0], [1.0, 1.0]])
    dv = np.array([[2.0, 0.0], [1.0, 1.0], [np.nan, 0.0]])
    add_error_columns(out, ["g", "r"], conf, flux_marg_var=dv, rchisq_scale_marg_tot=False,
                      fcorr=fcorr)
    G = feff * np.array([4.0, 1.0, 1.0])
    np.testing.assert_allclose(out["FLUXERR_MARG_TOT_g_fit"][:2],
                               np.sqrt(1.0 + dv[:2, 0] * G[:2] + np.nan_to_num(conf[:2, 0]) ** 2),
                               rtol=1e-6)
    np.testing.assert_allclose(out["FLUXERR_MARG_TOT_r_fit"][0], np.sqrt(1.0), rtol=1e-6)
