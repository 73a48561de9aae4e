"""PSF model build on the device (uph_psf_*) against the oracle restatement
(oracle/psf_build.py) on a synthetic star field."""
import numpy as np
import pytest

from oracle import psf_build as ob

pytestmark = pytest.mark.gpu


def _field(seed=5, H=1100, W=1200, n=64):
    rng = np.random.default_rng(seed)
    img = 100.0 + rng.standard_normal((H, W))
    x = rng.uniform(40, W - 40, n)
    y = rng.uniform(40, H - 40, n)
    x[0], y[0] = 12.0, 500.3                    # off the frame for a 61-px cutout
    x[1], y[1] = 600.2, 20.7                    # likewise
    flux = rng.uniform(2e4, 2e5, n)
    yy, xx = np.mgrid[:H, :W]
    for k in range(n):
        s = 1.6 + 0.3 * (x[k] - W / 2) / (W / 2) + 0.1 * (y[k] - H / 2) / (H / 2)
        y0, y1 = int(max(0, y[k] - 30)), int(min(H, y[k] + 31))
        x0, x1 = int(max(0, x[k] - 30)), int(min(W, x[k] + 31))
        g = np.exp(-0.5 * ((xx[y0:y1, x0:x1] - x[k]) ** 2 + (yy[y0:y1, x0:x1] - y[k]) ** 2)
                   / s ** 2)
        img[y0:y1, x0:x1] += flux[k] * g / (2 * np.pi * s * s)
    img[int(y[5]) + 8, int(x[5]) + 8] += 3e4       # a contaminated star
    bad = np.zeros((H, W), bool)
    bad[int(y[9]) - 3:int(y[9]) + 2, int(x[9]) + 5:int(x[9]) + 9] = True   # bad pixels in a box
    bad[300:305, :] |= False
    return img.astype(np.float32), bad, x, y


@pytest.mark.parametrize("order", [2, 4])
def test_psf_build_matches_oracle(order):
    import torch
    from uniphot_amd.psf_build import build_psf_from_stars
    img, bad, x, y = _field()
    print(f"psf build order {order}: {x.size} stars", flush=True)
    r = build_psf_from_stars(torch.from_numpy(img).cuda(),
                             torch.from_numpy(bad.astype(np.uint8)).cuda(), x, y, fwhm=4.0,
                             polynomial_order=order)
    st, fx, fy, kept = ob.extract_centered_stamps(img, bad, x, y, fwhm=4.0)
    np.testing.assert_array_equal(r.kept, kept)
    assert 0 not in kept and 1 not in kept and 9 not in kept
    np.testing.assert_allclose(r.fitted_x, fx, rtol=0, atol=1e-9)
    np.testing.assert_allclose(r.fitted_y, fy, rtol=0, atol=1e-9)
    np.testing.assert_allclose(r.stamps, st, rtol=0, atol=2e-6 * float(np.abs(st).max()))
    coefs, acc, it, chi2 = ob.fit_polynomial(fx, fy, st, img.shape, order=order)
    np.testing.assert_array_equal(r.accepted, acc)
    assert r.n_iters == it
    np.testing.assert_allclose(r.chi2, chi2, rtol=1e-4, atol=1e-12)
    ref = ob.wing_blend(coefs, st, acc, 4.0)
    scale = float(np.abs(ref).max())
    np.testing.assert_allclose(r.psf.coefs, ref, rtol=0, atol=1e-5 * scale)
    assert r.psf.x_origin == 600.0 and r.psf.y_scale == 550.0
    assert not acc[list(kept).index(5)]                 # the contaminated star is rejected
