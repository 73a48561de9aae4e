"""Band-flag PSF diagnostics on the device (uph_band_psf_flags) against the
float64 oracle: lobe zone and uncovered fraction of 15-term PSF models at
random positions (edges, corners, off the image, NaN), per-band no-data
masks with footprint edges and holes, and the reference's known answers
through the single-band facades."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

from oracle import flags as oflags
from uniphot_amd.band_flags import (FLAG_NO_DATA, FLAG_PARTIAL_COVERAGE, FLAG_PSF_LOBE,
                                    apply_band_flags, lobe_from_values, psf_flag_values,
                                    psf_uncovered_fraction)
from uniphot_amd.psf_polynomial import PsfPolynomial

pytestmark = pytest.mark.gpu

H, W = 240, 320


def _psf(rng, S, lobe):
    """Order-4 PSF: a Gaussian core whose width varies over the field plus
    a ring whose depth runs from +lobe to -lobe of the peak across X."""
    yy, xx = np.mgrid[:S, :S] - S // 2
    r2 = xx ** 2 + yy ** 2
    sig = rng.uniform(1.5, 2.5)
    g = np.exp(-0.5 * r2 / sig ** 2)
    g /= g.sum()
    ring = (np.sqrt(r2) > S // 3).astype(float) * g.max()
    c = np.zeros((S, S, 15))
    c[..., 0] = g
    c[..., 1] = -lobe * ring                     # X
    c[..., 2] = 0.05 * (r2 * g) / (r2 * g).sum()  # Y: a slowly varying wing
    c[..., 3:] = rng.normal(0, 1e-5, (S, S, 12))
    return PsfPolynomial(coefs=c.astype(np.float32), x_origin=W / 2, x_scale=W / 2,
                         y_origin=H / 2, y_scale=H / 2, H_psf=S, W_psf=S)


def _odict(p):
    return dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                y_origin=p.y_origin, y_scale=p.y_scale)


def _masks(rng, nb):
    out = []
    for b in range(nb):
        if b == 2:
            out.append(None)                     # a band without a mask
            continue
        m = np.zeros((H, W), bool)
        m[:, W - 40 - 10 * b:] = True            # footprint edge
        m[:15 + 5 * b, :] = True
        for _ in range(30):                      # holes
            y0, x0 = rng.integers(0, H), rng.integers(0, W)
            m[y0:y0 + rng.integers(1, 6), x0:x0 + rng.integers(1, 6)] = True
        out.append(m)
    return out


def test_psf_flags_match_oracle(cuda):
    rng = np.random.default_rng(11)
    nb = 4
    sizes = [21, 21, 21, 21]
    psfs = [_psf(rng, S, lobe) for S, lobe in zip(sizes, (0.004, 0.01, 0.003, 0.0))]
    masks = _masks(rng, nb)
    n = 1500
    x = rng.uniform(-30, W + 30, n)
    y = rng.uniform(-30, H + 30, n)
    x[:40] = np.round(x[:40]) + 0.5               # round-half-to-even ties
    x[40:50], y[50:60] = np.nan, np.inf
    x[60:70] = rng.uniform(0, 3, 10)              # left border
    pk, mn, unc = psf_flag_values(psfs, x, y, masks, device=cuda)
    for b, p in enumerate(psfs):
        od = _odict(p)
        fin = np.isfinite(x) & np.isfinite(y)
        st = [oflags.psf_stamp(od, a, c) for a, c in zip(x[fin], y[fin])]
        opk = np.array([s.max() for s in st])
        omn = np.array([s.min() for s in st])
        scale = np.abs(opk)
        np.testing.assert_allclose(pk[b][fin], opk, rtol=0, atol=1e-6 * scale.max())
        np.testing.assert_allclose(mn[b][fin], omn, rtol=0, atol=1e-6 * scale.max())
        assert np.isnan(pk[b][~fin]).all()
        zone = lobe_from_values(pk[b], mn[b])
        ozone = oflags.lobe_zone(od, x, y)
        tie = np.zeros(n, bool)
        tie[fin] = np.abs(omn + 0.002 * opk) < 1e-6 * scale.max()
        assert np.array_equal(zone[~tie], ozone[~tie]), b
        if masks[b] is None:
            assert (unc[b] == 0).all()
            continue
        ou = oflags.uncovered_fraction(od, masks[b], x, y)
        np.testing.assert_allclose(unc[b], ou, rtol=0, atol=1e-6)
        assert (unc[b][~fin] == 0).all()
    # the data exercise every case
    assert (unc[:2] == 1.0).any() and ((unc[:2] > 0) & (unc[:2] < 1)).any()
    assert lobe_from_values(pk[1], mn[1]).any() and not lobe_from_values(pk[3], mn[3]).all()


def test_uncovered_fraction_known_answers_on_the_device(cuda):
    S = 21
    g = np.exp(-0.5 * (np.arange(S) - S // 2) ** 2 / 4.0)
    g /= g.sum()
    c = g[S // 2]
    psf = PsfPolynomial(coefs=np.outer(g, g)[..., None].astype(np.float32), x_origin=W / 2,
                        x_scale=W / 2, y_origin=H / 2, y_scale=H / 2, H_psf=S, W_psf=S)
    nodata = np.zeros((H, W), bool)
    u = psf_uncovered_fraction(psf, nodata, np.array([0.0, 150.0, np.nan, -40.0, 0.0]),
                               np.array([100.0, 100.0, 5.0, 100.0, 0.0]), device=cuda)
    np.testing.assert_allclose(u[0], (1 - c) / 2, atol=1e-7)
    assert u[1] == 0.0 and u[2] == 0.0 and u[3] == 1.0
    np.testing.assert_allclose(u[4], 1 - ((1 + c) / 2) ** 2, atol=1e-7)
    nodata[:, 201:] = True
    u = psf_uncovered_fraction(psf, nodata, np.array([200.0, 200.5]), np.array([60.0, 60.0]),
                               device=cuda)
    np.testing.assert_allclose(u, (1 - c) / 2, atol=1e-7)


def test_apply_band_flags_on_a_catalogue(cuda):
    rng = np.random.default_rng(5)
    psfs = [_psf(rng, 21, 0.01), _psf(rng, 21, 0.0)]
    m = np.zeros((H, W), bool)
    m[:, 250:] = True
    x = np.array([10.0, 100.0, 249.0, 300.0, 200.0])
    y = np.array([120.0, 120.0, 120.0, 120.0, 120.0])
    out = pd.DataFrame({"x_pix_white_fit": x, "y_pix_white_fit": y,
                        "FLUX_g_fit": [1.0, 2.0, 3.0, 4.0, 5.0],
                        "FLUXERR_g_fit": [0.1, 0.1, 0.1, 0.1, np.nan],
                        "FLUX_r_fit": [1.0, 2.0, 3.0, 4.0, 5.0],
                        "FLUXERR_r_fit": [0.1] * 5})
    rec = apply_band_flags(out, ["g", "r"], psfs, nodata=[m, None], device=cuda)
    zone = oflags.lobe_zone(_odict(psfs[0]), x, y)
    part = oflags.uncovered_fraction(_odict(psfs[0]), m, x, y) > 0.1
    want = (np.where(zone, FLAG_PSF_LOBE, 0) | np.where(part, FLAG_PARTIAL_COVERAGE, 0)
            | np.array([0, 0, 0, 0, FLAG_NO_DATA]))
    assert out["FLAGS_g_fit"].tolist() == want.tolist()
    assert zone[3] and not zone[0]            # the ring deepens towards +X
    assert part[2] and part[3] and not part[1]
    assert out["FLAGS_r_fit"].tolist() == [0] * 5
    assert np.isnan(out.loc[4, "FLUX_g_fit"]) and rec["blanked"]["g"] == 1
    assert out["n_bands_usable"].tolist() == (1 + (want == 0)).tolist()
