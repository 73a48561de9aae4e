"""Band flags on the CPU: the oracle's PSF diagnostics against known answers
(closed forms and the reference's own test cases,
tract7dt/tests/test_partial_coverage.py, test_output_contract_nodata.py,
test_frames_and_saturation.py), and the column rules of
uniphot_amd.band_flags on hand-built catalogues."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

from oracle import flags as oflags
from uniphot_amd import _lib
from uniphot_amd.band_flags import (BAND_FLUX_COLUMN_PREFIXES, FLAG_NO_DATA,
                                    FLAG_PARTIAL_COVERAGE, FLAG_PSF_LOBE, FLAG_SATURATED,
                                    PARTIAL_COVERAGE_MAX, PSF_LOBE_FRAC, blank_nodata_band_cells,
                                    compose_band_flags, lobe_from_values)

S = 21
H, W = 200, 300


def gauss1d(sig=2.0):
    g = np.exp(-0.5 * (np.arange(S) - S // 2) ** 2 / sig ** 2)
    return g / g.sum()


def psf_dict(stamp, extra=None):
    """Constant PSF (order 0), or order 1 with ``extra`` as the X term."""
    c = stamp[..., None] if extra is None else np.stack([stamp, extra, np.zeros_like(stamp)], -1)
    return dict(coefs=c.astype(np.float64), x_origin=W / 2, x_scale=W / 2, y_origin=H / 2,
                y_scale=H / 2)


def test_uncovered_fraction_known_answers():
    g = gauss1d()
    stamp = np.outer(g, g)
    psf = psf_dict(stamp)
    c = g[S // 2]  # the centre column's share of a separable unit stamp
    nodata = np.zeros((H, W), bool)
    u = oflags.uncovered_fraction(psf, nodata, [150.0, 40.0], [100.0, 60.0])
    assert u.tolist() == [0.0, 0.0]
    # footprint ending right of the centre column: the right half is missing
    nd = np.zeros((H, W), bool)
    nd[:, 201:] = True
    u = oflags.uncovered_fraction(psf, nd, [200.0, 200.4, 199.6], [60.0, 60.0, 60.0])
    np.testing.assert_allclose(u, (1 - c) / 2, rtol=0, atol=1e-14)
    nd[:, 200] = True  # ... and from the centre column on
    u = oflags.uncovered_fraction(psf, nd, [200.0], [60.0])
    np.testing.assert_allclose(u, (1 + c) / 2, rtol=0, atol=1e-14)
    # the image border counts as no data; NaN never flagged; off the image 1
    u = oflags.uncovered_fraction(psf, np.zeros((H, W), bool), [0.0, 150.0, np.nan, -40.0],
                                  [100.0, 100.0, 5.0, 100.0])
    np.testing.assert_allclose(u[0], (1 - c) / 2, atol=1e-14)
    assert u[1] == 0.0 and u[2] == 0.0 and u[3] == 1.0
    # a corner source: two independent halves missing -> 1 - ((1 + c) / 2)^2
    u = oflags.uncovered_fraction(psf, np.zeros((H, W), bool), [0.0], [0.0])
    np.testing.assert_allclose(u, 1 - ((1 + c) / 2) ** 2, atol=1e-14)
    # round half to even: x = 200.5 centres on column 200
    nd2 = np.zeros((H, W), bool)
    nd2[:, 201:] = True
    np.testing.assert_allclose(oflags.uncovered_fraction(psf, nd2, [200.5], [60.0]),
                               (1 - c) / 2, atol=1e-14)
    # no positive part and a no-data hit: 1
    neg = psf_dict(-stamp)
    assert oflags.uncovered_fraction(neg, nd2, [200.0], [60.0])[0] == 1.0
    # no positive part and no hit: 0 (never evaluated)
    assert oflags.uncovered_fraction(neg, nd2, [100.0], [60.0])[0] == 0.0


def test_lobe_zone_known_answers():
    g = gauss1d()
    stamp = np.outer(g, g)
    pk = stamp.max()
    yy, xx = np.mgrid[:S, :S] - S // 2
    ring = (np.hypot(xx, yy) > 8).astype(float)
    xs, ys = np.array([150.0, 10.0]), np.array([100.0, 20.0])
    assert not oflags.lobe_zone(psf_dict(stamp), xs, ys).any()
    assert oflags.lobe_zone(psf_dict(stamp - 0.003 * pk * ring), xs, ys).all()
    assert not oflags.lobe_zone(psf_dict(stamp - 0.001 * pk * ring), xs, ys).any()
    assert oflags.lobe_zone(psf_dict(np.zeros_like(stamp)), xs, ys).all()
    assert not oflags.lobe_zone(psf_dict(stamp - 0.003 * pk * ring), [np.nan], [1.0]).any()
    # order 1: the lobe grows with X = (x - 150) / 150; -1 % of the peak at
    # X = 1, so it crosses -0.2 % at X = 0.2 (x = 180)
    var = psf_dict(stamp, extra=-0.01 * pk * ring)
    z = oflags.lobe_zone(var, [0.0, 150.0, 179.0, 181.0, 299.0], [5.0] * 5)
    assert z.tolist() == [False, False, False, True, True]


def test_lobe_rule_of_values():
    pk = np.array([1.0, 1.0, 1.0, 0.0, -1.0, np.nan])
    mn = np.array([0.0, -0.0019, -0.0021, 0.0, -2.0, np.nan])
    assert lobe_from_values(pk, mn).tolist() == [False, False, True, True, True, False]
    assert PSF_LOBE_FRAC == 0.002 and PARTIAL_COVERAGE_MAX == 0.1


def _band_block(bn, flux, ferr, win=None, conf=None):
    n = len(flux)
    ferr = np.asarray(ferr, np.float64)
    tot = np.sqrt(ferr ** 2 + 0.01)
    return {
        f"FLUX_{bn}_fit": np.asarray(flux, np.float64),
        f"FLUXERR_{bn}_fit": ferr.astype(np.float32),
        f"FLUXERR_CONF_{bn}_fit": np.full(n, 0.1) if conf is None else np.asarray(conf, float),
        f"FLUX_WIN_{bn}_fit": (np.asarray(flux, np.float64) * 0.9 if win is None
                               else np.asarray(win, float)),
        f"FLUXERR_TOT_{bn}_fit": tot.astype(np.float32),
        f"FLUXERR_MARG_{bn}_fit": ferr,
        f"FLUX_MARG_FRAC_{bn}_fit": np.zeros(n, np.float32),
        f"FLUXERR_MARG_TOT_{bn}_fit": tot.astype(np.float32),
    }


def _catalog():
    nan = np.nan
    cols = {"ID": ["a", "b", "c", "d", "e"], "grid_id_fit": [7, 7, 7, 8, -1]}
    cols.update(_band_block("g", flux=[10.0, 0.37, 5.0, 5.5e6, nan],
                            ferr=[1.0, nan, 0.5, nan, nan], win=[9.0, 0.0, 4.5, 3.0, nan],
                            conf=[0.1, 0.2, 0.1, nan, nan]))
    cols.update(_band_block("r", flux=[20.0, 3.0, -1.0, 7.0, nan],
                            ferr=[2.0, 0.4, 0.3, 0.7, nan]))
    return pd.DataFrame(cols)


def test_nodata_cells_blank_every_band_column():
    out = _catalog()
    before = out.copy()
    n = blank_nodata_band_cells(out, ["g", "r"])
    for band, bad in (("g", [False, True, False, True, True]),
                      ("r", [False, False, False, False, True])):
        bad = np.array(bad)
        for p in BAND_FLUX_COLUMN_PREFIXES:
            col = f"{p}{band}_fit"
            assert out.loc[bad, col].isna().all(), col
            pd.testing.assert_series_equal(out.loc[~bad, col], before.loc[~bad, col])
    pd.testing.assert_series_equal(out["grid_id_fit"], before["grid_id_fit"])
    assert n == {"g": 2, "r": 0}
    # a fully measured catalogue is untouched; band names match exactly
    full = _catalog().iloc[[0, 2]].reset_index(drop=True)
    ref = full.copy()
    assert blank_nodata_band_cells(full, ["g", "r"]) == {"g": 0, "r": 0}
    pd.testing.assert_frame_equal(full, ref)
    two = pd.DataFrame({})
    for bn in ("m400", "m400w"):
        for k, v in _band_block(bn, flux=[2.0, 3.0], ferr=[np.nan, 1.0]).items():
            two[k] = v
    keep = two["FLUX_m400w_fit"].copy()
    assert blank_nodata_band_cells(two, ["m400"]) == {"m400": 1}
    assert np.isnan(two.loc[0, "FLUX_m400_fit"])
    pd.testing.assert_series_equal(two["FLUX_m400w_fit"], keep)


def test_compose_saturation_partial_and_lobe_bits():
    out = pd.DataFrame({"FLUX_g_fit": [10.0, 20.0, np.nan], "FLUXERR_g_fit": [1.0, 2.0, np.nan],
                        "FLUX_WIN_g_fit": [9.0, 18.0, np.nan], "FLUX_r_fit": [5.0, 6.0, 7.0],
                        "FLUXERR_r_fit": [0.5, 0.6, 0.7]})
    counts = compose_band_flags(out, ["g", "r"], saturated={"g": np.array([True, False, True])})
    assert out["FLAGS_g_fit"].tolist() == [FLAG_SATURATED, 0, FLAG_NO_DATA | FLAG_SATURATED]
    assert np.isnan(out.loc[0, "FLUX_g_fit"]) and np.isnan(out.loc[0, "FLUX_WIN_g_fit"])
    assert out.loc[1, "FLUX_g_fit"] == 20.0
    assert out["FLAGS_r_fit"].tolist() == [0, 0, 0]
    assert out["n_bands_usable"].tolist() == [1, 2, 1]
    assert counts["g"]["saturated"] == 1
    out = pd.DataFrame({"FLUX_g_fit": [10.0, 20.0, np.nan], "FLUXERR_g_fit": [1.0, 2.0, np.nan],
                        "FLUX_r_fit": [5.0, 6.0, 7.0], "FLUXERR_r_fit": [1.0, 1.0, 1.0]})
    counts = compose_band_flags(out, ["g", "r"], partial={"g": np.array([False, True, True])},
                                psf_zone={"r": np.array([True, False, False])})
    assert out["FLAGS_g_fit"].tolist() == [0, FLAG_PARTIAL_COVERAGE,
                                           FLAG_NO_DATA | FLAG_PARTIAL_COVERAGE]
    assert out["FLAGS_r_fit"].tolist() == [FLAG_PSF_LOBE, 0, 0]
    assert out.loc[1, "FLUX_g_fit"] == 20.0
    assert out["n_bands_usable"].tolist() == [1, 1, 1]
    assert counts["g"]["partial_coverage"] == 1 and counts["r"]["psf_lobe"] == 1
    assert out["FLAGS_g_fit"].dtype == np.int16


def test_band_flags_argument_errors():
    L = _lib.lib()
    with pytest.raises(_lib.UphError) as ei:
        _lib.check(L.uph_band_psf_flags(4, 1, 22, 21, *([None] * 5), 10, 10, *([None] * 5),
                                        None))
    assert ei.value.code == 1 and "n_terms=22" in str(ei.value)
    with pytest.raises(_lib.UphError) as ei:
        _lib.check(L.uph_band_psf_flags(4, 1, 1, 21, *([None] * 5), 10, 10, *([None] * 5),
                                        None))
    assert "null argument" in str(ei.value)
    assert L.uph_band_psf_flags(0, 1, 1, 21, *([None] * 5), 10, 10, *([None] * 5), None) == 0


def test_reliability_contract():
    from uniphot_amd.band_flags import compute_reliability_flag
    out = pd.DataFrame({
        "x_pix_white_fit": [1.0, np.nan, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0],
        "excluded_any": [False, False, False, False, False, False, False, True],
        "opt_hit_max_iters": [False, False, True, False, False, False, False, False],
        "opt_converged": [True, False, False, False, True, True, True, True],
        "flag_bound_stalled": [False] * 8,
        "island_stalled": [False, False, False, False, True, True, False, False],
        "island_plateau": [False, False, False, False, True, False, False, False],
        "island_diverged": [False, False, False, False, False, True, False, False],
        "island_giant": [False, False, False, False, False, False, True, False],
        "n_bands_usable": [2, 0, 2, 0, 2, 2, 2, 0],
        "shape_frozen": [False, False, False, True, False, False, False, False],
    })
    a = out.copy()
    compute_reliability_flag(a)
    assert a["fit_is_reliable"].tolist() == [True, False, False, False, False, False, False,
                                             False]
    assert a["fit_quality_issues"].tolist() == [
        "", "no_fit_result", "hit_max_iters", "no_usable_band,not_converged,shape_frozen",
        "island_stalled", "island_stalled,island_diverged", "giant_frozen_geometry",
        "excluded"]
    b = out.copy()
    compute_reliability_flag(b, island_plateau_mode="converged", island_giant_mode="warn")
    assert b["fit_is_reliable"].tolist() == [True, False, False, False, True, False, True, False]
    assert b["fit_quality_issues"].tolist()[4] == "island_plateau"
    assert b["fit_quality_issues"].tolist()[6] == "giant_frozen_geometry"
    c = out.copy()
    compute_reliability_flag(c, island_stalled_mode="ignore", island_giant_mode="ignore")
    assert c["fit_is_reliable"].tolist()[4:7] == [True, True, True]
    assert c["fit_quality_issues"].tolist()[5] == "island_diverged"
    with pytest.raises(ValueError):
        compute_reliability_flag(out.copy(), high_rchisq_mode="loud")
    empty = pd.DataFrame({"x_pix_white_fit": []})
    compute_reliability_flag(empty)
    assert list(empty.columns) == ["x_pix_white_fit", "fit_is_reliable", "fit_quality_issues"]
