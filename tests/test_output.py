"""Catalogue writeback of an island-mode fit: the column rules of the
reference's pipeline.py (:2498-2910) and error_columns.py (:255-341) on a
hand-built fit result."""
import numpy as np
import pandas as pd

from uniphot_amd.output import slow_converge_flag, write_fit_columns
from uniphot_amd.pipeline import FitResult
from uniphot_amd.source_model import ellipse_soft_from_re_ab_phi_np


def _fit():
    n, nb = 4, 2
    st = np.array([0, 1, 3, 2], np.int32)
    lr, e1, e2 = ellipse_soft_from_re_ab_phi_np(np.array([1.0, 2.0, 3.0, 41.25]),
                                                np.array([1.0, 0.5, 0.8, 0.9]),
                                                np.array([0.0, 30.0, 120.0, 60.0]))
    var = np.array([[4.0, 9.0], [0.0, 16.0], [np.nan, 1.0], [25.0, 36.0]])
    f = FitResult(flux=np.arange(8.0).reshape(n, nb), flux_err=np.sqrt(np.nan_to_num(var)),
                  s1_flux=np.zeros((n, nb)), s1_var=np.ones((n, nb)),
                  x=np.array([10.0, 20.0, 30.0, 45.0]), y=np.array([5.0, 6.0, 7.0, 8.0]),
                  logre=lr, ee1=e1, ee2=e2, sersic_n=np.array([1.0, 1.0, 2.5, 4.0]),
                  island_id=np.array([0, 1, 1, 2]), island_conv=np.array([1, 0, 0, 1], bool),
                  island_stall=np.array([0, 1, 1, 0], bool), island_nst=np.array([3, 12, 12, 0]),
                  island_giant=np.array([0, 0, 0, 1], bool),
                  param_var=np.full((n, 7), 0.01), metrics=np.full((n, 4), 0.5),
                  flux_window=np.ones((n, nb)), island_diverged=np.array([0, 1, 1, 0], bool),
                  island_stall_dchi2=np.array([np.nan, -0.2, -0.2, np.nan]),
                  island_plateau=np.array([0, 1, 1, 0], bool), flux_var=var,
                  flux_marg_var=np.array([[0.0, 0.0], [5.0, 0.0], [1.0, 1.0], [np.nan] * 2]),
                  re_upper_pix=np.array([41.25, 41.25, 83.25, np.nan]),
                  free_n=np.array([0, 0, 1, 0], bool), grid_id=np.array([0, 0, 1, 1]))
    return f, st


def test_columns_follow_the_reference_rules():
    f, st = _fit()
    cat = pd.DataFrame(dict(ID=np.arange(4)))
    out = write_fit_columns(cat, f, ["g", "r"], x_init=[10, 20, 25.0, 45], y_init=[5, 6, 7, 8],
                            source_type=st, max_steps_main=8, max_steps=30)
    # FLUXERR = sqrt(var) only where var > 0 and finite
    np.testing.assert_allclose(out["FLUXERR_g_fit"], [2.0, np.nan, np.nan, 5.0])
    np.testing.assert_allclose(out["FLUX_r_fit"], [1, 3, 5, 7])
    # shape columns on extended rows, ELL = 1 - ab, THETA = phi - 90
    assert np.isnan(out["Re_fit"][0])
    np.testing.assert_allclose(out["Re_fit"][1:], [2.0, 3.0, 41.25], rtol=1e-5)
    np.testing.assert_allclose(out["ELL_fit"][1:], [0.5, 0.2, 0.1], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(out["THETA_fit"][1:], [-60.0, 30.0, -30.0], atol=1e-3)
    # Sersic n on DEV/SERSIC rows; frozen unless freed
    assert np.isnan(out["sersic_n_fit"][0]) and np.isnan(out["sersic_n_fit"][1])
    assert out["sersic_n_frozen"].tolist() == [False, False, False, True]
    assert out["stype_fit"].tolist() == ["star", "exp", "sersic", "dev"]
    # LM flags from the island stage; the giant counts as converged, 0 steps
    assert out["opt_converged"].tolist() == [True, False, False, True]
    assert out["opt_niters"].tolist() == [3, 12, 12, 0]
    assert out["position_stranded"].tolist() == [False, True, True, False]
    assert out["island_nsrc"].tolist() == [1, 2, 2, 1]
    # marginal errors: a zero increment copies FLUXERR, NaN FLUXERR stays NaN
    np.testing.assert_allclose(out["FLUXERR_MARG_g_fit"], [2.0, np.nan, np.nan, np.nan])
    np.testing.assert_allclose(out["FLUXERR_MARG_r_fit"][:3], [3.0, 4.0, np.sqrt(2.0)],
                               rtol=1e-6)
    np.testing.assert_allclose(out["FLUX_MARG_FRAC_r_fit"][:3], [0.0, 0.0, 1.0])
    # parameter errors: star x error = sqrt(var), shapes on extended rows
    np.testing.assert_allclose(out["x_pix_white_err"][:3], 0.1, rtol=1e-6)
    assert np.isnan(out["Re_err"][0]) and np.isfinite(out["Re_err"][1])
    assert np.isnan(out["x_pix_white_err"][3])             # giant
    assert np.isnan(out["sersic_n_err"][1]) and np.isfinite(out["sersic_n_err"][2])
    # bound flags: x moved exactly pos_max_shift; Re at 0.75 x tier
    assert out["flag_bound_stalled"].tolist() == [False, False, True, False]
    assert out["flag_shape_bound"].tolist() == [False, False, False, True]
    assert out["flag_slow_converge"].tolist() == [False, True, True, False]


def test_slow_converge_flag_without_requeue():
    assert slow_converge_flag([30, 29, 31], [1, 1, 0], None, 30).tolist() == [True, False, False]
    assert slow_converge_flag([9, 8], [1, 1], 8, 30).tolist() == [True, False]
