"""Host behaviour of the reference-shaped entry points that needs no GPU:
the per-grid operator's error contract and the resilient chunk runner's
halving on device-memory failures."""
import numpy as np
import pytest

from uniphot_amd import _lib
from uniphot_amd import island_solver as isv
from uniphot_amd.psf_polynomial import PsfPolynomial
from uniphot_amd.runner import PerGridTask, run_one_grid


def _task(**kw):
    n, nb = 5, 2
    psf = PsfPolynomial(np.ones((55, 55, 1), np.float32) / 55 ** 2, 0.0, 1.0, 0.0, 1.0, 55, 55)
    base = dict(grid_id=7, grid_row=1, grid_col=2, bandnames=["g", "r"],
                data_per_band=[np.zeros((40, 50), np.float32)] * nb,
                invvar_per_band=[np.ones((40, 50), np.float32)] * nb,
                pix_good_per_band=[np.ones((40, 50), bool)] * nb, psfs_per_band=[psf] * nb,
                psf_shape=(55, 55), grid_x0=100, grid_y0=200,
                source_indices=np.arange(10, 10 + n), pos_x_init=np.linspace(110, 140, n),
                pos_y_init=np.linspace(210, 230, n), pos_err_pix=np.full(n, 0.1),
                source_type=np.zeros(n, np.int32), logre_init=np.zeros(n),
                ee1_init=np.zeros(n), ee2_init=np.zeros(n), pos_max_shift_pix=5.0,
                bg_x_origin=0.0, bg_x_scale=1.0, bg_y_origin=0.0, bg_y_scale=1.0,
                bg_poly_order=1)
    base.update(kw)
    return PerGridTask(**base)


def test_run_one_grid_turns_failures_into_error_and_nan():
    """runner.py:756-786: any exception -> error string, NaN fluxes,
    variances and background of the right shapes, seeds echoed."""
    t = _task(data_per_band=[np.zeros((40, 50), np.float32), np.zeros((41, 50), np.float32)])
    r = run_one_grid(t)
    assert r.error is not None and "Error" in r.error
    assert r.fluxes.shape == (5, 2) and np.isnan(r.fluxes).all()
    assert r.bg_coefs.shape == (2, 3) and np.isnan(r.bg_var).all()
    assert r.pos_var.shape == (5, 2)
    np.testing.assert_array_equal(r.pos_x, t.pos_x_init)
    np.testing.assert_array_equal(r.source_indices, t.source_indices)
    assert not (r.s1_converged or r.s2_converged or r.s3_converged)
    # an unsupported request fails the same way, with its reason
    r2 = run_one_grid(_task(compute_confusion=True))
    assert "error model" in r2.error and np.isnan(r2.fluxes).all()


def test_resilient_chunk_runner_halves_on_oom(monkeypatch):
    """island_solver.py:3165-3230: a device-memory failure re-runs the chunk
    as two halves (recursively); rows come back in order, init_state is
    split with them; any other error propagates."""
    calls = []

    def fake(kern, tree, host, rows, width, max_steps, ds, ws, pad, init_state=None, **kw):
        rows = np.asarray(rows)
        calls.append(rows.tolist())
        if rows.size > 3:
            raise _lib.UphOOM(_lib.UPH_ERR_OOM, "out of memory")
        if init_state is not None:
            assert np.array_equal(init_state["nst"], rows * 10)
        return dict(p=rows[:, None].astype(np.float32), chi2=rows * 2.0, wall_s=1.0,
                    steps_run=int(rows.max()), perr=None)

    monkeypatch.setattr(isv, "run_class_chunk", fake)
    rows = np.arange(10)
    out = isv.run_class_chunk_resilient(None, None, None, rows, 10, 30, None, None, 0,
                                        init_state=dict(nst=rows * 10), log_fn=lambda m: None)
    assert out["halved"] and np.array_equal(out["p"][:, 0], rows)
    assert np.array_equal(out["chi2"], rows * 2.0) and out["perr"] is None
    assert out["wall_s"] == sum(1.0 for c in calls if len(c) <= 3)
    assert max(len(c) for c in calls if len(c) <= 3) <= 3

    def bad(*a, **k):
        raise ValueError("not a memory error")
    monkeypatch.setattr(isv, "run_class_chunk", bad)
    with pytest.raises(ValueError):
        isv.run_class_chunk_resilient(None, None, None, rows, 10, 30, None, None, 0)
