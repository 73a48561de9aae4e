"""GPU checks of the batching workload: independent island cutouts (1-64
members, P(k) ~ k^-2, 15-95 px windows) fitted in band-split launch
sequences.  Float32 fluxes are compared with the float64 fit of the same
islands; the float64 path itself is pinned to the oracle by the island
tests."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cut(cuda):
    from scenes.cutouts import scene_cutouts
    print("cutout scene seed 5, 240 islands, 3 bands", flush=True)
    return scene_cutouts(seed=5, n_isl=240, n_bands=3, device="cuda", mosaic_w=4096)


def test_launch_sequences_are_independent(cut):
    """Splitting the batch into several launch sequences (a small plane
    budget) changes no result bit: islands never interact."""
    import torch
    from uniphot_amd.pipeline import FitConfig, fit_cutouts, psf_bank
    psf_of = psf_bank(cut.psfs, cut.dstack.device)
    one = fit_cutouts(cut, FitConfig(), psf_of=psf_of, sync=torch.cuda.synchronize)
    many = fit_cutouts(cut, FitConfig(), psf_of=psf_of, sync=torch.cuda.synchronize,
                       budget_bytes=2e6)
    print("launch sequences", len(one.chunks), "->", len(many.chunks), flush=True)
    assert len(one.chunks) == 1 and len(many.chunks) > 3
    assert one.fitted.all()
    for k in ("flux", "flux_err", "p", "nst", "conv", "stall"):
        assert np.array_equal(getattr(one, k), getattr(many, k), equal_nan=True), k
    assert cut.island_set.n_src.max() > 16   # large classes are in the batch


def test_cutout_fluxes_against_fp64(cut):
    from uniphot_amd.pipeline import FitConfig, cutout_accuracy
    k = cut.island_set.n_src
    single = np.where(k == 1)[0]
    acc = cutout_accuracy(cut, FitConfig(), single)
    print("single-member islands", acc, flush=True)
    # the comparison itself: float64 at its own parameters reproduces itself
    assert acc["flux_solve_fp64_self"]["max"] == 0.0
    assert acc["flux_solve"]["max"] < 1e-5
    assert acc["fit"]["p99"] < 1e-5
    blends = np.where((k >= 2) & (k <= 16))[0]
    acc = cutout_accuracy(cut, FitConfig(), blends)
    print("blends of 2-16", acc, flush=True)
    assert acc["flux_solve_fp64_self"]["max"] == 0.0
    # float32 models in a blend: the flux solve at fixed parameters carries
    # the rendering precision times the blend's conditioning
    assert acc["flux_solve"]["frac_le_1e5"] >= 0.95
    assert acc["flux_solve"]["max"] < 1e-3


def test_bf16_gram_against_fp64(cut):
    """Island Gram on bfloat16 matrix cores: after the float64 flux solve's
    refinement step, the fluxes at fixed parameters stay at the float32
    Gram's accuracy; whole fits follow the float32 ones."""
    from uniphot_amd.pipeline import FitConfig, cutout_accuracy
    k = cut.island_set.n_src
    ids = np.where(k <= 4)[0]
    acc = cutout_accuracy(cut, FitConfig(), ids, grams=("fp32", "bf16"))
    print("bf16 vs fp32 Gram", acc, flush=True)
    assert acc["flux_solve_fp64_self"]["max"] == 0.0
    b16, f32 = acc["bf16"], acc["fp32"]
    assert b16["flux_solve"]["p50"] < 1e-6
    assert b16["flux_solve"]["frac_le_1e5"] >= f32["flux_solve"]["frac_le_1e5"] - 0.05
    assert b16["same_trajectory"] >= 0.8


def test_fp64_prerender_equals_inplace_with_larger_tiers(cut):
    """The float64 member pre-render against the in-place float64 render
    on the whole cutout batch, stamp tiers 55 and 111 included: the same
    steps and flags, parameters within float32 ulps, fluxes as far as an
    ulp of a position moves them (the pre-render's tier-55 transforms run
    in LDS with paired real rows and its Gram on float64 matrix cores)."""
    import dataclasses
    import torch
    from uniphot_amd.pipeline import FitConfig, fit_cutouts, psf_bank
    psf_of = psf_bank(cut.psfs, cut.dstack.device)
    c64 = FitConfig(precision="fp64", post_fit=False)
    a = fit_cutouts(cut, dataclasses.replace(c64, galaxy_prerender=False), psf_of=psf_of,
                    sync=torch.cuda.synchronize)
    b = fit_cutouts(cut, c64, psf_of=psf_of, sync=torch.cuda.synchronize)
    tiers = sorted({p.tier for p, _ in b.classes})
    print("tiers", tiers, flush=True)
    assert 111 in tiers
    for k in ("nst", "conv", "stall"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k
    assert np.nanmax(np.abs(a.p - b.p)) <= 1e-5
    scale = np.maximum(np.abs(a.flux), a.flux_err)
    assert np.nanmax(np.abs(a.flux - b.flux) / scale) <= 2e-6
