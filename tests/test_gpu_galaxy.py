"""GPU parity of extended-source (EXP/DEV/SERSIC) rendering and mixed-island
fits against the float64 oracle."""
import numpy as np
import pytest

from _scenes import mixed_classes, oracle_island, small_mixed_scene
from oracle import core as ocore
from oracle import island as oisl
from uniphot_amd.island_solver import FP64_REFIT_KAPPA

pytestmark = pytest.mark.gpu


def _psf(p):
    return dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                y_origin=p.y_origin, y_scale=p.y_scale)


# (stamp side, SERSIC n range): the three tiers, and n below the 0.8 splice
# where the mixture comes from the Tractor rows
@pytest.mark.parametrize("T,n_range", [(55, (1.0, 3.0)), (111, (1.0, 3.0)), (221, (1.0, 3.0)),
                                       (55, (0.29, 0.8))])
def test_galaxy_stamps_match_oracle(cuda, T, n_range):
    import torch
    from uniphot_amd import _lib
    from uniphot_amd.psf_polynomial import psf_device_arrays
    from uniphot_amd.sersic_table import device_table
    sc = small_mixed_scene(types=(1, 3, 3), n_range=n_range) if n_range[0] < 1 \
        else small_mixed_scene()
    c, g = psf_device_arrays(sc.psfs, T)
    n = sc.x.size
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
    out = torch.empty((n, T, T), dtype=torch.float32, device=cuda)
    L = _lib.lib()
    P = _lib.ptr
    sb = L.uph_render_scratch_bytes(T, 8)
    scr = torch.empty(sb, dtype=torch.uint8, device=cuda)
    band = (np.arange(n) % 2).astype(np.int32)
    gx0 = np.full(n, 5.0, np.float32)
    gy0 = np.full(n, 9.0, np.float32)
    tens = [up(a) for a in (c, g, band, sc.x, sc.y, gx0, gy0, sc.stype, sc.logre, sc.ee1,
                            sc.ee2, sc.sersic_n)]
    _lib.check(L.uph_render_stamps(n, T, c.shape[1], P(tens[0]), P(tens[1]),
                                   P(device_table(cuda)), *[P(t) for t in tens[2:]], 0, P(out),
                                   P(scr), sb, _lib.stream_ptr()))
    got = out.cpu().numpy()
    for i in range(n):
        ps = _psf(sc.psfs[band[i]])
        ps["coefs"] = ocore.pad_psf_coefs(ps["coefs"], T)
        ref, _ = ocore.member_stamp(ps, float(sc.x[i]), float(sc.y[i]), 5.0, 9.0,
                                    int(sc.stype[i]), float(sc.logre[i]), float(sc.ee1[i]),
                                    float(sc.ee2[i]), float(sc.sersic_n[i]))
        assert np.abs(got[i] - ref).max() < 2e-6 * np.abs(ref).max(), i
        assert abs(got[i].sum() - ref.sum()) < 1e-6


@pytest.mark.parametrize("n_range", [(1.0, 3.0), (0.4, 0.9)])
@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_mixed_island_fit_matches_oracle(cuda, precision, n_range):
    import torch
    from uniphot_amd.island_solver import make_class_kernels, psf_tree_device, run_island_batch
    from uniphot_amd.sersic_table import device_table
    sc = small_mixed_scene(n_range=n_range)
    plans, hosts, pad, d, w = mixed_classes(sc, free_n=(sc.stype == 3))
    assert any(p.kind == "mixed" for p in plans)
    batch = [(make_class_kernels(p, d.shape[0], sc.psfs[0].n_terms, precision=precision), h,
              np.arange(p.n_islands)) for p, h in zip(plans, hosts)]
    trees = {T: psf_tree_device(sc.psfs, T, cuda) for T in {p.tier for p in plans}}
    res = run_island_batch(batch, trees, 30, torch.from_numpy(d).to(cuda),
                           torch.from_numpy(w).to(cuda), precision=precision,
                           sersic_dev=device_table(cuda))
    o = oisl.LMOptions()
    for (plan, host), rr in zip(zip(plans, hosts), res):
        for r in range(plan.n_islands):
            isl = oracle_island(plan, host, r, d, w, sc.psfs)
            st = oisl.run_island(host["p0"][r], isl, o)
            if precision == "fp32":
                H, _, _, _ = oisl.hg(st["p"], isl, o)
                if oisl.kappa_scaled(H) > FP64_REFIT_KAPPA:
                    continue  # the fit refits such islands on the float64 kernels
            f, e = oisl.final_fluxes(st["p"], isl, o)
            mm = host["mmask"][r]
            assert rr["nst"][r] == st["nst"] and rr["conv"][r] == st["conv"]
            assert rr["stall"][r] == st["stall"]
            pg, po = rr["p"][r][mm], st["p"].reshape(-1, 6)[mm]
            assert np.abs(pg[:, :2] - po[:, :2]).max() < 1e-4
            keep = host["keep"][r][mm]
            assert np.abs((pg - po)[:, 2:][keep[:, 2:]]).max(initial=0) < 1e-4
            rel = np.abs(rr["fluxes"][r][:, mm] - f[:, mm]) / np.maximum(np.abs(f[:, mm]),
                                                                          e[:, mm])
            assert rel.max() < 1e-5


def test_mixed_band_split_equals_persistent(cuda):
    import torch
    from uniphot_amd.island_solver import make_class_kernels, psf_tree_device, run_island_batch
    from uniphot_amd.sersic_table import device_table
    sc = small_mixed_scene()
    plans, hosts, pad, d, w = mixed_classes(sc, free_n=(sc.stype == 3))
    batch = [(make_class_kernels(p, d.shape[0], sc.psfs[0].n_terms), h, np.arange(p.n_islands))
             for p, h in zip(plans, hosts)]
    trees = {T: psf_tree_device(sc.psfs, T, cuda) for T in {p.tier for p in plans}}
    args = (batch, trees, 30, torch.from_numpy(d).to(cuda), torch.from_numpy(w).to(cuda))
    ref = run_island_batch(*args, sersic_dev=device_table(cuda))
    print("split mixed batch", [(p.kind, p.n_islands) for p in plans], flush=True)
    got = run_island_batch(*args, sersic_dev=device_table(cuda), split=True, prerender=False)
    for ra, rb in zip(ref, got):
        for k in ("p", "fluxes", "ferr", "chi2", "lam", "nst", "conv", "stall", "n_good"):
            assert np.array_equal(ra[k], rb[k]), k


@pytest.mark.parametrize("tier", [55, 111])
def test_galaxy_prerender_matches_oracle(cuda, tier):
    """Band-split fit with the galaxy planes rendered by their own launches
    (float32 wave-level FFTs) against the float64 oracle, at the bar of the
    in-place render: native-tier galaxies (planes in the LDS pool, pad 128)
    and de Vaucouleurs galaxies of Re 3.6-5 px on the 111 stamp tier (pad
    256, planes in global scratch)."""
    import torch
    from uniphot_amd.island_solver import make_class_kernels, psf_tree_device, run_island_batch
    from uniphot_amd.sersic_table import device_table
    if tier == 55:
        sc = small_mixed_scene()
    else:
        sc = small_mixed_scene(seed=4, n_src=8, types=(0, 2), re_range=(3.6, 5.0))
    plans, hosts, pad, d, w = mixed_classes(sc, free_n=(sc.stype == 3))
    assert any(p.tier == tier and p.kind != "star" for p in plans)
    batch = [(make_class_kernels(p, d.shape[0], sc.psfs[0].n_terms), h, np.arange(p.n_islands))
             for p, h in zip(plans, hosts)]
    trees = {T: psf_tree_device(sc.psfs, T, cuda) for T in {p.tier for p in plans}}
    print("prerender mixed batch", [(p.kind, p.tier, p.n_islands) for p in plans], flush=True)
    res = run_island_batch(batch, trees, 30, torch.from_numpy(d).to(cuda),
                           torch.from_numpy(w).to(cuda), sersic_dev=device_table(cuda),
                           split=True, prerender=True)
    o = oisl.LMOptions()
    n_gal = 0
    for (plan, host), rr in zip(zip(plans, hosts), res):
        for r in range(plan.n_islands):
            isl = oracle_island(plan, host, r, d, w, sc.psfs)
            st = oisl.run_island(host["p0"][r].ravel().astype(np.float64), isl, o)
            f, e = oisl.final_fluxes(st["p"], isl, o)
            mm = host["mmask"][r]
            n_gal += int((host["st"][r][mm] != 0).sum())
            assert rr["nst"][r] == st["nst"] and rr["conv"][r] == st["conv"]
            assert rr["stall"][r] == st["stall"]
            pg, po = rr["p"][r][mm], st["p"].reshape(-1, 6)[mm]
            assert np.abs(pg[:, :2] - po[:, :2]).max() < 1e-4
            keep = host["keep"][r][mm]
            assert np.abs((pg - po)[:, 2:][keep[:, 2:]]).max(initial=0) < 1e-4
            rel = np.abs(rr["fluxes"][r][:, mm] - f[:, mm]) / np.maximum(np.abs(f[:, mm]),
                                                                          e[:, mm])
            assert rel.max() < 1e-5
    assert n_gal > 0


def test_large_island_gram_launch_matches_in_kernel_gram(cuda):
    """Every island's Gram taken by the many-workgroup Gram launch
    (threshold 0) against the evaluation's own Gram on a mixed scene, with
    the post-fit pass: same LM step counts and flags, parameters and fluxes
    at float32 rounding (the two sum the same products in the same order,
    except islands of <= 16 columns, which the in-kernel Gram splits over
    pixels)."""
    import torch
    from uniphot_amd.island_solver import make_class_kernels, psf_tree_device, run_island_batch
    from uniphot_amd.sersic_table import device_table
    sc = small_mixed_scene(seed=21, n_src=24)
    plans, hosts, pad, d, w = mixed_classes(sc, free_n=(sc.stype == 3))
    batch = [(make_class_kernels(p, d.shape[0], sc.psfs[0].n_terms), h, np.arange(p.n_islands))
             for p, h in zip(plans, hosts)]
    trees = {T: psf_tree_device(sc.psfs, T, cuda) for T in {p.tier for p in plans}}
    args = (batch, trees, 30, torch.from_numpy(d).to(cuda), torch.from_numpy(w).to(cuda))
    kw = dict(sersic_dev=device_table(cuda), split=True, prerender=True, post=True)
    print("gram launch classes", [(p.kind, p.bbox_class, p.nsrc_class, p.n_islands)
                                  for p in plans], flush=True)
    ref = run_island_batch(*args, gram_split_min=float("inf"), **kw)
    got = run_island_batch(*args, gram_split_min=0.0, **kw)
    for (plan, host), ra, rb in zip(zip(plans, hosts), ref, got):
        for k in ("nst", "conv", "stall", "n_good"):
            assert np.array_equal(ra[k], rb[k]), k
        for r in range(plan.n_islands):
            mm = host["mmask"][r]
            assert np.abs(ra["p"][r][mm] - rb["p"][r][mm]).max() < 1e-5
            fa, fb = ra["fluxes"][r][:, mm], rb["fluxes"][r][:, mm]
            scale = np.maximum(np.abs(fa), ra["ferr"][r][:, mm])
            assert (np.abs(fa - fb) / scale).max() < 1e-5
            va, vb = ra["post_var"][r][mm], rb["post_var"][r][mm]
            assert np.allclose(va, vb, rtol=1e-4, atol=1e-12, equal_nan=True)


def test_fp64_large_island_gram_launch_matches_in_kernel_gram(cuda):
    """The float64 sequence with every island's Gram on the many-workgroup
    float64 Gram launch (threshold 0) against the evaluation's own float64
    matrix-core Gram: same steps and flags, fluxes to float64 rounding
    (islands above 16 columns sum the same products in the same order;
    smaller ones are split over pixels in the evaluation)."""
    import torch
    from uniphot_amd.island_solver import make_class_kernels, psf_tree_device, run_island_batch
    from uniphot_amd.sersic_table import device_table
    sc = small_mixed_scene(seed=21, n_src=24)
    plans, hosts, pad, d, w = mixed_classes(sc, free_n=(sc.stype == 3))
    batch = [(make_class_kernels(p, d.shape[0], sc.psfs[0].n_terms, precision="fp64"), h,
              np.arange(p.n_islands)) for p, h in zip(plans, hosts)]
    trees = {T: psf_tree_device(sc.psfs, T, cuda) for T in {p.tier for p in plans}}
    args = (batch, trees, 30, torch.from_numpy(d).to(cuda), torch.from_numpy(w).to(cuda))
    kw = dict(sersic_dev=device_table(cuda), split=True, prerender=True, precision="fp64")
    ref = run_island_batch(*args, gram_split_min=float("inf"), **kw)
    got = run_island_batch(*args, gram_split_min=0.0, **kw)
    for (plan, host), ra, rb in zip(zip(plans, hosts), ref, got):
        for k in ("nst", "conv", "stall", "n_good"):
            assert np.array_equal(ra[k], rb[k]), k
        for r in range(plan.n_islands):
            mm = host["mmask"][r]
            assert np.abs(ra["p"][r][mm] - rb["p"][r][mm]).max() < 1e-5
            fa, fb = ra["fluxes"][r][:, mm], rb["fluxes"][r][:, mm]
            scale = np.maximum(np.abs(fa), ra["ferr"][r][:, mm])
            assert (np.abs(fa - fb) / scale).max() < 1e-7


@pytest.mark.parametrize("n_range", [(1.0, 3.0), (0.4, 0.9)])
def test_fp64_prerender_equals_inplace(cuda, n_range):
    """The float64 member pre-render (spectra + (member, band) planes
    launches, then the evaluation places the planes) performs the in-place
    float64 render's operations, and its Gram runs on float64 matrix cores
    (the in-place evaluation's is a VALU loop): the band-split fits take the
    same steps and flags, and agree to float64 rounding."""
    import torch
    from uniphot_amd.island_solver import make_class_kernels, psf_tree_device, run_island_batch
    from uniphot_amd.sersic_table import device_table
    sc = small_mixed_scene(n_range=n_range)
    plans, hosts, pad, d, w = mixed_classes(sc, free_n=(sc.stype == 3))
    batch = [(make_class_kernels(p, d.shape[0], sc.psfs[0].n_terms, precision="fp64"), h,
              np.arange(p.n_islands)) for p, h in zip(plans, hosts)]
    trees = {T: psf_tree_device(sc.psfs, T, cuda) for T in {p.tier for p in plans}}
    dd, ww = torch.from_numpy(d).to(cuda), torch.from_numpy(w).to(cuda)
    kw = dict(precision="fp64", sersic_dev=device_table(cuda), split=True)
    ref = run_island_batch(batch, trees, 30, dd, ww, prerender=False, **kw)
    got = run_island_batch(batch, trees, 30, dd, ww, prerender=True, **kw)
    for ra, rb in zip(ref, got):
        _same_fit(ra, rb)


def _same_fit(ra, rb):
    """Two float64 fits of one batch that differ only in float64 rounding:
    identical steps and flags, the float32 parameter state within a few
    ulps, and the fluxes (float32 outputs) and chi2 as far as one ulp of a
    position moves them (~3e-7 of a bright flux)."""
    for k in ("nst", "conv", "stall", "n_good"):
        assert np.array_equal(ra[k], rb[k]), k
    assert np.nanmax(np.abs(ra["p"] - rb["p"]), initial=0.0) <= 1e-5
    scale = np.maximum(np.abs(ra["fluxes"]), ra["ferr"])
    assert np.nanmax(np.abs(ra["fluxes"] - rb["fluxes"]) / scale, initial=0.0) <= 2e-6
    assert np.nanmax(np.abs(ra["chi2"] - rb["chi2"]) / np.abs(ra["chi2"]), initial=0.0) <= 1e-7
