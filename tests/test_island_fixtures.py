"""Island membership and planning against the reference's own outputs
(tests/golden/island_fixtures.npz, written by make_island_fixtures.py from
the reference's islands.py / island_plan.py on seeded catalogues): radii,
labels, boxes, buckets, classes, host arrays, the free-n gate, giant rows
and the unpacked source rows must be equal, bit for bit."""
from pathlib import Path

import numpy as np
import pytest

from uniphot_amd import island_plan as ip
from uniphot_amd import islands as isl

GOLD = np.load(Path(__file__).parent / "golden" / "island_fixtures.npz", allow_pickle=False)
PAD, R_PSF, KAPPA = 392, 27.5, 0.5
TAGS = ("c1", "c2")


def g(tag, key):
    return GOLD[f"{tag}/{key}"]


def cat(tag):
    return {k: g(tag, "in_" + k) for k in ("x", "y", "st", "lr", "e1", "e2", "sn", "s1",
                                           "s1_var")}


def eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    assert np.array_equal(a, b, equal_nan=a.dtype.kind == "f"), what


def islands_of(tag):
    c = cat(tag)
    with np.errstate(all="ignore"):
        fmed = np.nanmedian(c["s1"], axis=1)
    pr, pp = isl.psf_radial_profile(g(tag, "in_stamp"))
    radii = isl.flux_aware_radii(c["st"], c["lr"], c["sn"], fmed, float(g(tag, "in_sigma_sky")),
                                 KAPPA, pr, pp, R_PSF)
    use = np.full(c["x"].size, 4.0) if tag == "c1" else radii
    return c, (pr, pp), radii, isl.build_islands(c["x"], c["y"], c["st"], c["lr"], c["sn"],
                                                  R_PSF, radius=use)


@pytest.mark.parametrize("tag", TAGS)
def test_radii_islands_and_buckets_equal_the_reference(tag):
    c, (pr, pp), radii, iset = islands_of(tag)
    eq(pr, g(tag, "prof_r"), "profile r")
    eq(pp, g(tag, "prof_p"), "profile p")
    eq(radii, g(tag, "radii"), "flux-aware radii")
    geo = isl.build_islands(c["x"], c["y"], c["st"], c["lr"], c["sn"], R_PSF)
    eq(geo.labels, g(tag, "geo_labels"), "geometric labels")
    eq(geo.bbox, g(tag, "geo_bbox"), "geometric boxes")
    eq(iset.labels, g(tag, "labels"), "labels")
    eq(iset.bbox, g(tag, "bbox"), "boxes")
    eq(iset.n_src, g(tag, "n_src"), "member counts")
    dflt = isl.bucket_islands(iset)
    eq(np.asarray(sorted(dflt.buckets), np.int64).reshape(-1, 2),
       g(tag, "default_bucket_keys"), "default keys")
    eq(dflt.giant_ids, g(tag, "default_giant_ids"), "default giants")
    bk = isl.bucket_islands(iset, isl.ISLAND_BBOX_CLASSES, isl.ISLAND_NSRC_CLASSES)
    eq(np.asarray(sorted(bk.buckets), np.int64).reshape(-1, 2), g(tag, "bucket_keys"), "keys")
    for B, N in sorted(bk.buckets):
        eq(bk.buckets[(B, N)].island_ids, g(tag, f"bucket_{B}_{N}_ids"), (B, N))
        eq(bk.buckets[(B, N)].member_idx, g(tag, f"bucket_{B}_{N}_midx"), (B, N))
    eq(bk.giant_ids, g(tag, "giant_ids"), "giants")


@pytest.mark.parametrize("tag", TAGS)
def test_classes_host_arrays_gate_and_unpack_equal_the_reference(tag):
    c, _, _, iset = islands_of(tag)
    nb = c["s1"].shape[1]
    free = ip.free_n_gate(c["st"], c["s1"], c["s1_var"], True, 30.0)
    eq(free, g(tag, "free_n"), "free-n gate")
    bk = isl.bucket_islands(iset, isl.ISLAND_BBOX_CLASSES, isl.ISLAND_NSRC_CLASSES)
    plans, capped = ip.plan_island_classes(iset, bk, c["st"], c["lr"], c["sn"], 55)
    eq(np.asarray(capped, np.int64), g(tag, "capped"), "capped giants")
    assert len(plans) == int(g(tag, "n_plans"))
    for i, pl in enumerate(plans):
        pre = f"plan{i}_"
        eq([pl.bbox_class, pl.nsrc_class, pl.tier, pl.n_gal_class], g(tag, pre + "key"), i)
        assert pl.kind == str(g(tag, pre + "kind"))
        eq(pl.isl_ids, g(tag, pre + "isl_ids"), (i, "isl_ids"))
        eq(pl.midx, g(tag, pre + "midx"), (i, "midx"))
        eq(pl.mmask, g(tag, pre + "mmask"), (i, "mmask"))
        host = ip.class_host_arrays(pl, pos_x=c["x"], pos_y=c["y"], source_type=c["st"],
                                    logre=c["lr"], ee1=c["e1"], ee2=c["e2"], sersic_n=c["sn"],
                                    free_n_all=free, pos_max_shift_pix=5.0, pad=PAD)
        for k in ("ox", "oy", "x0b", "y0b", "st", "lr", "e1", "e2", "sn", "fn_flag", "is_ext",
                  "mmask", "keep", "gsel", "gvalid", "p0", "lower", "upper"):
            eq(host[k][:64], g(tag, pre + "host_" + k), (i, k))
        # unpack on the class's first rows with the stored synthetic results
        k = g(tag, pre + "res_p").shape[0]
        import dataclasses
        sub = dataclasses.replace(pl, isl_ids=pl.isl_ids[:k], midx=pl.midx[:k],
                                  mmask=pl.mmask[:k], bbox=pl.bbox[:k], n_src_of=pl.n_src_of[:k])
        hsub = {kk: (v[:k] if isinstance(v, np.ndarray) and v.ndim and
                     v.shape[0] == pl.n_islands else v) for kk, v in host.items()}
        res = {kk: GOLD[f"{tag}/{pre}{kk}"] for kk in (
            "res_p", "res_f", "res_fe", "res_chi2", "res_conv", "res_stall", "res_nst",
            "res_perr", "res_pdchi2", "res_smetrics", "res_swin", "res_femarg", "res_div",
            "res_ngood")}
        up = ip.unpack_class_results(sub, hsub, iset, n_bands=nb, **res)
        for name in [f.split(pre + "out_")[1] for f in GOLD.files
                     if f.startswith(f"{tag}/{pre}out_")]:
            eq(up[name], g(tag, pre + "out_" + name), (i, name))


@pytest.mark.parametrize("tag", TAGS)
def test_giant_rows_equal_the_reference(tag):
    c, _, _, iset = islands_of(tag)
    rows = ip.giant_seed_rows(g(tag, "giant_in_ids"), iset, pos_x=c["x"], pos_y=c["y"],
                              logre=c["lr"], ee1=c["e1"], ee2=c["e2"], sersic_n=c["sn"],
                              n_bands=c["s1"].shape[1])
    for k in ("src", "pos_x", "pos_y", "logre", "ee1", "ee2", "sersic_n", "island_id",
              "island_nsrc"):
        got = np.concatenate([r[k] for r in rows]) if rows else np.empty(0)
        eq(got, g(tag, "giant_" + k), k)
    if tag == "c2":
        assert len(rows) >= 1   # the dense cluster
        r = rows[0]
        assert np.all(np.isnan(r["fluxes"])) and r["island_giant"].all() and r["converged"].all()
