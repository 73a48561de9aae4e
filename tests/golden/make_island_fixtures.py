"""Write tests/golden/island_fixtures.npz: inputs and outputs of the
reference's island membership and planning functions on seeded catalogues.

The functions (``tract7dt/fit/grid_gpu/islands.py``: psf_radial_profile,
flux_aware_radii, build_islands, bucket_islands; ``island_plan.py``:
free_n_gate, plan_island_classes, class_host_arrays, giant_seed_rows,
unpack_class_results) use NumPy and SciPy only; the two files are loaded
here as stand-alone modules from the reference checkout, in this
container, and never travel anywhere.  Nothing else pins their integer and
index outputs bit for bit, and identical island membership is required.

Catalogues (seeded, generated here):
  * ``c1``: 256^2 px, 40 stars + 10 exponential discs, one band; islands by
    fixed 8-px linking (radius 4) and by flux-aware radii;
  * ``c2``: 2048^2 px, 5000 sources from a Thomas process (70 % stars, 15 %
    exponential, 10 % de Vaucouleurs, 5 % Sersic with n in [0.5, 6]), 20
    bands of noisy pass-1-like fluxes, plus one dense 80-source cluster that
    makes a giant island; flux-aware radii.
Run in the container that holds the reference checkout:
  python tests/golden/make_island_fixtures.py
"""
from __future__ import annotations

import dataclasses
import importlib.util
import sys
from pathlib import Path

import numpy as np

REF = Path("/root/reference/tract7dt/fit/grid_gpu")
OUT = Path(__file__).with_name("island_fixtures.npz")
PAD = 392          # max bbox class + 8
R_PSF = 27.5
KAPPA = 0.5


def load(name):
    spec = importlib.util.spec_from_file_location(f"ref_{name}", REF / f"{name}.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[spec.name] = mod
    spec.loader.exec_module(mod)
    return mod


def psf_stamp(T=55, sig=(1.0, 2.0, 4.0), w=(0.7, 0.25, 0.05)):
    from scipy.special import erf
    edges = np.arange(T + 1) - T // 2 - 0.5
    img = np.zeros((T, T))
    for s, a in zip(sig, w):
        c = np.diff(0.5 * erf(edges / (np.sqrt(2) * s)))
        img += a * np.outer(c, c)
    return img / img.sum()


def shapes(rng, n):
    re = rng.uniform(1.0, 4.0, n)
    q = rng.uniform(0.4, 1.0, n)
    phi = np.radians(rng.uniform(0.0, 180.0, n))
    e = (1 - q) / (1 + q)
    ee = -np.log(1 - e)
    return (np.log(re).astype(np.float32), (ee * np.cos(-2 * phi)).astype(np.float32),
            (ee * np.sin(-2 * phi)).astype(np.float32))


def catalog_c1(rng):
    n = 50
    x = rng.uniform(0, 255, n)
    y = rng.uniform(0, 255, n)
    st = np.r_[np.zeros(40), np.ones(10)].astype(np.int32)
    lr, e1, e2 = shapes(rng, n)
    lr, e1, e2 = [np.where(st > 0, a, 0).astype(np.float32) for a in (lr, e1, e2)]
    sn = np.ones(n, np.float32)
    flux = np.exp(rng.uniform(np.log(1e3), np.log(1e6), (n, 1)))
    sig = np.sqrt(125.0 * 30.0)
    s1 = flux + rng.normal(0, sig, flux.shape)
    return dict(x=x, y=y, st=st, lr=lr, e1=e1, e2=e2, sn=sn, s1=s1.astype(np.float32),
                s1_var=np.full(s1.shape, sig ** 2, np.float32), sigma_sky=np.sqrt(125.0))


def catalog_c2(rng):
    W = 2048
    n_par = 1725
    px, py = rng.uniform(0, W - 1, n_par), rng.uniform(0, W - 1, n_par)
    kids = 1 + rng.poisson(1 / 0.3 - 1, n_par)
    xs = np.concatenate([px[i] + rng.normal(0, 6, k) * (np.arange(k) > 0)
                         for i, k in enumerate(kids)])
    ys = np.concatenate([py[i] + rng.normal(0, 6, k) * (np.arange(k) > 0)
                         for i, k in enumerate(kids)])
    ok = (xs >= 0) & (xs <= W - 1) & (ys >= 0) & (ys <= W - 1)
    x, y = xs[ok][:4920], ys[ok][:4920]
    # a dense cluster: one island of 80 members (a giant)
    x = np.r_[x, 1000 + rng.normal(0, 12, 80)]
    y = np.r_[y, 1000 + rng.normal(0, 12, 80)]
    n = x.size
    u = rng.uniform(0, 1, n)
    st = np.where(u < 0.7, 0, np.where(u < 0.85, 1, np.where(u < 0.95, 2, 3))).astype(np.int32)
    lr, e1, e2 = shapes(rng, n)
    gal = st > 0
    lr, e1, e2 = [np.where(gal, a, 0).astype(np.float32) for a in (lr, e1, e2)]
    sn = np.where(st == 2, 4.0, np.where(st == 3, rng.uniform(0.5, 6.0, n), 1.0)).astype(np.float32)
    nb = 20
    sky = rng.uniform(50, 500, nb)
    flux = np.exp(rng.uniform(np.log(1e2), np.log(1e6), (n, 1))) * rng.uniform(0.5, 2, (n, nb))
    sig = np.sqrt((sky + 25.0) * 30.0)[None, :] * np.ones((n, 1))
    s1 = flux + rng.normal(0, 1, flux.shape) * sig
    s1[rng.uniform(0, 1, s1.shape) < 0.002] = np.nan   # a few unconstrained bands
    return dict(x=x, y=y, st=st, lr=lr, e1=e1, e2=e2, sn=sn, s1=s1.astype(np.float32),
                s1_var=(sig ** 2).astype(np.float32),
                sigma_sky=float(np.median(np.sqrt(sky + 25.0))))


def fake_results(rng, plan, host, nb):
    n_c, N = plan.n_islands, plan.nsrc_class
    p = host["p0"] + rng.normal(0, 0.05, host["p0"].shape).astype(np.float32)
    perr = np.abs(rng.normal(0, 0.01, (n_c, N, 13))).astype(np.float32)
    perr[..., 6] = rng.normal(0, 0.001, (n_c, N)).astype(np.float32)
    perr[rng.uniform(0, 1, (n_c, N)) < 0.05, 3] = -1.0     # degenerate variances
    return dict(
        res_p=p, res_f=rng.normal(1e3, 300, (n_c, nb, N)).astype(np.float32),
        res_fe=np.abs(rng.normal(30, 5, (n_c, nb, N))).astype(np.float32),
        res_chi2=np.abs(rng.normal(1e4, 1e3, n_c)),
        res_conv=rng.uniform(0, 1, n_c) < 0.9, res_stall=rng.uniform(0, 1, n_c) < 0.1,
        res_nst=rng.integers(1, 30, n_c).astype(np.int32), res_perr=perr,
        res_pdchi2=rng.normal(0, 1, n_c).astype(np.float32),
        res_smetrics=rng.uniform(0, 2, (n_c, N, 4)).astype(np.float32),
        res_swin=rng.normal(900, 100, (n_c, nb, N)).astype(np.float32),
        res_femarg=np.abs(rng.normal(0, 10, (n_c, nb, N))).astype(np.float32),
        res_div=rng.uniform(0, 1, n_c) < 0.05,
        res_ngood=rng.integers(500, 50000, n_c).astype(np.int64))


def record(out, tag, cat, isl_mod, plan_mod, rng, link_radius=None):
    nb = cat["s1"].shape[1]
    out[f"{tag}/in_x"], out[f"{tag}/in_y"] = cat["x"], cat["y"]
    for k in ("st", "lr", "e1", "e2", "sn", "s1", "s1_var"):
        out[f"{tag}/in_{k}"] = cat[k]
    out[f"{tag}/in_sigma_sky"] = np.float64(cat["sigma_sky"])
    stamp = psf_stamp()
    out[f"{tag}/in_stamp"] = stamp
    pr, pp = isl_mod.psf_radial_profile(stamp)
    out[f"{tag}/prof_r"], out[f"{tag}/prof_p"] = pr, pp
    with np.errstate(all="ignore"):
        fmed = np.nanmedian(cat["s1"], axis=1)
    radii = isl_mod.flux_aware_radii(cat["st"], cat["lr"], cat["sn"], fmed, cat["sigma_sky"],
                                     KAPPA, pr, pp, R_PSF)
    out[f"{tag}/radii"] = radii
    geo = isl_mod.build_islands(cat["x"], cat["y"], cat["st"], cat["lr"], cat["sn"], R_PSF)
    out[f"{tag}/geo_labels"], out[f"{tag}/geo_bbox"] = geo.labels, geo.bbox
    use = radii if link_radius is None else np.full(cat["x"].size, link_radius)
    iset = isl_mod.build_islands(cat["x"], cat["y"], cat["st"], cat["lr"], cat["sn"], R_PSF,
                                 radius=use)
    out[f"{tag}/labels"], out[f"{tag}/bbox"], out[f"{tag}/n_src"] = \
        iset.labels, iset.bbox, iset.n_src
    # the function's default ladders, then the island stage's (which the
    # class planning below uses; island_stage.py:133-134)
    dflt = isl_mod.bucket_islands(iset)
    out[f"{tag}/default_bucket_keys"] = np.asarray(sorted(dflt.buckets), np.int64).reshape(-1, 2)
    out[f"{tag}/default_giant_ids"] = dflt.giant_ids
    bk = isl_mod.bucket_islands(iset, (32, 48, 64, 96, 128, 192, 256, 384),
                                (1, 2, 4, 8, 16, 32, 64))
    keys = sorted(bk.buckets)
    out[f"{tag}/bucket_keys"] = np.asarray(keys, np.int64).reshape(-1, 2)
    for B, N in keys:
        b = bk.buckets[(B, N)]
        out[f"{tag}/bucket_{B}_{N}_ids"] = b.island_ids
        out[f"{tag}/bucket_{B}_{N}_midx"] = b.member_idx
    out[f"{tag}/giant_ids"] = bk.giant_ids
    free = plan_mod.free_n_gate(cat["st"], cat["s1"], cat["s1_var"], True, 30.0)
    out[f"{tag}/free_n"] = free
    plans, capped = plan_mod.plan_island_classes(iset, bk, cat["st"], cat["lr"], cat["sn"], 55)
    out[f"{tag}/capped"] = np.asarray(capped, np.int64)
    out[f"{tag}/n_plans"] = np.int64(len(plans))
    host_keys = ("ox", "oy", "x0b", "y0b", "st", "lr", "e1", "e2", "sn", "fn_flag", "is_ext",
                 "mmask", "keep", "gsel", "gvalid", "p0", "lower", "upper")
    unpack_keys = ("src", "pos_x", "pos_y", "logre", "ee1", "ee2", "sersic_n", "fluxes",
                   "flux_err", "chi2", "chi2_nu", "n_steps", "converged", "stalled",
                   "diverged", "stall_dchi2", "railed", "free_n", "island_id", "island_nsrc",
                   "island_giant", "re_upper_pix", "pos_x_err", "pos_y_err", "re_err",
                   "ell_err", "theta_err", "sersic_n_err", "rchisq", "fracflux", "fracin",
                   "fracmasked", "flux_win", "flux_marg_var")
    for i, pl in enumerate(plans):
        pre = f"{tag}/plan{i}_"
        out[pre + "key"] = np.array([pl.bbox_class, pl.nsrc_class, pl.tier, pl.n_gal_class])
        out[pre + "kind"] = np.array(pl.kind)
        out[pre + "isl_ids"], out[pre + "midx"], out[pre + "mmask"] = \
            pl.isl_ids, pl.midx, pl.mmask
        host = plan_mod.class_host_arrays(pl, pos_x=cat["x"], pos_y=cat["y"],
                                          source_type=cat["st"], logre=cat["lr"],
                                          ee1=cat["e1"], ee2=cat["e2"], sersic_n=cat["sn"],
                                          free_n_all=free, pos_max_shift_pix=5.0, pad=PAD)
        for k in host_keys:   # the first 64 rows of each class
            out[pre + "host_" + k] = host[k][:64]
        # unpack on the class's first rows only (keeps the file small)
        k_rows = min(pl.n_islands, 4)
        sub = dataclasses.replace(pl, isl_ids=pl.isl_ids[:k_rows], midx=pl.midx[:k_rows],
                                  mmask=pl.mmask[:k_rows], bbox=pl.bbox[:k_rows],
                                  n_src_of=pl.n_src_of[:k_rows])
        hsub = {k: (v[:k_rows] if isinstance(v, np.ndarray) and v.ndim and
                    v.shape[0] == pl.n_islands else v) for k, v in host.items()}
        res = fake_results(rng, sub, hsub, nb)
        for k, v in res.items():
            out[pre + k] = v
        up = plan_mod.unpack_class_results(sub, hsub, iset, n_bands=nb, **res)
        for k in unpack_keys:
            out[pre + "out_" + k] = np.asarray(up[k])
    gids = np.concatenate([bk.giant_ids.astype(np.int64), np.asarray(capped, np.int64)])
    rows = plan_mod.giant_seed_rows(gids, iset, pos_x=cat["x"], pos_y=cat["y"],
                                    logre=cat["lr"], ee1=cat["e1"], ee2=cat["e2"],
                                    sersic_n=cat["sn"], n_bands=nb)
    out[f"{tag}/giant_in_ids"] = gids
    for k in ("src", "pos_x", "pos_y", "logre", "ee1", "ee2", "sersic_n", "island_id",
              "island_nsrc"):
        out[f"{tag}/giant_{k}"] = (np.concatenate([r[k] for r in rows]) if rows
                                   else np.empty(0))


def main():
    isl_mod = load("islands")
    plan_mod = load("island_plan")
    out = {}
    rng = np.random.default_rng(20261201)
    record(out, "c1", catalog_c1(rng), isl_mod, plan_mod, rng, link_radius=4.0)
    rng = np.random.default_rng(20261202)
    record(out, "c2", catalog_c2(rng), isl_mod, plan_mod, rng)
    np.savez_compressed(OUT, **out)
    print(OUT, len(out), OUT.stat().st_size)


if __name__ == "__main__":
    main()
