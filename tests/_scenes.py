"""Small seeded scenes and island builders shared by the tests."""
from __future__ import annotations

import numpy as np

from oracle import island as oisl
from scenes import field as synth
from uniphot_amd.island_plan import class_host_arrays, free_n_gate, plan_island_classes
from uniphot_amd.islands import ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES, bucket_islands, build_islands
from uniphot_amd.source_model import SOURCE_STAR


def small_star_scene(seed=7, H=160, W=160, n_bands=2, n_src=14, order=2, flux_lo=3e3,
                     flux_hi=3e5, clustered=True):
    """Star-only scene with a spatially varying PSF fitted at ``order``."""
    rng = np.random.default_rng(seed)
    fields = [synth.MixturePsfField(H, W, (1.0, 2.0, 4.0), (0.7, 0.25, 0.05), amp=0.1, rng=rng)
              for _ in range(n_bands)]
    psfs = [f.fit_polynomial(55, order=order, n_grid=9) for f in fields]
    if clustered:
        par = rng.uniform(25, W - 26, (n_src // 2, 2))
        pts = np.concatenate([par, par + rng.normal(0, 3.5, par.shape)])
    else:
        pts = rng.uniform(25, W - 26, (n_src, 2))
    tx, ty = pts[:, 0], pts[:, 1]
    n = tx.size
    st = np.zeros(n, np.int32)
    flux = np.exp(rng.uniform(np.log(flux_lo), np.log(flux_hi), (n, 1))) * rng.uniform(
        0.5, 2.0, (1, n_bands))
    z = np.zeros(n, np.float32)
    truths = [synth.render_truth(H, W, fields[b], tx, ty, st, z, z, z, np.ones(n), flux[:, b])
              for b in range(n_bands)]
    skies = rng.uniform(50, 300, n_bands)
    data, invvar, good = synth._finish(rng, truths, skies)
    x, y, lr, e1, e2 = synth._seed_catalog(rng, tx, ty, st, z, z, z)
    return synth.Scene(data, invvar, good, psfs, x, y, st, lr, e1, e2, np.ones(n, np.float32),
                       tx, ty, flux, meta=dict(seed=seed))


def star_classes(sc, link_radius=6.0, pos_max_shift=5.0):
    """Islands by fixed linking radius -> (plans, hosts, pad, padded numpy stacks)."""
    n = sc.x.size
    iset = build_islands(sc.x, sc.y, sc.stype, sc.logre, sc.sersic_n, 27.5,
                         radius=np.full(n, link_radius))
    bk = bucket_islands(iset, ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES)
    plans, _ = plan_island_classes(iset, bk, sc.stype, sc.logre, sc.sersic_n, 55)
    pad = max(p.bbox_class for p in plans) + 8
    nb, H, W = sc.data.shape
    d = np.zeros((nb, H + 2 * pad, W + 2 * pad), np.float32)
    w = np.zeros_like(d)
    d[:, pad:pad + H, pad:pad + W] = sc.data
    w[:, pad:pad + H, pad:pad + W] = np.where(sc.good, sc.invvar, 0)
    free = np.zeros(n, bool)
    hosts = [class_host_arrays(p, pos_x=sc.x, pos_y=sc.y, source_type=sc.stype,
                               logre=sc.logre, ee1=sc.ee1, ee2=sc.ee2, sersic_n=sc.sersic_n,
                               free_n_all=free, pos_max_shift_pix=pos_max_shift, pad=pad)
             for p in plans]
    return plans, hosts, pad, d, w


def oracle_island(plan, host, row, d, w, psfs):
    """oracle.island.Island for row ``row`` of a class."""
    B, T = plan.bbox_class, plan.tier
    ox, oy = int(host["ox"][row]), int(host["oy"][row])
    from oracle.core import pad_psf_coefs
    ps = [dict(coefs=pad_psf_coefs(p.coefs.astype(np.float64), T), x_origin=p.x_origin,
               x_scale=p.x_scale, y_origin=p.y_origin, y_scale=p.y_scale) for p in psfs]
    return oisl.Island(d[:, oy:oy + B, ox:ox + B], w[:, oy:oy + B, ox:ox + B],
                       host["x0b"][row], host["y0b"][row], host["st"][row], host["sn"][row],
                       host["fn_flag"][row], host["mmask"][row], host["keep"][row],
                       host["lower"][row].ravel(), host["upper"][row].ravel(), ps, T)


def small_mixed_scene(seed=21, H=200, W=200, n_bands=2, n_src=16, order=2, flux_lo=5e3,
                      flux_hi=2e5, types=(0, 1, 2, 3), re_range=(1.2, 3.0),
                      n_range=(1.0, 3.0)):
    """Scene with stars and EXP/DEV/SERSIC galaxies (Re uniform in
    ``re_range`` px, SERSIC n uniform in ``n_range``)."""
    from uniphot_amd.source_model import ellipse_soft_from_re_ab_phi_np
    rng = np.random.default_rng(seed)
    fields = [synth.MixturePsfField(H, W, (1.0, 2.0, 4.0), (0.7, 0.25, 0.05), amp=0.1, rng=rng)
              for _ in range(n_bands)]
    psfs = [f.fit_polynomial(55, order=order, n_grid=9) for f in fields]
    par = rng.uniform(30, W - 31, (n_src // 2, 2))
    pts = np.concatenate([par, par + rng.normal(0, 4.0, par.shape)])
    tx, ty = pts[:, 0], pts[:, 1]
    n = tx.size
    st = np.asarray(types, np.int32)[rng.integers(len(types), size=n)]
    lr, e1, e2 = ellipse_soft_from_re_ab_phi_np(rng.uniform(*re_range, n), rng.uniform(0.5, 1, n),
                                                rng.uniform(0, 180, n))
    gal = st != SOURCE_STAR
    lr = np.where(gal, lr, 0).astype(np.float32)
    e1 = np.where(gal, e1, 0).astype(np.float32)
    e2 = np.where(gal, e2, 0).astype(np.float32)
    sn = np.where(st == 2, 4.0, np.where(st == 3, rng.uniform(*n_range, n), 1.0)).astype(np.float32)
    flux = np.exp(rng.uniform(np.log(flux_lo), np.log(flux_hi), (n, 1))) * rng.uniform(
        0.5, 2.0, (1, n_bands))
    truths = [synth.render_truth(H, W, fields[b], tx, ty, st, lr, e1, e2, sn, flux[:, b])
              for b in range(n_bands)]
    skies = rng.uniform(50, 300, n_bands)
    data, invvar, good = synth._finish(rng, truths, skies)
    x, y, lr_s, e1_s, e2_s = synth._seed_catalog(rng, tx, ty, st, lr, e1, e2)
    return synth.Scene(data, invvar, good, psfs, x, y, st, lr_s, e1_s, e2_s, sn, tx, ty, flux,
                       meta=dict(seed=seed))


def mixed_classes(sc, link_radius=7.0, pos_max_shift=5.0, free_n=None):
    n = sc.x.size
    iset = build_islands(sc.x, sc.y, sc.stype, sc.logre, sc.sersic_n, 27.5,
                         radius=np.full(n, link_radius))
    bk = bucket_islands(iset, ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES)
    plans, _ = plan_island_classes(iset, bk, sc.stype, sc.logre, sc.sersic_n, 55)
    pad = max(p.bbox_class for p in plans) + 8
    nb, H, W = sc.data.shape
    d = np.zeros((nb, H + 2 * pad, W + 2 * pad), np.float32)
    w = np.zeros_like(d)
    d[:, pad:pad + H, pad:pad + W] = sc.data
    w[:, pad:pad + H, pad:pad + W] = np.where(sc.good, sc.invvar, 0)
    free = np.zeros(n, bool) if free_n is None else free_n
    hosts = [class_host_arrays(p, pos_x=sc.x, pos_y=sc.y, source_type=sc.stype,
                               logre=sc.logre, ee1=sc.ee1, ee2=sc.ee2, sersic_n=sc.sersic_n,
                               free_n_all=free, pos_max_shift_pix=pos_max_shift, pad=pad)
             for p in plans]
    return plans, hosts, pad, d, w
