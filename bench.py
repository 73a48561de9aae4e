"""Forced-photometry fit benchmark on one MI355X.

One step = the whole island-mode fit of one synthetic field with its inputs
already resident in HBM: pass-1 grid-band linear solves, island
decomposition, the band-split island VarPro-LM (member pre-render,
(island, band) evaluations and per-island LM steps), pass-2 linear solves.
Metric: sources fitted per second (whole job).  Also reported: the
dominant kernel's roofline position -- algorithmic HBM bytes and fp32 flops
per launch over its measured average duration (HIP events on its stream),
placed against whichever of the 8 TB/s HBM and 157.3 TFLOP/s fp32 peaks it
is closer to -- and a CPU baseline (the CPU restatement of the same method,
float32 pixels and OpenMP, timed on a stratified sample of the same
workload).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c2-stars|c1|c3|c4|c5]
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X dense fp32, vector and matrix alike
FP64_PEAK_TFLOPS = 78.6   # MI355X dense fp64, vector and matrix alike


class KernelTimer:
    """HIP events around each library launch on torch's current stream (the
    stream the C ABI is handed)."""

    def __init__(self):
        self.recs = []
        self.deferred = []
        self.enabled = True

    def begin(self, name):
        import torch
        if not self.enabled:
            return None
        e0 = torch.cuda.Event(enable_timing=True)
        e0.record()
        return (name, e0)

    def end(self, ev, info):
        import torch
        if ev is None:
            return
        if not info:               # a launch with no cost record (the persistent batch)
            return
        e1 = torch.cuda.Event(enable_timing=True)
        e1.record()
        self.recs.append((info.get("name") or ev[0], ev[1], e1, info))

    def record(self, name, e0, e1, info):
        """An event pair recorded inside the library (per-kernel launches)."""
        self.recs.append((name, e0, e1, info))

    def defer(self, fn):
        """Bookkeeping that records event pairs, run when the timer reports
        rather than inside the timed steps."""
        self.deferred.append(fn)

    def _flush(self):
        while self.deferred:
            self.deferred.pop(0)()

    def per_launch(self):
        import torch
        torch.cuda.synchronize()
        self._flush()
        rows = []
        for name, e0, e1, info in self.recs:
            rows.append((name, e0.elapsed_time(e1), {k: v for k, v in info.items()
                                                     if isinstance(v, (int, float, str))}))
        return rows

    def summary(self):
        import torch
        torch.cuda.synchronize()
        self._flush()
        out = {}
        for name, e0, e1, info in self.recs:
            ms = e0.elapsed_time(e1)
            d = out.setdefault(name, dict(launches=0, ms=0.0, bytes=0.0, flops=0.0))
            d["launches"] += 1
            d["ms"] += ms
            b, f = algorithmic_cost(info)
            d["bytes"] += b
            d["flops"] += f
        return out


def algorithmic_cost(info):
    """Algorithmic HBM bytes and flops of one launch.
    island: sum_i nb * A_i * 8 * (2 s_i + 2) (data + weight of the island box
            per evaluation pass: initial chi2, 2 per LM step, final fluxes)
            + PSF coefficients once; flops per band per step 4N R + 24 N^2 A
            with R = 2 T^2 n_t + 28 T^2 (star members).
    island_eval: one band-split evaluation launch: nb * A_i * 8 for every
            island box it evaluates (bytes computed by the island facade).
    band:   data + invvar + good per window pixel (9 B) + stamps read.
    render: PSF coefficients + stamp written per stamp."""
    k = info["kind"]
    if k == "island":
        host, rows = info["host"], info["rows"]
        nst = info["outs"]["nst"].cpu().numpy().astype(np.float64)
        area = host["area"][rows].astype(np.float64)
        nb = info["nb"]
        by = float(np.sum(nb * area * 8.0 * (2.0 * nst + 2.0)))
        nreal = host["mmask"][rows].sum(axis=1).astype(np.float64)
        T = 55.0
        R = 2 * T * T * 15 + 28 * T * T
        fl = float(np.sum(nb * (2 * nst + 2) * (4 * nreal * R / 2 + 24 * nreal ** 2 * area)))
        return by, fl
    if k == "island_batch":
        nst = info["outs"]["nst"].cpu().numpy().astype(np.float64)
        area = info["area"].astype(np.float64)
        nb = info["nb"]
        nreal = info["nreal"].astype(np.float64)
        by = float(np.sum(nb * area * 8.0 * (2.0 * nst + 2.0)))
        T = 55.0
        R = 2 * T * T * 15 + 28 * T * T
        fl = float(np.sum(nb * (2 * nst + 2) * (4 * nreal * R / 2 + 24 * nreal ** 2 * area)))
        return by, fl
    if k == "island_eval":
        return info["bytes"], info.get("flops", 0.0)
    if k == "band":
        return info["pix"] * 9.0 + info["n_src"] * info["T"] ** 2 * 4.0, 0.0
    if k == "render":
        T = info["T"]
        return info["n"] * T * T * 4.0, info["n"] * (2.0 * T * T * info["n_terms"] + 28 * T * T)
    return 0.0, 0.0


def measured_traffic(workload, kernel):
    """HBM bytes per launch of ``kernel`` from the committed counter summary
    (FETCH_SIZE and WRITE_SIZE runs, each scaled by what the same counter
    read for a plain copy of known size), or None if there is none."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    fetch_scale = t["copy_bytes"] / (t["copy_fetch_kib"] * 1024.0)
    write_scale = t["copy_bytes"] / (t["copy_write_kib"] * 1024.0)
    return (t["fetch_kib"] * 1024.0 * fetch_scale + t["write_kib"] * 1024.0 * write_scale)


def make_scene(config):
    from scenes import field as synth
    if config == "c2-stars":
        return synth.scene_c2(star_only=True), dict(n_rows=8, n_cols=8)
    if config == "c2":
        return synth.scene_c2(), dict(n_rows=8, n_cols=8)
    if config == "c1":
        return synth.scene_c1(), dict(n_rows=2, n_cols=2, link_radius=4.0)
    raise SystemExit(f"unknown config {config}")


def _cpu_threads():
    """Threads of the CPU baseline: the process's CPU share (OMP_NUM_THREADS
    when set, else the affinity mask)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, min(n, int(env))) if env and env.isdigit() else n


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _device_cutouts(art, picks):
    """The island cutouts the GPU fit used, rebuilt by its own mosaic code
    (window of the scene-subtracted stack plus the member add-back at the
    class tier): {(class, row): (data, weights)}."""
    from uniphot_amd.scene import island_mosaic
    by_cls = {}
    for pi, r in picks:
        by_cls.setdefault(pi, []).append(int(r))
    keys = sorted(by_cls)
    chunk = [(None, art["classes"][pi][1], np.asarray(by_cls[pi])) for pi in keys]
    plan_of = {id(h): p for p, h in art["classes"]}
    dmos, wmos, moved = island_mosaic(chunk, plan_of, art["dstack"], art["wstack"], art["scene"])
    dm, wm = dmos.cpu().numpy(), wmos.cpu().numpy()
    out = {}
    for pi, (_, h2, rows) in zip(keys, moved):
        B = art["classes"][pi][0].bbox_class
        for r in rows:
            mx, my = int(h2["ox"][r]), int(h2["oy"][r])
            out[(pi, int(r))] = (dm[:, my:my + B, mx:mx + B], wm[:, my:my + B, mx:mx + B])
    return out


def _island_items(art, scene, psfs, picks, cuts=None):
    """CPU-restatement island inputs for (class, row) picks: the window
    cutouts with the scene subtraction's member add-back (restated in
    float64), or the given ``cuts`` {(class, row): (data, weights)}."""
    from oracle import core as ocore
    from oracle import island as oisl
    d, w = art["dstack"], art["wstack"]
    items, meta = [], []
    for pi, r in picks:
        plan, host = art["classes"][pi]
        B, T = plan.bbox_class, plan.tier
        ox, oy = int(host["ox"][r]), int(host["oy"][r])
        cut = d[:, oy:oy + B, ox:ox + B]
        wcut = w[:, oy:oy + B, ox:ox + B]
        mm = host["mmask"][r]
        if cuts is not None:
            cut, wcut = cuts[(pi, int(r))]
        elif scene is not None:
            mem = plan.midx[r][mm]
            ps = [dict(q, coefs=ocore.pad_psf_coefs(q["coefs"], T)) for q in psfs]
            cut = cut + oisl.member_addback(ps, host["x0b"][r], host["y0b"][r], B,
                                            scene.x[mem], scene.y[mem], scene.stype[mem],
                                            scene.flux[mem])
        items.append(dict(data=cut, w=wcut, x0b=host["x0b"][r],
                          y0b=host["y0b"][r], st=host["st"][r], sn=host["sn"][r],
                          fn=host["fn_flag"][r], mmask=mm, keep=host["keep"][r],
                          p0=host["p0"][r].ravel(), lower=host["lower"][r].ravel(),
                          upper=host["upper"][r].ravel(), T=T))
        meta.append((pi, plan.midx[r][mm], int(mm.sum())))
    return items, meta


def cpu_baseline(fld, res, sc, cfg, n_islands=120, n_problems=48, seed=0, oracle_cmp=None):
    """The CPU restatement (oracle/cpu_fit.cpp: float32 pixels, float64
    solves, OpenMP over islands and grid-band problems) timed on a sample
    of the same field: ``n_islands`` islands stratified by class (each
    class in proportion to its island count, at least one) fitted on all
    threads, each island's time measured on its own thread, extrapolated
    per class; ``n_problems`` random grid-band problems (stamp render +
    solve) on one thread, extrapolated to the two linear passes.  value =
    sources/s on all threads (per-island thread times over the thread
    count); one_core = the same on one thread.  ``oracle_cmp`` (a list)
    collects float64 fits of 150 random islands, on the cutouts the GPU fit
    used, by the float64 build of the same source, which reproduces the
    float64 oracle (tests/test_oracle_cpu.py)."""
    from oracle import core as ocore
    from oracle import cpu as ocpu
    from oracle import island as oisl
    from oracle import linear as olin
    from uniphot_amd import forced_phot_sparse as fps
    rng = np.random.default_rng(seed)
    art_dev = res.record["_artifacts"]
    art = dict(art_dev, dstack=art_dev["dstack"].cpu().numpy(),
               wstack=art_dev["wstack"].cpu().numpy())
    scene = art.get("scene")
    psfs = [dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                 y_origin=p.y_origin, y_scale=p.y_scale) for p in fld.psfs]
    C = _cpu_threads()
    o = oisl.LMOptions(reg=cfg.reg)
    ocore.sersic_table()  # cached before any timing
    counts = np.array([plan.n_islands for plan, _ in art["classes"]])
    n_tot = int(counts.sum())
    picks = []
    for pi, nc in enumerate(counts):
        k = int(min(nc, max(1, round(n_islands * nc / max(n_tot, 1)))))
        picks += [(pi, int(r)) for r in rng.choice(nc, k, replace=False)]
    items, meta = _island_items(art, scene, psfs, picks)
    # each sampled island runs at most `cap` LM steps; its time is scaled to
    # the step count the GPU fit took, counting the initial chi2 and the
    # final fluxes as one step together
    cap = 3
    nst_gpu = np.array([max(int(res.island_nst[src[0]]), 0) for _, src, _ in meta])
    out = []
    t0 = time.perf_counter()
    for s_cap in sorted(set(np.minimum(nst_gpu, cap).tolist())):
        sel = np.flatnonzero(np.minimum(nst_gpu, cap) == s_cap)
        rr = ocpu.fit_islands([items[i] for i in sel], psfs, dataclasses.replace(o, max_steps=s_cap),
                              n_threads=C)
        out += [(int(i), r) for i, r in zip(sel, rr)]
    wall_isl = time.perf_counter() - t0
    per_class = {}
    for i, r in out:
        scale = (nst_gpu[i] + 1.0) / (min(nst_gpu[i], cap) + 1.0)
        per_class.setdefault(meta[i][0], []).append(r["sec"] * scale)
    t1_isl = float(sum(counts[pi] * np.mean(v) for pi, v in per_class.items()))
    # linear passes: random (grid, band) problems, one thread
    grids = art["grids"]
    data = fld.data.cpu().numpy()
    iv = fld.invvar.cpu().numpy()
    gd = fld.good.cpu().numpy().astype(bool)
    geo = (0.5 * (fld.W - 1), 0.5 * fld.W, 0.5 * (fld.H - 1), 0.5 * fld.H)
    t_lin = 0.0
    for _ in range(n_problems):
        g = grids[rng.integers(len(grids))]
        b = int(rng.integers(fld.nb))
        fi = g.fit_indices
        T = art["gtier"][g.grid_id]
        xs, ys = sc.x[fi], sc.y[fi]
        t0 = time.perf_counter()
        all_star = bool(np.all(sc.stype[fi] == 0))
        nbr, msk = fps.build_neighbor_list(xs, ys, T, T)
        st = ocpu.render_stamps([psfs[b]], T, 0, xs, ys, g.x0_pad, g.y0_pad, sc.stype[fi],
                                sc.logre[fi], sc.ee1[fi], sc.ee2[fi], sc.sersic_n[fi],
                                renorm=all_star, n_threads=1)
        ix, iy = fps.stamp_top_left(xs, ys, g.x0_pad, g.y0_pad, T, T)
        sl = (b, slice(g.y0_pad, g.y1_pad), slice(g.x0_pad, g.x1_pad))
        ocpu.band_solve([dict(data=data[sl], invvar=iv[sl], good=gd[sl], stamps=st, ix=ix,
                              iy=iy, nbr=np.where(msk, nbr, -1), gx0=g.x0_pad, gy0=g.y0_pad)],
                        geo, k_max=nbr.shape[1], reg=cfg.reg, n_threads=1)
        t_lin += time.perf_counter() - t0
    t1_lin = t_lin / n_problems * 2 * len(grids) * fld.nb
    n_src = int(sc.x.size)
    one = n_src / (t1_isl + t1_lin)
    if oracle_cmp is not None:
        order = [(pi, r) for pi, (plan, _) in enumerate(art["classes"])
                 for r in range(plan.n_islands)]
        sel = rng.choice(len(order), min(150, len(order)), replace=False)
        picks64 = [order[i] for i in sel]
        items64, meta64 = _island_items(art, scene, psfs, picks64,
                                        cuts=_device_cutouts(art_dev, picks64))
        fits = ocpu.fit_islands(items64, psfs, o, n_threads=C, precision="fp64")
        for it, (_, src, _), r in zip(items64, meta64, fits):
            mm = np.asarray(it["mmask"], bool)
            oracle_cmp.append((src, r["fluxes"][:, mm].T, r["ferr"][:, mm].T,
                               r["p"].reshape(-1, 6)[mm], r["nst"], r["conv"]))
    return dict(value=one * C, unit="sources/s", cores=C, kind="port",
                one_core=one, cpu_model=_cpu_model(),
                sample=(f"CPU restatement of the per-island VarPro-LM and the grid-band solve "
                        f"(oracle/cpu_fit.cpp, g++ -O3 x86-64-v3, float32 pixels, float64 "
                        f"solves, OpenMP): {len(items)} islands stratified over "
                        f"{len(counts)} classes on {C} threads in {wall_isl:.1f}s (per-island "
                        f"thread time {sum(r['sec'] for _, r in out):.1f}s for at most {cap} LM "
                        f"steps each, scaled to the fit's step count and per class to {n_tot} "
                        f"islands: {t1_isl:.1f} core-s) + {n_problems} random "
                        f"grid-band problems (stamps + solve) on 1 thread in {t_lin:.1f}s "
                        f"(extrapolated to 2 passes x {len(grids)} grids x {fld.nb} bands: "
                        f"{t1_lin:.1f} core-s); value = one-core rate x {C} threads"))


def oracle_accuracy(res, cmp):
    """GPU island fit (island VarPro fluxes and positions of ``res``)
    against float64 fits of random islands by the float64 build of the CPU
    restatement (which reproduces the float64 oracle's fits): per source, max over bands of |f - f_o| /
    max(|f_o|, sigma_o), position error in px, and the share of islands
    with the oracle's step count."""
    from uniphot_amd.pipeline import _rel_stats
    if not cmp:
        return dict(n=0)
    src = np.concatenate([c[0] for c in cmp])
    fo = np.concatenate([c[1] for c in cmp])
    eo = np.concatenate([c[2] for c in cmp])
    po = np.concatenate([c[3] for c in cmp])
    st = _rel_stats(res.island_flux[src].astype(np.float64), fo, eo)
    st["position_max_px"] = float(np.max(np.hypot(res.x[src] - po[:, 0], res.y[src] - po[:, 1])))
    same = [int(np.all(res.island_nst[c[0]] == c[4])) for c in cmp]
    st["same_steps"] = float(np.mean(same))
    st["islands"] = len(cmp)
    return st


def field_accuracy(fld, res, cat, cfg, n_sample, sync, seed=2):
    """Float32 island fit (with its float64 refit) against the float64
    kernels on ``n_sample`` random fitted islands of the field."""
    from uniphot_amd.pipeline import island_accuracy
    art = res.record["_artifacts"]
    iset = art["island_set"]
    rng = np.random.default_rng(seed)
    fit_ids = np.setdiff1d(np.arange(iset.n_islands), art["giants"])
    sample = np.sort(rng.choice(fit_ids, min(n_sample, fit_ids.size), replace=False))
    acc = island_accuracy(fld.psf, fld.nb, fld.n_terms, fld.native, art["dstack"],
                          art["wstack"], art["pad"], iset, sample, cat, art["free_n"], cfg,
                          sync=sync, scene=art["scene"],
                          flux_scale=(res.s1_flux, np.sqrt(np.maximum(res.s1_var, 0.0))))
    acc.update(acc.pop("fp32"))
    return acc


def roofline_of(ks, config):
    """Dominant kernel (most total time; the split sequence's own record is
    for reference only) placed against the HBM or fp32 roof it is closer to."""
    kern = max((k for k in ks if k != "island_split_sequence"), key=lambda k: ks[k]["ms"])
    isl = ks[kern]
    avg_s = isl["ms"] / 1e3 / max(isl["launches"], 1)
    per_launch_bytes = isl["bytes"] / max(isl["launches"], 1)
    achieved = per_launch_bytes / avg_s / 1e9 if avg_s > 0 else 0.0
    tflops = isl["flops"] / max(isl["ms"] / 1e3, 1e-12) / 1e12
    peak = (FP64_PEAK_TFLOPS if kern.endswith("_fp64") or kern.startswith("pre64")
            else FP32_PEAK_TFLOPS)
    compute = tflops / peak > achieved / HBM_PEAK_GBS
    traffic = measured_traffic(config, kern)
    if compute:
        roof = {"bound": "mfma", "achieved": round(tflops, 4), "peak": peak,
                "unit": "TFLOP/s", "frac": round(tflops / peak, 6),
                "traffic": traffic, "hbm_gbs": round(achieved, 3),
                "hbm_frac": round(achieved / HBM_PEAK_GBS, 6)}
    else:
        roof = {"bound": "hbm", "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6), "traffic": traffic,
                "fp32_tflops": round(tflops, 4),
                "fp32_frac": round(tflops / FP32_PEAK_TFLOPS, 6)}
    roof.update({"kernel": kern, "avg_launch_ms": round(avg_s * 1e3, 4),
                 "launches": isl["launches"], "algorithmic_bytes_per_launch": per_launch_bytes,
                 "algorithmic_flops_per_launch": isl["flops"] / max(isl["launches"], 1)})
    return roof


def cpu_baseline_cutouts(cut, res, budget_s=12.0, seed=0):
    """Float64 oracle on random islands of the cutout batch (bounded time),
    extrapolated to seconds per source."""
    from oracle import island as oisl
    from oracle import core as ocore
    rng = np.random.default_rng(seed)
    d = cut.dstack
    w = cut.wstack
    o = oisl.LMOptions()
    psfs = [dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                 y_origin=p.y_origin, y_scale=p.y_scale) for p in cut.psfs]
    items = [(pi, r) for pi, (plan, host) in enumerate(res.classes)
             for r in range(plan.n_islands)]
    order = rng.permutation(len(items))
    t0 = time.perf_counter()
    n_src = n_isl = 0
    for q in order:
        if time.perf_counter() - t0 > budget_s:
            break
        plan, host = res.classes[items[q][0]]
        r = items[q][1]
        B, T = plan.bbox_class, plan.tier
        ox, oy = int(host["ox"][r]), int(host["oy"][r])
        ps = [dict(p, coefs=ocore.pad_psf_coefs(p["coefs"], T)) for p in psfs]
        dw = d[:, oy:oy + B, ox:ox + B].cpu().numpy()
        ww = w[:, oy:oy + B, ox:ox + B].cpu().numpy()
        isl = oisl.Island(dw, ww, host["x0b"][r], host["y0b"][r], host["st"][r], host["sn"][r],
                          host["fn_flag"][r], host["mmask"][r], host["keep"][r],
                          host["lower"][r].ravel(), host["upper"][r].ravel(), ps, T)
        st = oisl.run_island(host["p0"][r].ravel().astype(np.float64), isl, o)
        oisl.final_fluxes(st["p"], isl, o)
        n_src += int(host["mmask"][r].sum())
        n_isl += 1
    el = time.perf_counter() - t0
    return dict(value=n_src / max(el, 1e-9), unit="sources/s", cores=1, kind="port",
                sample=f"float64 oracle, 1 thread: {n_isl} random islands ({n_src} sources) of "
                       f"the batch in {el:.1f}s, per source")


def bench_cutouts(args, torch):
    """Batching workload: independent island cutouts fitted in band-split
    launch sequences; the step is the whole batch's island fit."""
    from uniphot_amd.pipeline import FitConfig, cutout_accuracy, fit_cutouts, psf_bank
    from scenes.cutouts import scene_cutouts
    sync = torch.cuda.synchronize
    t_scene = time.perf_counter()
    cut = scene_cutouts(n_isl=args.islands, n_bands=10, device="cuda")
    sync()
    n = int(cut.x.size)
    print(f"[bench] scene c3: {args.islands} islands, {n} sources, 10 bands, mosaic "
          f"{cut.H}x{cut.W} in {time.perf_counter() - t_scene:.1f}s", file=sys.stderr, flush=True)
    psf_of = psf_bank(cut.psfs, cut.dstack.device)
    cfg = FitConfig()
    for i in range(args.warmup):
        r = fit_cutouts(cut, cfg, sync=sync, psf_of=psf_of)
        print(f"[bench] warmup {i}: chunks {r.chunks}", file=sys.stderr, flush=True)
    timer = KernelTimer()
    sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        res = fit_cutouts(cut, cfg, sync=sync, psf_of=psf_of, timer=timer)
        print(f"[bench] step {i}: chunks {res.chunks}", file=sys.stderr, flush=True)
    sync()
    el = time.perf_counter() - t0
    ks = timer.summary()
    rng = np.random.default_rng(1)
    sample = np.sort(rng.choice(args.islands, min(args.sample, args.islands), replace=False))
    acc = cutout_accuracy(cut, cfg, sample, sync=sync, psf_of=psf_of)
    print(f"[bench] accuracy {acc}", file=sys.stderr, flush=True)
    # the same sample through one persistent workgroup per island, one launch
    # per class (the plain per-class form)
    import dataclasses
    sub_src = int(cut.island_set.n_src[sample].sum())
    t1 = time.perf_counter()
    fit_cutouts(cut, cfg, sync=sync, psf_of=psf_of, island_ids=sample)
    t_split = time.perf_counter() - t1
    t1 = time.perf_counter()
    fit_cutouts(cut, dataclasses.replace(cfg, island_split=False), sync=sync, psf_of=psf_of,
                island_ids=sample)
    t_plain = time.perf_counter() - t1
    line = {
        "metric": "sources fitted/sec at <=1e-5 flux rel-err vs FP64; p50/p99 tile latency",
        "value": round(n * args.steps / el, 2), "unit": "sources/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c3", "n_islands": args.islands, "n_sources": n, "n_bands": 10,
                   "members": "1-64, P(k)~k^-2", "windows": "15-95 px",
                   "launch_sequences": [c for c, _ in res.chunks]},
        "roofline": roofline_of(ks, "c3"),
        "accuracy_vs_fp64": acc,
        "sample_split_vs_persistent": {"islands": int(sample.size), "sources": sub_src,
                                       "band_split_s": round(t_split, 3),
                                       "persistent_per_island_s": round(t_plain, 3)},
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                        for kk, vv in v.items()} for k, v in ks.items()},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_cutouts(cut, res, budget_s=args.cpu_budget)
    print(json.dumps(line), flush=True)


def bench_serving(args, torch):
    """Serving workload: a Poisson stream of field tiles (side 256-4096,
    1-40 bands, 1e-3 sources/px^2) admitted continuously by HBM budget,
    each fitted on its own stream; the tiles are generated on the device
    before the stream starts.  value = sources served per second of
    makespan; p50/p99 request latency beside it."""
    from uniphot_amd.serving import TileServer, latency_summary
    from scenes.tiles import PsfBank, make_tile, tile_stream
    t_scene = time.perf_counter()
    reqs, bank = tile_stream(n_req=args.requests, rate=args.rate)
    torch.cuda.synchronize()
    print(f"[bench] c5: {len(reqs)} requests "
          f"{[(r.side, r.n_bands, r.n_sources) for r in reqs]} generated in "
          f"{time.perf_counter() - t_scene:.1f}s", file=sys.stderr, flush=True)
    srv = TileServer(bank, hbm_budget=200e9, max_inflight=args.inflight)
    warm = make_tile(np.random.default_rng(0), -1, 0.0, 512, 4, bank)
    srv.serve([warm], realtime=False)
    served = srv.serve(reqs, realtime=True)
    summ = latency_summary(served)
    print(f"[bench] served {summ}", file=sys.stderr, flush=True)
    line = {
        "metric": "sources fitted/sec at <=1e-5 flux rel-err vs FP64; p50/p99 tile latency",
        "value": round(summ["sources_per_s"], 2), "unit": "sources/s", "n_gpus": 1,
        "steps": 1, "warmup": 1, "ms_per_step": round(summ["makespan_s"] * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c5", "requests": len(reqs), "rate_per_s": args.rate,
                   "sides": [256, 512, 1024, 2048, 4096], "bands": "1-40",
                   "density_per_px2": 1e-3, "hbm_budget_gb": 200, "max_inflight": args.inflight},
        "latency": {k: round(v, 4) if isinstance(v, float) else v for k, v in summ.items()},
        "requests": [{"rid": s.rid, "side": s.side, "bands": s.n_bands, "sources": s.n_sources,
                      "arrival_s": round(s.arrival, 4), "admitted_s": round(s.admitted, 4),
                      "latency_s": round(s.latency, 4)} for s in served],
    }
    print(json.dumps(line), flush=True)


def bench_precision(args, torch):
    """Precision workload: the 20-band 2048^2 field scaled to 8192^2 with 40
    bands and 80 k sources (same source density, mix and PSF), fitted with
    the island Gram on float32 matrix cores (value) and on bfloat16 ones
    (three refinement steps, islands whose refinement does not settle
    refitted on float32 operands); per-source flux and position error of
    both against the float64 island fit on a sample of islands."""
    import dataclasses
    from uniphot_amd.pipeline import DeviceField, FitConfig, fit_field, island_accuracy
    from scenes.field import SEEDS
    from scenes.tiles import PsfBank, make_tile
    sync = torch.cuda.synchronize
    t_scene = time.perf_counter()
    bank = PsfBank(SEEDS["c4"], n_bands=40)
    side = args.side
    req = make_tile(np.random.default_rng(SEEDS["c4"]), 0, 0.0, side, 40, bank,
                    density=80000.0 / 8192.0 ** 2)
    fld = DeviceField.from_device(req.data, req.invvar, req.good, req.psfs)
    sync()
    n = req.n_sources
    print(f"[bench] scene c4: {side}^2, 40 bands, {n} sources in "
          f"{time.perf_counter() - t_scene:.1f}s", file=sys.stderr, flush=True)
    g = max(1, side // 256)
    base = FitConfig(n_rows=g, n_cols=g)
    cat = (req.x, req.y, req.stype, req.logre, req.ee1, req.ee2, req.sersic_n)
    runs = {}
    for i, gram in enumerate(("fp32", "bf16")):
        cfg = dataclasses.replace(base, gram=gram)
        if i == 0 and args.warmup:
            r = fit_field(fld, *cat, cfg, sync=sync)
            print(f"[bench] warmup: {r.timings}", file=sys.stderr, flush=True)
        timer = KernelTimer()
        sync()
        t0 = time.perf_counter()
        res = fit_field(fld, *cat, cfg, sync=sync, timer=timer, keep_artifacts=True)
        sync()
        el = time.perf_counter() - t0
        print(f"[bench] {gram}: {el:.2f}s {res.timings}", file=sys.stderr, flush=True)
        runs[gram] = (res, el, timer)
    res, el, timer = runs["fp32"]
    art = res.record["_artifacts"]
    iset = art["island_set"]
    rng = np.random.default_rng(2)
    fit_ids = np.setdiff1d(np.arange(iset.n_islands), art["giants"])
    sample = np.sort(rng.choice(fit_ids, min(args.sample, fit_ids.size), replace=False))
    acc = island_accuracy(fld.psf, fld.nb, fld.n_terms, fld.native, art["dstack"],
                          art["wstack"], art["pad"], iset, sample, cat, art["free_n"], base,
                          grams=("bf16", "fp32"), sync=sync, scene=art["scene"],
                          flux_scale=(res.s1_flux, np.sqrt(np.maximum(res.s1_var, 0.0))))
    print(f"[bench] accuracy {acc}", file=sys.stderr, flush=True)
    ks = timer.summary()
    line = {
        "metric": "sources fitted/sec at <=1e-5 flux rel-err vs FP64; p50/p99 tile latency",
        "value": round(n / el, 2), "unit": "sources/s", "n_gpus": 1, "steps": 1,
        "warmup": int(args.warmup), "ms_per_step": round(el * 1e3, 2), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "c4", "image": [side, side], "n_bands": 40, "n_sources": n,
                   "grid": [g, g], "islands": res.record["n_islands"],
                   "stage_s": {k: round(v, 4) for k, v in res.timings.items()}},
        "bf16_gram": {"value": round(n / runs["bf16"][1], 2),
                      "ms_per_step": round(runs["bf16"][1] * 1e3, 2),
                      "stage_s": {k: round(v, 4) for k, v in runs["bf16"][0].timings.items()}},
        "roofline": roofline_of(ks, "c4"),
        "accuracy_vs_fp64": acc,
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                        for kk, vv in v.items()} for k, v in ks.items()},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", help="c2 (default), c2-stars, c1, c3, c4 or c5")
    ap.add_argument("--islands", type=int, default=200_000, help="c3: island cutouts")
    ap.add_argument("--requests", type=int, default=16, help="c5: tile requests")
    ap.add_argument("--rate", type=float, default=50.0, help="c5: mean arrivals per second")
    ap.add_argument("--inflight", type=int, default=4, help="c5: tiles in flight at most")
    ap.add_argument("--side", type=int, default=8192, help="c4: field side (px)")
    ap.add_argument("--sample", type=int, default=300,
                    help="islands of the FP64 accuracy comparison (c2, c3, c4; 0 = none)")
    ap.add_argument("--cpu-budget", type=float, default=12.0,
                    help="c3: seconds of float64-oracle CPU work for its baseline")
    ap.add_argument("--cpu-islands", type=int, default=120,
                    help="islands of the CPU baseline's stratified sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--island-split", type=int, default=1,
                    help="1 = band-split island LM (each island's bands spread over all CUs), "
                         "0 = one persistent workgroup per island")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB before the steps, as the "
                         "known-size reference for FETCH_SIZE / WRITE_SIZE counter runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the fit would shard by independent islands and grid tiles across ranks)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if args.config == "c3":
        bench_cutouts(args, torch)
        return
    if args.config == "c5":
        bench_serving(args, torch)
        return
    if args.config == "c4":
        bench_precision(args, torch)
        return
    from uniphot_amd.pipeline import DeviceField, FitConfig, fit_field
    t_scene = time.perf_counter()
    sc, gcfg = make_scene(args.config)
    print(f"[bench] scene {args.config}: {sc.x.size} sources, {sc.n_bands} bands, "
          f"{sc.data.shape[1]}x{sc.data.shape[2]} in {time.perf_counter() - t_scene:.1f}s",
          file=sys.stderr, flush=True)
    fld = DeviceField(sc.data, sc.invvar, sc.good, sc.psfs)
    cfg = FitConfig(**gcfg, island_split=bool(args.island_split))
    sync = torch.cuda.synchronize
    if args.copy_calib_mib > 0:
        src = torch.ones(args.copy_calib_mib << 20, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        sync()
        dst.copy_(src)
        sync()
        del src, dst
    args_fit = (sc.x, sc.y, sc.stype, sc.logre, sc.ee1, sc.ee2, sc.sersic_n, cfg)
    for i in range(args.warmup):
        r = fit_field(fld, *args_fit, sync=sync)
        print(f"[bench] warmup {i}: {r.timings}", file=sys.stderr, flush=True)
    timer = KernelTimer()
    sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        res = fit_field(fld, *args_fit, sync=sync, timer=timer, keep_artifacts=(i == args.steps - 1))
        print(f"[bench] step {i}: {res.timings}", file=sys.stderr, flush=True)
    sync()
    el = time.perf_counter() - t0
    n = sc.x.size
    value = n * args.steps / el
    ks = timer.summary()
    if os.environ.get("UPH_BENCH_VERBOSE"):
        for name, ms, info in timer.per_launch()[:400]:
            print(f"[bench] {name} {ms:.3f} ms {info}", file=sys.stderr)
    roof = roofline_of(ks, args.config)
    from uniphot_amd import autotune
    line = {
        "metric": "sources fitted/sec at <=1e-5 flux rel-err vs FP64; p50/p99 tile latency",
        "value": round(value, 2), "unit": "sources/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.config, "n_sources": int(n), "n_bands": sc.n_bands,
                   "image": [int(sc.data.shape[1]), int(sc.data.shape[2])],
                   "grid": [cfg.n_rows, cfg.n_cols], "islands": res.record["n_islands"],
                   "stage_s": {k: round(v, 4) for k, v in res.timings.items()}},
        "roofline": roof,
        "autotune": autotune.log(),
        "kernels": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv)
                        for kk, vv in v.items()} for k, v in ks.items()},
    }
    if args.sample > 0:
        t_acc = time.perf_counter()
        cat = (sc.x, sc.y, sc.stype, sc.logre, sc.ee1, sc.ee2, sc.sersic_n)
        acc = field_accuracy(fld, res, cat, cfg, args.sample, sync)
        frac = acc.get("fit", {}).get("frac_le_1e5", float("nan"))
        acc["value_within_1e5"] = round(value * frac, 2)
        line["accuracy_vs_fp64"] = acc
        print(f"[bench] accuracy vs fp64 ({time.perf_counter() - t_acc:.1f}s) {acc}",
              file=sys.stderr, flush=True)
        # the same fit with every island on the float32 kernels (no float64
        # routing or refit): its throughput and its accuracy on the sample
        from uniphot_amd import autotune
        b32 = [t["choice"] for t in autotune.log() if "|fp32|" in t["key"]]
        c32 = dataclasses.replace(cfg, fp64_contrast=None, fp64_refit_kappa=None,
                                  island_blocks=dict(fp32=b32[-1]) if b32 else None,
                                  autotune=not b32)
        sync()
        t32 = time.perf_counter()
        r32 = fit_field(fld, *args_fit[:-1], c32, sync=sync, keep_artifacts=True)
        sync()
        el32 = time.perf_counter() - t32
        a32 = field_accuracy(fld, r32, cat, c32, args.sample, sync)
        v32 = n / el32
        line["fp32_only"] = {"value": round(v32, 2), "ms_per_step": round(el32 * 1e3, 2),
                             "stage_s": {k: round(v, 4) for k, v in r32.timings.items()},
                             "accuracy_vs_fp64": a32,
                             "value_within_1e5": round(
                                 v32 * a32.get("fit", {}).get("frac_le_1e5", float("nan")), 2)}
        print(f"[bench] fp32 only {line['fp32_only']}", file=sys.stderr, flush=True)
    line["config"]["fp64_refit_islands"] = res.record["n_fp64_refit_islands"]
    if "fcorr" in res.record:
        fc = res.record["fcorr"]
        line["config"]["fcorr"] = {"F": [round(f, 5) for f in fc["F"]], "source": fc["source"]}
    if not args.no_cpu:
        cmp = []
        t_cpu = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline(fld, res, sc, cfg, n_islands=args.cpu_islands,
                                            oracle_cmp=cmp)
        print(f"[bench] cpu baseline ({time.perf_counter() - t_cpu:.1f}s) "
              f"{line['cpu_baseline']}", file=sys.stderr, flush=True)
        line["accuracy_vs_oracle"] = oracle_accuracy(res, cmp)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
