"""Per-island VarPro-LM on the GPU (host facade over ``uph_island_fit``).

Mirrors the reference pair ``make_class_kernels`` / ``run_class_chunk``
(``tract7dt/fit/grid_gpu/island_solver.py:1874-1922``, ``:2793-3131``) for
the legacy step protocol: same host dict layout (``class_host_arrays``),
same padded field stacks, same per-row output dict.  Differences by
design: there is no vmap width (one persistent kernel pulls islands from a
device work queue, so ``width`` is accepted and ignored), and the LM loop
runs on the device with no per-step host synchronisation.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from . import _lib
from .island_plan import (DEFAULT_CHI2_FTOL_REL, DEFAULT_MAX_STEPS, DEFAULT_POS_STEP_MAX,
                          DEFAULT_REG, DEFAULT_SHAPE_STEP_MAX, DEFAULT_SHAPE_STEP_TOL,
                          DEFAULT_STEP_TOL_PIX, DEFAULT_UN_STEP_MAX)

# Cap on the persistent workgroups of one launch (4 per CU on 256 CUs) and
# the default cap on the workspace they may hold.  The per-workgroup
# workspace is sized for the largest island of the batch (about 34 MB for
# 32 members in a 192^2 window), so the default keeps 4 evaluation
# workgroups per CU there; a caller sharing the device passes less.
MAX_BLOCKS = 1024
MAX_SCRATCH_BYTES = 40 << 30
# Islands whose largest Gram (C^2 x window pixels, C = 1 + N + live
# parameters) reaches this get it computed by the many-workgroup Gram
# launch instead of inside their one evaluation workgroup.
GRAM_SPLIT_MIN = 3e7
GRAM_SPLIT_BYTES = 2e9  # cap on their Gram tile-pair buffer per launch sequence


@dataclass
class ClassKernels:
    """Static description of one island class (the reference's cache key)."""
    B: int
    T: int
    N: int
    n_bands: int
    n_terms: int
    kind: str = "star"
    reg: float = DEFAULT_REG
    pos_step_max: float = DEFAULT_POS_STEP_MAX
    chi2_ftol_rel: float = DEFAULT_CHI2_FTOL_REL
    step_tol_pix: float = DEFAULT_STEP_TOL_PIX
    shape_step_max: float = DEFAULT_SHAPE_STEP_MAX
    un_step_max: float = DEFAULT_UN_STEP_MAX
    shape_step_tol: float = DEFAULT_SHAPE_STEP_TOL
    precision: str = "fp32"

    @property
    def n_live_max(self) -> int:
        return 2 * self.N if self.kind == "star" else 6 * self.N

    def options(self, max_steps: int, resume: bool = False) -> _lib.LMOptions:
        return _lib.LMOptions(self.reg, self.pos_step_max, self.shape_step_max,
                              self.un_step_max, self.chi2_ftol_rel, self.step_tol_pix,
                              self.shape_step_tol, int(max_steps), int(bool(resume)))


def make_class_kernels(plan, n_bands: int, n_terms: int, *, reg: float = DEFAULT_REG,
                       pos_step_max_pix: float = DEFAULT_POS_STEP_MAX,
                       chi2_ftol_rel: float = DEFAULT_CHI2_FTOL_REL,
                       step_tol_pix: float = DEFAULT_STEP_TOL_PIX,
                       precision: str = "fp32", **_ignored) -> ClassKernels:
    return ClassKernels(B=int(plan.bbox_class), T=int(plan.tier), N=int(plan.nsrc_class),
                        n_bands=int(n_bands), n_terms=int(n_terms), kind=str(plan.kind),
                        reg=float(reg), pos_step_max=float(pos_step_max_pix),
                        chi2_ftol_rel=float(chi2_ftol_rel), step_tol_pix=float(step_tol_pix),
                        precision=precision)


@dataclass
class PsfTree:
    """Device PSF of one tier: coefs (nb, nt, T, T), geo (nb, 4)."""
    coefs: object
    geo: object
    T: int
    n_terms: int


def psf_tree_device(psfs, tier: int, device="cuda") -> PsfTree:
    import torch
    from .psf_polynomial import psf_device_arrays
    c, g = psf_device_arrays(psfs, tier)
    return PsfTree(torch.from_numpy(c).to(device), torch.from_numpy(g).to(device), tier,
                   c.shape[1])


def run_class_chunk(kernels: ClassKernels, psf_tree: PsfTree, host: dict, rows,
                    width: int | None, max_steps: int, dstack_dev, wstack_dev, pad: int,
                    init_state=None, compute_param_errors: bool = False,
                    want_residual: bool = False, compute_source_metrics: bool = False,
                    compute_flux_marg: bool = False, compute_re_profile: bool = False, *,
                    stream=None, sersic_dev=None, timer=None, form: str = "split",
                    stall_probes_on: bool = True) -> dict:
    """Fit rows ``rows`` of a class host dict on the device and return the
    reference's per-row output dict (island_solver.py:2793-3131): p (n_c,
    N, 6), fluxes and ferr (n_c, nb, N), perr (n_c, N, 13: marginal
    variances of x, y, logre, ee1, ee2, u_n, cov(ee1, ee2), then six prior
    fractions -- NaN, no priors in this build), femarg (n_c, nb, N),
    smetrics (n_c, N, 4), swin (n_c, nb, N), n_good, chi2, conv, stall,
    diverged, stall_dchi2, railed, n_same, prej, nst, lam, consec, nrej,
    wall_s, steps_run; None for what was not requested.  ``init_state``
    (p, lam, chi2, consec, nrej, nst per row) resumes the LM from that
    state.  ``width`` is accepted for the call signature: the device work
    queue replaces the vmap width.  ``form="persistent"`` runs the one-class
    persistent kernel (no post-fit pass, no probes)."""
    if host.get("has_priors"):
        raise NotImplementedError("catalog priors are not part of this build")
    if want_residual or compute_re_profile:
        raise NotImplementedError("island residual export and the Re profile scan (opt-in in "
                                  "the reference) are not part of this build")
    if form == "persistent":
        return _run_class_persistent(kernels, psf_tree, host, rows, max_steps, dstack_dev,
                                     wstack_dev, init_state, stream, sersic_dev, timer)
    if psf_tree.T != kernels.T:
        raise ValueError(f"psf tree tier {psf_tree.T} != class tier {kernels.T}")
    t0 = time.perf_counter()
    rows = np.asarray(rows, np.int64)
    n_c, N, nb = int(rows.size), kernels.N, kernels.n_bands
    want_post = compute_param_errors or compute_source_metrics or compute_flux_marg
    post = want_post or stall_probes_on
    trees = {kernels.T: psf_tree}
    cls = [(kernels, host, rows)]
    inits = None if init_state is None else [init_state]
    if kernels.precision == "fp64":
        res = fit_float64(cls, trees, max_steps, dstack_dev, wstack_dev, sersic_dev=sersic_dev,
                          post=post, init_states=inits)[0]
    else:
        res = run_island_batch(cls, trees, max_steps, dstack_dev, wstack_dev,
                               sersic_dev=sersic_dev, timer=timer, split=True, prerender=True,
                               post=post, init_states=inits, stream=stream)[0]
    if stall_probes_on and "post_hg" in res:
        stall_probes(cls, [res], trees, dstack_dev, wstack_dev, sersic_dev=sersic_dev)
    out = {k: res[k] for k in ("p", "fluxes", "ferr", "n_good", "chi2", "conv", "stall", "nst",
                               "lam", "consec", "nrej")}
    out["diverged"] = res.get("diverged", np.zeros(n_c, bool))
    out["stall_dchi2"] = res.get("stall_dchi2", np.full(n_c, np.nan, np.float32))
    out["railed"] = np.zeros(n_c, bool)          # trust-region mode only
    out["n_same"] = np.zeros(n_c, np.int32)      # "changed" stall counting only
    out["prej"] = np.zeros((n_c, 6 * N), np.float32)
    out["perr"] = out["femarg"] = out["smetrics"] = out["swin"] = None
    out["reprof"] = out["resid"] = out["resid_w"] = None
    if want_post and "post_var" in res:
        from .pipeline import metrics_from_sums
        if compute_param_errors:
            perr = np.full((n_c, N, 13), np.nan, np.float32)
            perr[:, :, :7] = res["post_var"]
            out["perr"] = perr
        if compute_flux_marg:
            out["femarg"] = res["post_dvar"]
        if compute_source_metrics:
            m4, fw = metrics_from_sums(res["post_metrics"])
            out["smetrics"] = m4.astype(np.float32)
            out["swin"] = np.transpose(fw, (0, 2, 1)).astype(np.float32)
    out["steps_run"] = res["steps_run"]
    out["wall_s"] = time.perf_counter() - t0
    return out


def _is_device_mem_error(exc: BaseException) -> bool:
    """A device-memory failure a smaller chunk may avoid: the library's
    UPH_ERR_OOM, or torch's own out-of-memory error."""
    if isinstance(exc, _lib.UphOOM):
        return True
    e = f"{type(exc).__name__}: {exc}".upper()
    return any(k in e for k in ("OUT OF MEMORY", "OUT_OF_MEMORY", "RESOURCE_EXHAUSTED"))


def _merge_chunk_outputs(parts: list) -> dict:
    """Row-concatenate the result dicts of a split retry."""
    out = dict(parts[0])
    for k, v0 in parts[0].items():
        if k in ("wall_s", "steps_run"):
            continue
        if v0 is None:
            out[k] = None
        elif isinstance(v0, np.ndarray):
            out[k] = np.concatenate([p[k] for p in parts], axis=0)
    out["wall_s"] = float(sum(p["wall_s"] for p in parts))
    out["steps_run"] = int(max(p["steps_run"] for p in parts))
    return out


def run_class_chunk_resilient(kernels, psf_tree, host, rows, width, max_steps, dstack_dev,
                              wstack_dev, pad, *, init_state=None, log_fn=None, **kw) -> dict:
    """:func:`run_class_chunk`, retried as two half chunks (recursively) on
    a device-memory failure (island_solver.py:3165-3230); anything else
    propagates, and a single-row chunk that fails re-raises."""
    rows = np.asarray(rows)
    try:
        out = run_class_chunk(kernels, psf_tree, host, rows, width, max_steps, dstack_dev,
                              wstack_dev, pad, init_state=init_state, **kw)
        out["halved"] = False
        return out
    except Exception as exc:                                    # noqa: BLE001
        if not _is_device_mem_error(exc) or rows.size <= 1:
            raise
        half = (rows.size + 1) // 2
        msg = (f"island chunk of {rows.size} hit {type(exc).__name__} ({str(exc)[:80]}) -- "
               f"retrying as two chunks")
        (log_fn or (lambda m: print("island: " + m, flush=True)))(msg)
        parts = []
        for sl in (slice(0, half), slice(half, rows.size)):
            sub = None if init_state is None else {k: np.asarray(v)[sl]
                                                   for k, v in init_state.items()}
            parts.append(run_class_chunk_resilient(kernels, psf_tree, host, rows[sl],
                                                   width, max_steps, dstack_dev, wstack_dev,
                                                   pad, init_state=sub, log_fn=log_fn, **kw))
        merged = _merge_chunk_outputs(parts)
        merged["halved"] = True
        return merged


def _run_class_persistent(kernels: ClassKernels, psf_tree: PsfTree, host: dict, rows,
                          max_steps: int, dstack_dev, wstack_dev, init_state, stream,
                          sersic_dev, timer) -> dict:
    """One class on the persistent single-class kernel (uph_island_fit)."""
    import torch
    t0 = time.perf_counter()
    rows = np.asarray(rows, np.int64)
    n_c = int(rows.size)
    N, B, T = kernels.N, kernels.B, kernels.T
    nb = kernels.n_bands
    nP = 6 * N
    if psf_tree.T != T:
        raise ValueError(f"psf tree tier {psf_tree.T} != class tier {T}")
    if tuple(dstack_dev.shape) != tuple(wstack_dev.shape) or dstack_dev.shape[0] != nb:
        raise ValueError("dstack/wstack must both be (n_bands, Hp, Wp)")
    Hp, Wp = int(dstack_dev.shape[1]), int(dstack_dev.shape[2])
    ox = host["ox"][rows]
    oy = host["oy"][rows]
    if n_c and (ox.min() < 0 or oy.min() < 0 or ox.max() + B > Wp or oy.max() + B > Hp):
        raise ValueError("island window outside the padded stacks")
    if kernels.kind != "star" and sersic_dev is None:
        raise ValueError("mixed classes need the device Sersic table")
    dev = dstack_dev.device

    def up(a, dt):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev, non_blocking=False)

    mm = host["mmask"][rows]
    # LPT order: most real members first
    order = np.argsort(-mm.sum(axis=1), kind="stable").astype(np.int32)
    ins = dict(
        ox=up(ox, np.int32), oy=up(oy, np.int32), x0b=up(host["x0b"][rows], np.float32),
        y0b=up(host["y0b"][rows], np.float32), stype=up(host["st"][rows], np.int32),
        n_frozen=up(host["sn"][rows], np.float32), free_n=up(host["fn_flag"][rows], np.uint8),
        mmask=up(mm, np.uint8), keep=up(host["keep"][rows], np.uint8),
        p0=up((host["p0"][rows] if init_state is None else init_state["p"]).reshape(n_c, nP),
              np.float32),
        lower=up(host["lower"][rows].reshape(n_c, nP), np.float32),
        upper=up(host["upper"][rows].reshape(n_c, nP), np.float32), order=up(order, np.int32))
    f32 = dict(dtype=torch.float32, device=dev)
    outs = dict(
        p=torch.empty(n_c, nP, **f32), fluxes=torch.empty(n_c, nb, N, **f32),
        ferr=torch.empty(n_c, nb, N, **f32), chi2=torch.empty(n_c, dtype=torch.float64, device=dev),
        lam=torch.empty(n_c, dtype=torch.float64, device=dev), conv=torch.empty(n_c, dtype=torch.uint8, device=dev),
        stall=torch.empty(n_c, dtype=torch.uint8, device=dev),
        nst=torch.empty(n_c, dtype=torch.int32, device=dev),
        consec=torch.empty(n_c, dtype=torch.int32, device=dev),
        nrej=torch.empty(n_c, dtype=torch.int32, device=dev),
        n_good=torch.empty(n_c, dtype=torch.int32, device=dev))
    if init_state is not None:
        for k, dt in (("chi2", np.float64), ("lam", np.float64), ("consec", np.int32),
                      ("nrej", np.int32), ("nst", np.int32)):
            outs[k].copy_(torch.from_numpy(np.asarray(init_state[k], dt).reshape(-1)))
    prec = _lib.UPH_PRECISION_FP64 if kernels.precision == "fp64" else _lib.UPH_PRECISION_FP32
    L = _lib.lib()
    per = L.uph_island_scratch_per_block(N, B, T, kernels.n_live_max, prec)
    nblk = max(1, min(n_c, MAX_BLOCKS, MAX_SCRATCH_BYTES // max(per, 1)))
    scratch = torch.empty(per * nblk, dtype=torch.uint8, device=dev)
    counter = torch.zeros(1, dtype=torch.int32, device=dev)
    cls = _lib.IslandClass(n_c, N, B, T, psf_tree.n_terms, nb, kernels.n_live_max, Hp, Wp,
                           0 if kernels.kind == "star" else 1)
    P = _lib.ptr
    inp = _lib.IslandInputs(P(dstack_dev), P(wstack_dev), P(psf_tree.coefs), P(psf_tree.geo),
                            P(sersic_dev), *(P(ins[k]) for k in (
                                "ox", "oy", "x0b", "y0b", "stype", "n_frozen", "free_n",
                                "mmask", "keep", "p0", "lower", "upper", "order")))
    out = _lib.IslandOutputs(*(P(outs[k]) for k in (
        "p", "fluxes", "ferr", "chi2", "lam", "conv", "stall", "nst", "consec", "nrej",
        "n_good")))
    if n_c:
        ev = timer.begin("island_lm_kernel") if timer is not None else None
        _lib.check(L.uph_island_fit(C.byref(cls), C.byref(inp), C.byref(out),
                                    C.byref(kernels.options(max_steps, init_state is not None)),
                                    prec, P(scratch),
                                    scratch.numel(), P(counter), _lib.stream_ptr(stream)))
        if ev is not None:
            timer.end(ev, dict(kind="island", n=n_c, B=B, N=N, nb=nb, outs=outs, host=host,
                               rows=rows))
    res = {k: v.cpu().numpy() for k, v in outs.items()}
    res["p"] = res["p"].reshape(n_c, N, 6)
    for k in ("conv", "stall"):
        res[k] = res[k].astype(bool)
    res["n_good"] = res["n_good"].astype(np.int64)
    res["steps_run"] = int(res["nst"].max()) if n_c else 0
    res["wall_s"] = time.perf_counter() - t0
    return res


def _fft_pad(T: int) -> int:
    """FFT size of a stamp of side T: the power of two >= T + T // 2
    (sersic_mog.py:196-207)."""
    n = 1
    while n < T + T // 2:
        n <<= 1
    return n


def _prerender_arrays(order, shape, moff, cat, tiers, nb, m_acc, precision, prerender, dev, up):
    """Index arrays and buffers of the band-split member pre-render: every
    real member, extended members first, each kind in the islands'
    work-queue order; extended members also get a block of Fourier
    spectra.  Planes per band: 7 for an extended member, 3 for a star, each
    min(T, B)^2 floats."""
    import torch
    L = _lib.lib()
    P = _lib.ptr
    empty = dict(args=(0, 0, *([None] * 6), 0, None, 0, None, None),
                 isl=np.zeros(0, np.int64), T=np.zeros(0, np.int64), star=np.zeros(0, bool),
                 spec_floats=np.zeros(0, np.float64), n_ext=0)
    if not prerender:
        return empty
    f64 = precision == "fp64"
    st_all, mm_all = cat["st"], cat["mm"].astype(bool)
    ext, stars = [], []
    for i in order:
        m0, N = moff[i], shape[i][0]
        for k in np.nonzero(mm_all[m0:m0 + N])[0]:
            (stars if st_all[m0 + k] == 0 else ext).append((int(i), int(k)))
    members = ext + stars
    if not members:
        return empty
    pre_isl = [i for i, _ in members]
    pre_slot = [k for _, k in members]
    slot_pre = np.full(m_acc, -1, np.int32)
    for g, (i, k) in enumerate(members):
        slot_pre[moff[i] + k] = g
    gi = np.asarray(pre_isl, np.int64)
    T = np.asarray([tiers[shape[i][2]] for i in pre_isl], np.int64)
    if not f64 and any(L.uph_prerender_scratch_bytes(int(t)) == 0 for t in set(T.tolist())):
        return empty
    star = np.arange(len(members)) >= len(ext)
    # one plane holds the window part of the stamp: at most min(T, B) a side
    side = np.minimum(T, np.asarray([shape[i][1] for i in pre_isl], np.int64))
    psz = nb * np.where(star, 3, 7) * side * side
    spec_of = L.uph_prerender64_spectra_doubles if f64 else L.uph_prerender_spectra_floats
    ssz = np.where(star, 0, [spec_of(int(t)) for t in T]).astype(np.int64)
    poff = np.concatenate([[0], np.cumsum(psz)[:-1]]).astype(np.int64)
    soff = np.concatenate([[0], np.cumsum(ssz)[:-1]]).astype(np.int64)
    soff = np.where(star, -1, soff).astype(np.int64)
    # float64 islands: planes and spectra in doubles, no taps/derivative scratch
    fdt = torch.float64 if f64 else torch.float32
    pbuf = torch.empty(int(psz.sum()), dtype=fdt, device=dev)
    fspec = torch.empty(max(int(ssz.sum()), 1), dtype=fdt, device=dev)
    gdt = torch.empty(1 if f64 else len(members) * 9, dtype=torch.float64, device=dev)
    gtap = torch.empty(1 if f64 else len(members) * 28, dtype=torch.float32, device=dev)
    keep = [up(np.asarray(pre_isl, np.int32), np.int32), up(np.asarray(pre_slot, np.int32), np.int32),
            up(poff, np.int64), up(soff, np.int64), up(slot_pre, np.int32)]
    args = (len(members), len(ext), *(P(t) for t in keep), P(pbuf), int(pbuf.numel()), P(fspec),
            int(fspec.numel()), P(gdt), P(gtap))
    # window part and shift-read box of each stamp at the initial position
    # (positions move by well under a pixel), for the bytes accounting
    slot = np.asarray(moff, np.int64)[gi] + np.asarray(pre_slot, np.int64)
    Bw = np.asarray([shape[i][1] for i in pre_isl], np.int64)
    p0 = cat["p0"].reshape(-1, 6)[slot]
    ix = np.rint(p0[:, 0].astype(np.float32) - cat["x0b"][gi]).astype(np.int64) - T // 2
    iy = np.rint(p0[:, 1].astype(np.float32) - cat["y0b"][gi]).astype(np.int64) - T // 2
    sw = np.clip(np.minimum(Bw, ix + T) - np.maximum(0, ix), 0, None)
    sh = np.clip(np.minimum(Bw, iy + T) - np.maximum(0, iy), 0, None)
    rw = np.minimum(T, np.minimum(Bw, ix + T) - ix + 3) - np.maximum(0, np.maximum(0, ix) - ix - 3)
    rh = np.minimum(T, np.minimum(Bw, iy + T) - iy + 3) - np.maximum(0, np.maximum(0, iy) - iy - 3)
    live_xy = cat["keep"].reshape(-1, 6)[slot][:, :2].sum(axis=1)
    live = cat["keep"].reshape(-1, 6)[slot].sum(axis=1)
    return dict(args=args, keepalive=(keep, pbuf, fspec, gdt, gtap), isl=gi, T=T, star=star,
                spec_floats=ssz.astype(np.float64), n_ext=len(ext),
                win=(sw * sh).astype(np.float64), reg=np.clip(rw * rh, 0, None).astype(np.float64),
                grad_planes=np.where(star, 1 + live_xy, 1 + live).astype(np.float64))


def prepare_island_batch(classes, psf_trees: dict, max_steps: int, dstack_dev, wstack_dev, *,
                         precision: str = "fp32", sersic_dev=None, timer=None,
                         split: bool = False, prerender: bool = True,
                         post: bool = False, gram: str = "fp32",
                         workspace_bytes: float = MAX_SCRATCH_BYTES,
                         gram_split_min: float = GRAM_SPLIT_MIN,
                         init_states=None, max_blocks: int | None = None
                         ) -> "PreparedIslandBatch":
    """Stage the islands of several classes for ONE launch sequence.

    ``classes`` is a list of (ClassKernels, host dict, rows); ``psf_trees``
    maps stamp tier -> PsfTree.  Islands are ordered longest-first
    (members x window area) for the work queue.  The returned batch's
    ``collect()`` gives one result dict per class, in the same layout as
    :func:`run_class_chunk`.

    ``split=False`` is the persistent form (``uph_island_fit_batch``, one
    workgroup per island).  ``split=True`` runs the band-split form
    (``uph_island_fit_split``): every evaluation pass is spread over
    (island, band) items, so one large island no longer sets the wall time;
    results are bit-identical.  With ``prerender`` every member is rendered
    and shifted by separate (member, band) launches before each pass and
    the evaluation only places the planes: float32 (extended members
    through LDS wave-level FFTs; results agree with the in-place render to
    float32 rounding) or float64 (the in-place render's operations, so the
    results are bit-identical to it).  A ``timer`` gets per-launch events of the
    split sequence.  ``post`` (split + pre-render only) adds the post-fit
    pass: per-member parameter variances (``post_var`` (n, N, 7)), the
    shape-marginalisation flux variance increment (``post_dvar`` (n, nb, N))
    and the per-band metric sums (``post_metrics`` (n, N, nb, 6)), plus the
    scaled condition number of each island's Gauss-Newton matrix
    (``post_kappa`` (n,)).
    ``max_blocks`` caps the persistent workgroups of the launches (default
    MAX_BLOCKS; :mod:`uniphot_amd.autotune` measures it per workload).
    ``gram="bf16"`` (split + pre-render) builds the island Gram matrices on
    bfloat16 matrix cores with float32 accumulation.  ``init_states`` (one
    dict per class with p (n_c, N, 6), lam, chi2, consec, nrej, nst) resumes
    every island from that LM state (island_solver.py:2902-2930); all
    classes or none carry one.
    """
    import torch
    t0 = time.perf_counter()
    resume = init_states is not None and any(st is not None for st in init_states)
    if resume and any(st is None for st in init_states):
        raise ValueError("init_states: every class or none resumes")
    if not classes:
        return PreparedIslandBatch(lambda stream=None: None, lambda: None, lambda: [])
    k0 = classes[0][0]
    nb, nt = k0.n_bands, k0.n_terms
    if any(k.kind != "star" for k, _, _ in classes) and sersic_dev is None:
        raise ValueError("mixed classes need the device Sersic table")
    tiers = sorted({k.T for k, _, _ in classes})
    if len(tiers) > 3:
        raise ValueError(f"at most 3 stamp tiers per batch, got {tiers}")
    tier_idx = {T: i for i, T in enumerate(tiers)}
    Hp, Wp = int(dstack_dev.shape[1]), int(dstack_dev.shape[2])
    dev = dstack_dev.device
    shape, moff, foff, cost = [], [], [], []
    cat = {k: [] for k in ("ox", "oy", "x0b", "y0b", "st", "sn", "fn", "mm", "keep", "p0",
                           "lo", "hi")}
    m_acc = f_acc = 0
    spans = []
    per = 0
    prec = _lib.UPH_PRECISION_FP64 if precision == "fp64" else _lib.UPH_PRECISION_FP32
    L = _lib.lib()
    for kern, host, rows in classes:
        rows = np.asarray(rows, np.int64)
        n_c, N, B = rows.size, kern.N, kern.B
        if kern.n_bands != nb or kern.n_terms != nt:
            raise ValueError("classes of one batch must share n_bands and n_terms")
        ox, oy = host["ox"][rows], host["oy"][rows]
        if n_c and (ox.min() < 0 or oy.min() < 0 or ox.max() + B > Wp or oy.max() + B > Hp):
            raise ValueError("island window outside the padded stacks")
        mm = host["mmask"][rows]
        spans.append((len(shape), n_c, m_acc, f_acc, N))
        for i in range(n_c):
            shape.append((N, B, tier_idx[kern.T], kern.n_live_max))
            moff.append(m_acc + i * N)
            foff.append(f_acc + i * nb * N)
        cost.extend((mm.sum(axis=1) * B * B).tolist())
        m_acc += n_c * N
        f_acc += n_c * nb * N
        cat["ox"].append(ox)
        cat["oy"].append(oy)
        cat["x0b"].append(host["x0b"][rows])
        cat["y0b"].append(host["y0b"][rows])
        cat["st"].append(host["st"][rows].ravel())
        cat["sn"].append(host["sn"][rows].ravel())
        cat["fn"].append(host["fn_flag"][rows].ravel())
        cat["mm"].append(mm.ravel())
        cat["keep"].append(host["keep"][rows].ravel())
        if resume:
            st0 = init_states[len(spans) - 1]
            cat["p0"].append(np.asarray(st0["p"], np.float32).reshape(n_c, -1).ravel())
        else:
            cat["p0"].append(host["p0"][rows].ravel())
        cat["lo"].append(host["lower"][rows].ravel())
        cat["hi"].append(host["upper"][rows].ravel())
        per = max(per, L.uph_island_scratch_per_block(N, B, kern.T, kern.n_live_max, prec))
    n_isl = len(shape)
    if split and prerender:
        for T in tiers:
            need = int(L.uph_prerender64_scratch_bytes(int(T)) if precision == "fp64" else
                       L.uph_prerender_scratch_bytes(int(T)))
            if need == 0:
                prerender = False  # a stamp tier the pre-render does not serve
            per = max(per, need)
    per = (per + 255) // 256 * 256

    def up(a, dt):
        return torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)

    order = np.argsort(-np.asarray(cost, np.float64), kind="stable").astype(np.int32)
    c = {k: np.concatenate(v) for k, v in cat.items()}
    ins = dict(ox=up(c["ox"], np.int32), oy=up(c["oy"], np.int32), x0b=up(c["x0b"], np.float32),
               y0b=up(c["y0b"], np.float32), stype=up(c["st"], np.int32),
               n_frozen=up(c["sn"], np.float32), free_n=up(c["fn"], np.uint8),
               mmask=up(c["mm"], np.uint8), keep=up(c["keep"], np.uint8),
               p0=up(c["p0"], np.float32), lower=up(c["lo"], np.float32),
               upper=up(c["hi"], np.float32), order=up(order, np.int32))
    d_shape = up(np.asarray(shape, np.int32), np.int32)
    d_moff = up(np.asarray(moff, np.int64), np.int64)
    d_foff = up(np.asarray(foff, np.int64), np.int64)
    f32 = dict(dtype=torch.float32, device=dev)
    outs = dict(
        p=torch.empty(m_acc * 6, **f32), fluxes=torch.empty(f_acc, **f32),
        ferr=torch.empty(f_acc, **f32), chi2=torch.empty(n_isl, dtype=torch.float64, device=dev),
        lam=torch.empty(n_isl, dtype=torch.float64, device=dev), conv=torch.empty(n_isl, dtype=torch.uint8, device=dev),
        stall=torch.empty(n_isl, dtype=torch.uint8, device=dev),
        nst=torch.empty(n_isl, dtype=torch.int32, device=dev),
        consec=torch.empty(n_isl, dtype=torch.int32, device=dev),
        nrej=torch.empty(n_isl, dtype=torch.int32, device=dev),
        n_good=torch.empty(n_isl, dtype=torch.int32, device=dev))
    n_work = n_isl * nb if split else n_isl
    nblk = max(1, min(n_work, int(max_blocks or MAX_BLOCKS), int(workspace_bytes) // per))
    scratch = torch.empty(per * nblk, dtype=torch.uint8, device=dev)
    counter = torch.zeros(1, dtype=torch.int32, device=dev)
    P = _lib.ptr
    bat = _lib.IslandBatch()
    bat.n_isl, bat.n_bands, bat.n_terms, bat.Hp, bat.Wp = n_isl, nb, nt, Hp, Wp
    bat.max_B = max(s[1] for s in shape)
    bat.mixed = int(any(k.kind != "star" for k, _, _ in classes))
    bat.n_tiers = len(tiers)
    for i, T in enumerate(tiers):
        bat.tier_T[i] = T
        bat.psf_coefs[i] = P(psf_trees[T].coefs)
    bat.shape, bat.mem_off, bat.flux_off = P(d_shape), P(d_moff), P(d_foff)
    geo = psf_trees[tiers[0]].geo
    inp = _lib.IslandInputs(P(dstack_dev), P(wstack_dev), None, P(geo), P(sersic_dev),
                            *(P(ins[k]) for k in ("ox", "oy", "x0b", "y0b", "stype",
                                                  "n_frozen", "free_n", "mmask", "keep", "p0",
                                                  "lower", "upper", "order")))
    out = _lib.IslandOutputs(*(P(outs[k]) for k in (
        "p", "fluxes", "ferr", "chi2", "lam", "conv", "stall", "nst", "consec", "nrej",
        "n_good")))
    opts = k0.options(max_steps, resume=resume)
    if resume:
        # the LM state the kernels resume from, in the outputs they update
        for k, dt in (("chi2", np.float64), ("lam", np.float64), ("consec", np.int32),
                      ("nrej", np.int32), ("nst", np.int32)):
            v = np.concatenate([np.asarray(st0[k], dt).reshape(-1) for st0 in init_states])
            outs[k].copy_(torch.from_numpy(v))
    if split:
        nlc = np.asarray([s_[3] for s_ in shape], np.int64)
        blk = nlc * nlc + nlc + 2
        hg_off = np.concatenate([[0], np.cumsum(blk * nb)[:-1]]).astype(np.int64)
        hg_n = int((blk * nb).sum())
        items = np.repeat(order.astype(np.int64), nb)
        bands = np.tile(np.arange(nb, dtype=np.int32), n_isl)
        hg = torch.empty(hg_n, dtype=torch.float64, device=dev)
        p_try = torch.empty(m_acc * 6, **f32)
        n_cnt = int(L.uph_island_split_counters(int(max_steps)))
        counters = torch.empty(n_cnt, dtype=torch.int32, device=dev)
        d_items, d_bands = up(items, np.int32), up(bands, np.int32)
        d_hgoff = up(hg_off, np.int64)
        n_eval = 2 * int(max_steps) + 2
        gal = _prerender_arrays(order, shape, moff, c, tiers, nb, m_acc, precision, prerender,
                                dev, up)
        post_on = bool(post and gal["args"][0] > 0 and precision != "fp64")
        n_eval += int(post_on)
        evs = None
        ev_arr = None
        if timer is not None:
            # per-pass timing events (created by a first record)
            evs = [torch.cuda.Event(enable_timing=True) for _ in range(6 * n_eval)]
            for e in evs:
                e.record()
            ev_arr = (C.c_void_p * len(evs))(*[e.cuda_event for e in evs])
        post_args = (0, None, None, 0, None, None, None)
        bf16_on = bool(gram == "bf16" and gal["args"][0] > 0 and precision != "fp64")
        unsettled = torch.zeros(n_isl, dtype=torch.int32, device=dev)
        if post_on:
            Ns = np.asarray([s_[0] for s_ in shape], np.int64)
            pblk = nlc * (nb * Ns + nlc)
            post_off = np.concatenate([[0], np.cumsum(pblk)[:-1]]).astype(np.int64)
            post_t = dict(off=up(post_off, np.int64),
                          buf=torch.empty(max(int(pblk.sum()), 1), dtype=torch.float64,
                                          device=dev),
                          var=torch.full((m_acc * 7,), float("nan"), **f32),
                          dvar=torch.full((f_acc,), float("nan"), **f32),
                          metrics=torch.full((m_acc * nb * 6,), float("nan"), **f32),
                          kappa=torch.full((n_isl,), float("nan"), **f32))
            hblk = nlc * nlc + nlc
            post_t["hg_off"] = up(np.concatenate([[0], np.cumsum(hblk)[:-1]]), np.int64)
            post_t["hg"] = torch.full((max(int(hblk.sum()), 1),), float("nan"),
                                      dtype=torch.float64, device=dev)
            post_hblk = hblk
            post_args = (1, P(post_t["off"]), P(post_t["buf"]), int(post_t["buf"].numel()),
                         P(post_t["var"]), P(post_t["dvar"]), P(post_t["metrics"]))
        # large islands: Gram on the many-workgroup launch (float32 or float64)
        ncap = np.asarray([1 + s_[0] + s_[3] for s_ in shape], np.float64)
        npx = np.asarray([s_[1] * s_[1] for s_ in shape], np.float64)
        cost = ncap * ncap * npx
        big = np.nonzero((cost >= gram_split_min) & (gal["args"][0] > 0) & (not bf16_on))[0]
        # largest first, while the tile-pair buffer stays within its budget
        big = big[np.argsort(-cost[big], kind="stable")]
        ntp_of = ((ncap[big].astype(np.int64) + 15) // 16)
        ntp_of = ntp_of * (ntp_of + 1) // 2
        keep_n = np.arange(1, big.size + 1) * np.maximum.accumulate(ntp_of) * nb * 2048.0
        big = np.sort(big[keep_n <= GRAM_SPLIT_BYTES])
        big_args = (0, 0, None, None, None)
        if big.size:
            ntp_cap = int(max(((c_ + 15) // 16) * ((c_ + 15) // 16 + 1) // 2
                              for c_ in ncap[big].astype(np.int64)))
            slot = np.full(n_isl, -1, np.int32)
            slot[big] = np.arange(big.size, dtype=np.int32)
            big_t = dict(slot=up(slot, np.int32), isl=up(big.astype(np.int32), np.int32),
                         part=torch.empty(big.size * nb * ntp_cap * 256, dtype=torch.float64,
                                          device=dev))
            big_args = (int(big.size), ntp_cap, P(big_t["slot"]), P(big_t["isl"]),
                        P(big_t["part"]))
        sp = _lib.IslandSplit(n_isl * nb, P(d_items), P(d_bands), P(d_hgoff), P(hg), hg_n,
                              P(p_try), P(counters),
                              C.cast(ev_arr, C.c_void_p) if ev_arr is not None else None,
                              *gal["args"], *post_args, int(bf16_on), P(unsettled), *big_args,
                              P(post_t["kappa"]) if post_on else None,
                              P(post_t["hg"]) if post_on else None,
                              P(post_t["hg_off"]) if post_on else None)
    else:
        evs = None
        post_on = False
        bf16_on = False

    def launch(stream=None):
        if split:
            _lib.check(L.uph_island_fit_split(C.byref(bat), C.byref(inp), C.byref(out),
                                              C.byref(opts), prec, C.byref(sp), P(scratch),
                                              scratch.numel(), per, _lib.stream_ptr(stream)))
        else:
            _lib.check(L.uph_island_fit_batch(C.byref(bat), C.byref(inp), C.byref(out),
                                              C.byref(opts), prec, P(scratch), scratch.numel(),
                                              per, P(counter), _lib.stream_ptr(stream)))

    area = np.concatenate([h["area"][np.asarray(r)] for _, h, r in classes])
    mmc = np.concatenate([h["mmask"][np.asarray(r)].sum(axis=1) for _, h, r in classes])

    def timer_info():
        return dict(kind="island_batch", n=n_isl, nb=nb, outs=outs, area=area, nreal=mmc,
                    name="island_split_sequence" if split else None)

    def _record_costs(nst, gh):
        # Algorithmic cost of every launch of the sequence, recorded with
        # its events.  Bytes are the unique HBM bytes the launch must move;
        # flops count multiply-adds as 2 and a complex FFT of n points as
        # 5 n log2 n.
        #   evaluation: data + weight of every island box it evaluates (INIT
        #     and FINAL: all islands; the H/g and trial launches of step it:
        #     the islands that took step it); per band the Gram of the
        #     1 + N + nL planes, the residual and its projections.
        #   taps/spectra: spectra written (all three on H/g passes, F alone
        #     otherwise) and 28 taps per member.
        #   extended planes: PSF coefficient table once, spectra read, window
        #     part of the planes written; per (member, band) the polynomial
        #     over the stamp, forward transforms of P (and dP/dx, dP/dy),
        #     one spectral product and inverse transform per output plane,
        #     the 7-tap shift of the window part.
        #   stars: coefficient table once, planes written; the polynomial
        #     over the shift-read box and the shift.
        box = nb * area.astype(np.float64) * 8.0
        kp = c["keep"].reshape(-1, 6).astype(np.float64) * c["mm"].reshape(-1, 1)
        nlive = np.asarray([kp[moff[i]:moff[i] + shape[i][0]].sum() for i in range(n_isl)])
        npix = area.astype(np.float64)
        nr = mmc.astype(np.float64)
        kinds = ["init"] + ["grad", "try"] * int(max_steps) + ["final"] + ["post"] * post_on
        g_isl = gh["isl"]
        pad = np.asarray([_fft_pad(int(t)) for t in gh["T"]], np.float64)
        coef_tbl = {T: nb * nt * T * T * 4.0 for T in tiers}

        def fft_cost(n):
            return 5.0 * n * np.log2(n)

        for k, kind in enumerate(kinds):
            it = (k - 1) // 2
            sel = np.ones(n_isl, bool) if kind in ("init", "final", "post") else nst > it
            e5 = evs[6 * k:6 * k + 6]
            grad = kind in ("grad", "post")
            if g_isl.size:
                gsel = sel[g_isl]
                st_, ex_ = gsel & gh["star"], gsel & ~gh["star"]
                tg = gh["T"].astype(np.float64)
                planes = gh["grad_planes"] if grad else np.ones(len(g_isl))
                nspec = 3.0 if grad else 1.0
                wb = 8.0 if precision == "fp64" else 4.0  # bytes per plane / spectra element
                spec_b = gh["spec_floats"] / 3.0 * nspec * wb
                sbytes = float(np.sum(spec_b[ex_]) + (0.0 if wb == 8.0 else 28 * 4.0 * gsel.sum()))
                hh = pad // 2 + 1
                npoly = 3.0 if grad else 1.0
                nfwd = np.where(grad & (gh["grad_planes"] > 1), 3.0, 1.0) if grad else 1.0
                e_fl = (2.0 * nt * tg * tg * npoly + nfwd * (tg + hh) * fft_cost(pad)
                        + planes * (6.0 * hh * pad + (hh + tg) * fft_cost(pad))
                        + 28.0 * gh["win"] * planes)
                t_fl = 2.0 * nt * gh["reg"] * npoly + 28.0 * gh["win"] * planes
                tiers_of = lambda m: sum(coef_tbl[int(t)] for t in set(gh["T"][m].tolist()))
                ebytes = float(tiers_of(ex_) + np.sum(spec_b[ex_])
                               + nb * np.sum((gh["win"] * planes * wb)[ex_]))
                tbytes = float(tiers_of(st_) + nb * np.sum((gh["win"] * planes * wb)[st_]))
                if wb == 8.0:
                    # float64: the spectra launch, the planes launch of the
                    # stars and tier-55 members (transforms in LDS), then
                    # the launches of the larger tiers
                    small = ex_ & (gh["T"] <= 64)
                    large = ex_ & (gh["T"] > 64)
                    bx = lambda m: float(tiers_of(m) + np.sum(spec_b[m])
                                         + nb * np.sum((gh["win"] * planes * wb)[m]))
                    timer.record("pre64_spectra_kernel", e5[0], e5[1],
                                 dict(kind="island_eval", phase=kind, bytes=sbytes, flops=0.0,
                                      n_pre=int(ex_.sum())))
                    timer.record("pre64_planes_kernel", e5[1], e5[2],
                                 dict(kind="island_eval", phase=kind,
                                      bytes=bx(small) + tbytes,
                                      flops=float(nb * (np.sum(e_fl[small]) + np.sum(t_fl[st_]))),
                                      n_pre=int(small.sum() + st_.sum())))
                    if large.any():
                        timer.record("pre64_planes_big_kernel", e5[2], e5[3],
                                     dict(kind="island_eval", phase=kind, bytes=bx(large),
                                          flops=float(nb * np.sum(e_fl[large])),
                                          n_pre=int(large.sum())))
                else:
                    timer.record("prerender_spectra_kernel", e5[0], e5[1],
                                 dict(kind="island_eval", phase=kind, bytes=sbytes, flops=0.0,
                                      n_pre=int(gsel.sum())))
                    if gh["n_ext"]:
                        timer.record("prerender_planes_kernel", e5[1], e5[2],
                                     dict(kind="island_eval", phase=kind, bytes=ebytes,
                                          flops=float(nb * np.sum(e_fl[ex_])),
                                          n_pre=int(ex_.sum())))
                    if gh["n_ext"] < g_isl.size:
                        timer.record("prerender_stars_kernel", e5[2], e5[3],
                                     dict(kind="island_eval", phase=kind, bytes=tbytes,
                                          flops=float(nb * np.sum(t_fl[st_])),
                                          n_pre=int(st_.sum())))
            ncol = 1 + nr + (nlive if grad else 0)
            gram_fl = nb * npix * ncol * (ncol + 1)
            ev_fl = nb * npix * (ncol * (ncol + 1) + 4.0 * nr + (2.0 * (nr + nlive) if grad else 0))
            if big.size:
                # the large islands' Gram (and its data + weight reads) is
                # the Gram launch's, the rest of their evaluation stays
                bsel = np.zeros(n_isl, bool)
                bsel[big] = True
                bsel &= sel
                timer.record("island_gram64_kernel" if precision == "fp64" else
                             "island_gram_kernel", e5[3], e5[4],
                             dict(kind="island_eval", phase=kind,
                                  bytes=float(box[bsel].sum()),
                                  flops=float(gram_fl[bsel].sum()), n_isl=int(bsel.sum())))
                ev_fl = np.where(bsel, ev_fl - gram_fl, ev_fl)
            timer.record("island_eval_kernel_fp64" if precision == "fp64" else
                         "island_eval_kernel", e5[4], e5[5],
                         dict(kind="island_eval", phase=kind, bytes=float(box[sel].sum()),
                              flops=float(ev_fl[sel].sum()), n_isl=int(sel.sum())))

    def collect():
        return _collect()

    def _collect():
        h = {k: v.cpu().numpy() for k, v in outs.items()}
        if split and int(counters[-1].item()) != 0:
            raise _lib.UphError(_lib.UPH_ERR_INTERNAL,
                                "island band blocks overflowed their buffer")
        if split and evs is not None:
            # the cost bookkeeping runs when the timer reports (``timer.defer``),
            # not between the fit and the next pass
            # host fields only: the deferred call must not keep the batch's
            # device buffers (gal holds them) alive
            nst_h = h["nst"].astype(np.int64)
            gh = {k: gal.get(k) for k in ("isl", "star", "T", "grad_planes", "spec_floats",
                                          "win", "reg", "n_ext")}
            defer = getattr(timer, "defer", None)
            if defer is not None:
                defer(lambda: _record_costs(nst_h, gh))
            else:
                _record_costs(nst_h, gh)
        hp = ({k: post_t[k].cpu().numpy() for k in ("var", "dvar", "metrics", "kappa", "hg")}
              if post_on
              else None)
        h_uns = unsettled.cpu().numpy() if split and bf16_on else None
        results = []
        for (kern, host, rows), (i0, n_c, m0, f0, N) in zip(classes, spans):
            r = dict(p=h["p"][m0 * 6:(m0 + n_c * N) * 6].reshape(n_c, N, 6),
                     fluxes=h["fluxes"][f0:f0 + n_c * nb * N].reshape(n_c, nb, N),
                     ferr=h["ferr"][f0:f0 + n_c * nb * N].reshape(n_c, nb, N))
            for k in ("chi2", "lam", "conv", "stall", "nst", "consec", "nrej", "n_good"):
                r[k] = h[k][i0:i0 + n_c]
            for k in ("conv", "stall"):
                r[k] = r[k].astype(bool)
            r["n_good"] = r["n_good"].astype(np.int64)
            if split and bf16_on:
                r["gram_unsettled"] = h_uns[i0:i0 + n_c].astype(bool)
            if post_on:
                r["post_var"] = hp["var"][m0 * 7:(m0 + n_c * N) * 7].reshape(n_c, N, 7)
                r["post_dvar"] = hp["dvar"][f0:f0 + n_c * nb * N].reshape(n_c, nb, N)
                r["post_metrics"] = hp["metrics"][m0 * nb * 6:(m0 + n_c * N) * nb * 6].reshape(
                    n_c, N, nb, 6)
                r["post_kappa"] = hp["kappa"][i0:i0 + n_c]
                # band sums of H (nL x nL) and g (nL) per island, nL the
                # island's live parameters (rows padded to the class's block)
                hb = int(post_hblk[i0]) if n_c else 0
                h0 = int(np.sum(post_hblk[:i0]))
                r["post_hg"] = hp["hg"][h0:h0 + n_c * hb].reshape(n_c, hb)
            r["steps_run"] = int(r["nst"].max()) if n_c else 0
            r["wall_s"] = time.perf_counter() - t0
            results.append(r)
        return results

    # every device buffer the launches address: the launch closure holds
    # raw pointers only, and a captured graph replays them later
    keep = [ins, d_shape, d_moff, d_foff, outs, scratch, counter, dstack_dev, wstack_dev,
            psf_trees, sersic_dev]
    if split:
        keep += [hg, p_try, counters, d_items, d_bands, d_hgoff, gal, unsettled]
        if big.size:
            keep.append(big_t)
        if post_on:
            keep.append(post_t)
    return PreparedIslandBatch(launch, timer_info, collect, keep)


@dataclass
class PreparedIslandBatch:
    """A batch of island classes staged on the device: ``launch(stream)``
    enqueues the fit (it can be captured into a graph), ``timer_info()``
    describes it for a kernel timer and ``collect()`` copies the results
    back (one dict per class).  ``keepalive`` holds every device buffer the
    launches address (the launch closure passes raw pointers)."""
    launch: object
    timer_info: object
    collect: object
    keepalive: object = None


def run_island_batch(classes, psf_trees: dict, max_steps: int, dstack_dev, wstack_dev, *,
                     precision: str = "fp32", stream=None, sersic_dev=None, timer=None,
                     split: bool = False, prerender: bool = True, post: bool = False,
                     gram: str = "fp32", workspace_bytes: float = MAX_SCRATCH_BYTES,
                     gram_split_min: float = GRAM_SPLIT_MIN, init_states=None,
                     max_blocks: int | None = None):
    """Fit the islands of several classes in one launch sequence (see
    :func:`prepare_island_batch`) and return one result dict per class, in
    the same layout as :func:`run_class_chunk`."""
    pb = prepare_island_batch(classes, psf_trees, max_steps, dstack_dev, wstack_dev,
                              precision=precision, sersic_dev=sersic_dev, timer=timer,
                              split=split, prerender=prerender, post=post, gram=gram,
                              workspace_bytes=workspace_bytes, gram_split_min=gram_split_min,
                              init_states=init_states, max_blocks=max_blocks)
    ev = timer.begin("island_lm_kernel") if timer is not None else None
    pb.launch(stream)
    if ev is not None:
        timer.end(ev, pb.timer_info())
    return pb.collect()


# Islands whose band-summed Gauss-Newton matrix, scaled to a unit diagonal,
# has a 1-norm condition number above this are refitted on the float64
# kernels: their LM steps amplify the float32 model error (about 1e-7 of
# the model) enough to move the trajectory by more than 1e-5 of a flux.
# Close pairs (separation below about 1 px) and flat shape directions land
# there; single sources and resolved blends sit near 1-30.
FP64_REFIT_KAPPA = 50.0
_STATE_KEYS = ("p", "fluxes", "ferr", "chi2", "lam", "conv", "stall", "nst", "consec", "nrej",
               "n_good")
_POST_KEYS = ("post_var", "post_dvar", "post_metrics", "post_kappa", "post_hg")


FP64_PLANE_BUDGET = 16e9  # bytes of float64 member planes per launch sequence


def _fp64_groups(k64, budget: float = FP64_PLANE_BUDGET):
    """Split (kernels, host, rows) classes into launch sequences whose
    float64 member planes (n_bands x (7 extended | 3 star) x min(T, B)^2
    doubles per real member) stay within ``budget`` bytes: a list of groups
    of (class index, first row, end row)."""
    groups, used = [[]], 0.0
    for ci, (kern, host, rows) in enumerate(k64):
        side = min(int(kern.T), int(kern.B))
        mm = np.asarray(host["mmask"])[rows].astype(bool)
        st = np.asarray(host["st"])[rows]
        planes = (np.where(st == 0, 3, 7) * mm).sum(axis=1)
        b = kern.n_bands * side * side * 8.0 * planes
        start = 0
        for i in range(rows.size):
            if used > 0 and used + b[i] > budget:
                if i > start:
                    groups[-1].append((ci, start, i))
                groups.append([])
                used, start = 0.0, i
            used += b[i]
        if rows.size > start:
            groups[-1].append((ci, start, rows.size))
    return [g for g in groups if g]


def _run_fp64_split(k64, psf_trees, max_steps, dstack_dev, wstack_dev, *, sersic_dev=None,
                    prerender=True, workspace_bytes=MAX_SCRATCH_BYTES, timer=None,
                    init_states=None, max_blocks=None):
    """:func:`run_island_batch` on the float64 band-split kernels in launch
    sequences of at most FP64_PLANE_BUDGET bytes of member planes (see
    :func:`_fp64_groups`); one result dict per class, rows in order."""
    groups = _fp64_groups(k64) if prerender else [[(ci, 0, r.size) for ci, (_, _, r)
                                                   in enumerate(k64)]]
    parts = [[] for _ in k64]
    for grp in groups:
        sub = [(k64[ci][0], k64[ci][1], k64[ci][2][a_:b_]) for ci, a_, b_ in grp]
        ist = None
        if init_states is not None:
            ist = [None if init_states[ci] is None else
                   {k: (np.asarray(v)[a_:b_] if np.ndim(v) else v)
                    for k, v in init_states[ci].items()} for ci, a_, b_ in grp]
        res = run_island_batch(sub, psf_trees, max_steps, dstack_dev, wstack_dev,
                               precision="fp64", sersic_dev=sersic_dev, split=True,
                               prerender=prerender, workspace_bytes=workspace_bytes,
                               timer=timer, init_states=ist, max_blocks=max_blocks)
        for (ci, _, _), r in zip(grp, res):
            parts[ci].append(r)
    return [p[0] if len(p) == 1 else _merge_chunk_outputs(p) for p in parts]


def fit_float64(classes, psf_trees: dict, max_steps: int, dstack_dev, wstack_dev, *,
                sersic_dev=None, post: bool = True,
                workspace_bytes: float = MAX_SCRATCH_BYTES, timer=None, init_states=None,
                prerender: bool = True, max_blocks: int | None = None):
    """Fit ``classes`` (kernels, host, rows) on the float64 kernels
    (band-split; members pre-rendered by the float64 (member, band)
    launches, or in place with ``prerender=False``: the same planes, bit
    for bit), then, with ``post``, the
    float32 post-fit pass (no LM step) at the float64 parameters for the
    parameter variances, marginal flux errors, metrics and conditioning.
    Fluxes, flux errors and the LM state are the float64 fit's.  Every row
    carries ``fp64_refit`` True."""
    import dataclasses
    if not classes:
        return []
    k64 = [(dataclasses.replace(k, precision="fp64"), h, np.asarray(r)) for k, h, r in classes]
    r64 = _run_fp64_split(k64, psf_trees, max_steps, dstack_dev, wstack_dev,
                          sersic_dev=sersic_dev, prerender=prerender,
                          workspace_bytes=workspace_bytes, timer=timer, init_states=init_states,
                          max_blocks=max_blocks)
    if post:
        at64 = []
        for (kern, host, rows), r in zip(k64, r64):
            h2 = dict(host)
            h2["p0"] = host["p0"].copy()
            h2["p0"][rows] = r["p"]
            at64.append((dataclasses.replace(kern, precision="fp32"), h2, rows))
        rpost = run_island_batch(at64, psf_trees, 0, dstack_dev, wstack_dev, precision="fp32",
                                 sersic_dev=sersic_dev, split=True, prerender=True, post=True,
                                 workspace_bytes=workspace_bytes, timer=timer)
        for r, rp in zip(r64, rpost):
            for k in _POST_KEYS:
                if k in rp:
                    r[k] = rp[k]
    for (_, _, rows), r in zip(k64, r64):
        r["fp64_refit"] = np.ones(rows.size, bool)
    return r64


def refit_ill_conditioned(classes, results, psf_trees: dict, max_steps: int, dstack_dev,
                          wstack_dev, *, kappa: float = FP64_REFIT_KAPPA, sersic_dev=None,
                          workspace_bytes: float = MAX_SCRATCH_BYTES, timer=None,
                          max_blocks: int | None = None):
    """Refit, on the float64 kernels from the same seeds, the islands of a
    float32 post-fit batch (``results`` of :func:`run_island_batch` with
    ``post``) whose ``post_kappa`` exceeds ``kappa`` (see
    :func:`fit_float64`).  Returns (results, number of islands refitted);
    every row carries ``fp64_refit``."""
    redo, sel_of = [], []
    for (kern, host, rows), res in zip(classes, results):
        rows = np.asarray(rows)
        k = res.get("post_kappa")
        sel = np.zeros(rows.size, bool) if k is None else ~(k <= kappa)  # NaN refits too
        sel_of.append(sel)
        if sel.any():
            redo.append((kern, host, rows[sel]))
    out = []
    for res, sel in zip(results, sel_of):
        r = dict(res)
        r["fp64_refit"] = np.zeros(sel.size, bool)
        out.append(r)
    if not redo:
        return out, 0
    r64 = fit_float64(redo, psf_trees, max_steps, dstack_dev, wstack_dev,
                      sersic_dev=sersic_dev, workspace_bytes=workspace_bytes, timer=timer,
                      max_blocks=max_blocks)
    j = 0
    for r, sel in zip(out, sel_of):
        if not sel.any():
            continue
        idx = np.nonzero(sel)[0]
        for k in _STATE_KEYS + _POST_KEYS + ("fp64_refit",):
            if k in r and k in r64[j]:
                v = r[k].copy()
                v[idx] = r64[j][k]
                r[k] = v
        r["steps_run"] = max(int(r.get("steps_run", 0)), int(r64[j]["steps_run"]))
        j += 1
    return out, int(sum(int(x[2].size) for x in redo))


# Stall probes (island_solver.py:1335-1465): a stalled island took
# STALL_REJECT_K consecutive rejected steps at its final parameters with the
# damping raised 10x each time, so the rejected proposals are a function of
# the final state; the plateau probe sweeps the full damping ladder.
STALL_REJECT_K = 4
PLATEAU_LAMBDAS = tuple(float(np.float32(v)) for v in
                        [1e-9] + [10.0 ** e for e in range(-8, 7)])
PLATEAU_DLNP_TOL = 0.5   # island_plateau: best dchi2 > -2 * tol (config.py:594-610)


def _live_slots(host, row) -> np.ndarray:
    """Flat (slot, parameter) indices of an island's live parameters in the
    kernels' order: real members in slot order, then x, y, logre, ee1, ee2,
    u_n as kept."""
    keep = host["keep"][row] & host["mmask"][row][:, None]
    return np.flatnonzero(keep.reshape(-1))


def _damped_proposal(p, H, g, lam, live, clipv, lower, upper):
    """The legacy step's trial point from (H, g) at p (one_step,
    island_solver.py:921-969): float64 solve of (H + lam diag(H) + 1e-8 I)
    delta = -g, the step cast to float32, clipped per parameter, added in
    float32 and clipped to the box."""
    Hd = H.astype(np.float64).copy()
    dg = np.diag(Hd).copy()
    Hd[np.diag_indices_from(Hd)] = lam * dg + dg + 1e-8
    d = (-np.linalg.solve(Hd, g.astype(np.float64))).astype(np.float32)
    cl = clipv[live % 6]
    d = np.clip(d, -cl, cl)
    pt = p.astype(np.float32).copy()
    pt[live] = pt[live] + d
    return np.clip(pt, lower, upper)


def stall_probes(classes, results, psf_trees: dict, dstack_dev, wstack_dev, *, sersic_dev=None,
                 workspace_bytes: float = MAX_SCRATCH_BYTES, timer=None):
    """Post-hoc probes of the stalled islands of a fitted batch whose
    results carry the post-fit pass (``post_hg``): ``diverged`` -- the
    STALL_REJECT_K rejected proposals, rebuilt at lam_final 10^(j - K), are
    all identical (a clip- or bound-pinned runaway); ``stall_dchi2`` -- the
    best chi2 change min_j chi2(proposal(lam_j)) - chi2(p) over the damping
    ladder 1e-9 ... 1e6 (negative: a damped step would still improve).
    chi2 values come from zero-step device passes on the island's own
    kernels (float64 where the island was fitted in float64).  Adds
    ``diverged`` (n_c,) and ``stall_dchi2`` (n_c,, NaN where not stalled) to
    each result; returns the number of islands probed."""
    import dataclasses
    jobs = {"fp32": [], "fp64": []}
    for ci, ((kern, host, rows), res) in enumerate(zip(classes, results)):
        rows = np.asarray(rows)
        res["diverged"] = np.zeros(rows.size, bool)
        res["stall_dchi2"] = np.full(rows.size, np.nan, np.float32)
        if "post_hg" not in res:
            continue
        clipv = np.array([kern.pos_step_max, kern.pos_step_max, kern.shape_step_max,
                          kern.shape_step_max, kern.shape_step_max, kern.un_step_max],
                         np.float32)
        f64 = res.get("fp64_refit", np.zeros(rows.size, bool))
        for i in np.flatnonzero(res["stall"]):
            live = _live_slots(host, rows[i])
            nL = live.size
            blk = res["post_hg"][i]
            if nL == 0 or not np.isfinite(blk[:nL * nL + nL]).all():
                continue
            H = blk[:nL * nL].reshape(nL, nL)
            g = blk[nL * nL:nL * nL + nL]
            p = res["p"][i].reshape(-1).astype(np.float32)
            lo = host["lower"][rows[i]].reshape(-1)
            hi = host["upper"][rows[i]].reshape(-1)
            try:
                lam = float(res["lam"][i])
                rej = [_damped_proposal(p, H, g, lam * 10.0 ** (j - STALL_REJECT_K), live,
                                        clipv, lo, hi) for j in range(STALL_REJECT_K)]
                res["diverged"][i] = all(np.array_equal(r, rej[0]) for r in rej[1:])
                trial = [p] + [_damped_proposal(p, H, g, lm, live, clipv, lo, hi)
                               for lm in PLATEAU_LAMBDAS]
            except np.linalg.LinAlgError:
                continue
            jobs["fp64" if f64[i] else "fp32"].append((ci, i, np.stack(trial)))
    n_probed = 0
    for prec, items in jobs.items():
        if not items:
            continue
        by_class = {}
        for ci, i, trial in items:
            by_class.setdefault(ci, []).append((i, trial))
        batch, back = [], []
        for ci, lst in by_class.items():
            kern, host, rows = classes[ci]
            rows = np.asarray(rows)
            n_host = host["p0"].shape[0]
            rep = np.repeat(rows[[i for i, _ in lst]], len(PLATEAU_LAMBDAS) + 1)
            h2 = {k: (v[rep] if isinstance(v, np.ndarray) and v.ndim and v.shape[0] == n_host
                      else v) for k, v in host.items()}
            h2["p0"] = np.concatenate([t for _, t in lst]).reshape(rep.size, -1, 6)
            batch.append((dataclasses.replace(kern, precision=prec), h2, np.arange(rep.size)))
            back.append((ci, [i for i, _ in lst]))
        if prec == "fp64":
            out = _run_fp64_split(batch, psf_trees, 0, dstack_dev, wstack_dev,
                                  sersic_dev=sersic_dev, workspace_bytes=workspace_bytes,
                                  timer=timer)
        else:
            out = run_island_batch(batch, psf_trees, 0, dstack_dev, wstack_dev, precision=prec,
                                   sersic_dev=sersic_dev, split=True, prerender=True,
                                   workspace_bytes=workspace_bytes, timer=timer)
        for (ci, idx), r in zip(back, out):
            c2 = r["chi2"].reshape(len(idx), len(PLATEAU_LAMBDAS) + 1)
            results[ci]["stall_dchi2"][idx] = (c2[:, 1:].min(axis=1) - c2[:, 0]).astype(
                np.float32)
            n_probed += len(idx)
    return n_probed


def island_fit(*, bandnames, data_minus_bg, invvar, psfs, island_set, buckets, pos_x, pos_y,
               source_type, logre, ee1, ee2, sersic_n, s1_flux, s1_flux_var, pos_err=None,
               re_prior_sig=None, ell_prior_sig=None, theta_prior_sig=None,
               sersic_n_prior_sig=None, render_backend: str = "fft", fit_sersic_n: bool = True,
               n_fit_min_snr: float = 30.0, pos_max_shift_pix: float = 5.0,
               pos_step_max_pix: float = DEFAULT_POS_STEP_MAX,
               max_steps: int = 30, max_steps_main=8,
               chi2_ftol_rel: float = DEFAULT_CHI2_FTOL_REL,
               step_tol_pix: float = DEFAULT_STEP_TOL_PIX, un_freeze_steps: int = 0,
               reg: float = DEFAULT_REG, nsrc_max=None, joint_nsrc_cap: int = 32,
               star_render_split: bool = True, stall_count_mode: str = "proposals",
               convergence: str = "legacy", plateau_dlnp_tol: float = 0.5,
               include_giants: bool = True, compute_param_errors: bool = False,
               compute_source_metrics: bool = False, compute_re_profile: bool = False,
               return_island_residual: bool = False, free_n_override=None, log_fn=None):
    """Fit every bucketed island and return per-source rows, giants
    included (island_solver.py:3233-3595): free-n gate from the S1 fluxes,
    class planning with the joint member cap, padded field stacks (PAD =
    largest window + 8), the band-split device fit of all classes (float64
    for islands the precision policy routes there), the stall probes,
    :func:`unpack_class_results` per class and :func:`giant_seed_rows`.
    ``max_steps_main`` (the reference's main pass / re-queue split) gives
    the same LM trajectories as one ``max_steps`` pass and is accepted for
    the signature.  The legacy step only: priors, u_n annealing, the
    "changed" stall counting, dlnp convergence, the Re profile scan and the
    residual export raise NotImplementedError."""
    import torch
    from .island_plan import (IslandFitResult, free_n_gate, giant_seed_rows,
                              plan_island_classes, unpack_class_results)
    from .pipeline import FitConfig, fit_island_set, psf_bank

    def has(a):
        return a is not None and np.any(np.isfinite(np.asarray(a, np.float64)))
    if any(has(a) for a in (pos_err, re_prior_sig, ell_prior_sig, theta_prior_sig,
                            sersic_n_prior_sig)):
        raise NotImplementedError("catalog priors are not part of this build")
    if un_freeze_steps or stall_count_mode != "proposals" or convergence != "legacy":
        raise NotImplementedError("only the legacy LM step (no u_n annealing, proposal stall "
                                  "counting, legacy convergence) is part of this build")
    if compute_re_profile or return_island_residual:
        raise NotImplementedError("the Re profile scan and residual export are not part of "
                                  "this build")
    if render_backend != "fft" and np.any(np.asarray(source_type) != 0):
        raise NotImplementedError("extended sources render through the 'fft' backend only")
    nb = len(bandnames)
    st = np.asarray(source_type, np.int32)
    x = np.asarray(pos_x, np.float32)
    y = np.asarray(pos_y, np.float32)
    lr, e1, e2 = (np.asarray(a, np.float32) for a in (logre, ee1, ee2))
    sn = np.asarray(sersic_n, np.float32)
    fl = np.nan_to_num(np.asarray(s1_flux, np.float64), nan=0.0)
    fv = np.nan_to_num(np.asarray(s1_flux_var, np.float64), nan=1e30)
    free_n = (np.asarray(free_n_override, bool) if free_n_override is not None else
              free_n_gate(st, fl, fv, bool(fit_sersic_n), float(n_fit_min_snr)))
    native = int(psfs[0].H_psf)
    sub = buckets
    if nsrc_max is not None:
        import dataclasses
        sub = dataclasses.replace(buckets, buckets={k: b for k, b in buckets.buckets.items()
                                                    if k[1] <= int(nsrc_max)})
    plans, capped = plan_island_classes(island_set, sub, st, lr, sn, native,
                                        joint_nsrc_cap=joint_nsrc_cap)
    dev = torch.device("cuda")
    pad = max((p.bbox_class for p in plans), default=0) + 8
    d = np.asarray(data_minus_bg, np.float32)
    w = np.asarray(invvar, np.float32)
    _, H, W = d.shape
    dstack = torch.zeros((nb, H + 2 * pad, W + 2 * pad), dtype=torch.float32, device=dev)
    wstack = torch.zeros_like(dstack)
    dstack[:, pad:pad + H, pad:pad + W] = torch.from_numpy(d).to(dev)
    wstack[:, pad:pad + H, pad:pad + W] = torch.from_numpy(w).to(dev)
    cfg = FitConfig(max_steps=int(max_steps), reg=float(reg),
                    pos_max_shift_pix=float(pos_max_shift_pix),
                    post_fit=bool(compute_param_errors or compute_source_metrics))
    if float(pos_step_max_pix) != DEFAULT_POS_STEP_MAX or \
            float(chi2_ftol_rel) != DEFAULT_CHI2_FTOL_REL or \
            float(step_tol_pix) != DEFAULT_STEP_TOL_PIX:
        raise NotImplementedError("non-default LM clip / tolerances through island_fit; use "
                                  "make_class_kernels + run_class_chunk")
    rows = []
    if plans:
        isf = fit_island_set(psf_bank(psfs, dev), nb, int(psfs[0].n_terms), native, dstack,
                             wstack, pad, island_set, plans, x, y, st, lr, e1, e2, sn, free_n,
                             cfg, flux_scale=(np.asarray(s1_flux, np.float64),
                                              np.sqrt(np.maximum(fv, 0.0))),
                             keep_class_results=True)
        from .pipeline import metrics_from_sums
        import dataclasses
        for plan, host, r_idx, res in isf.class_results:
            pl = dataclasses.replace(plan, isl_ids=plan.isl_ids[r_idx], midx=plan.midx[r_idx],
                                     mmask=plan.mmask[r_idx], bbox=plan.bbox[r_idx],
                                     n_src_of=plan.n_src_of[r_idx])
            h = {k: (v[r_idx] if isinstance(v, np.ndarray) and v.ndim and
                     v.shape[0] == plan.n_islands else v) for k, v in host.items()}
            perr = sm = sw = fm = None
            if "post_var" in res:
                if compute_param_errors:
                    perr = np.full(res["post_var"].shape[:2] + (13,), np.nan, np.float32)
                    perr[:, :, :7] = res["post_var"]
                    fm = res["post_dvar"]
                if compute_source_metrics:
                    m4, fw = metrics_from_sums(res["post_metrics"])
                    sm, sw = m4.astype(np.float32), np.transpose(fw, (0, 2, 1)).astype(np.float32)
            rows.append(unpack_class_results(
                pl, h, island_set, res["p"], res["fluxes"], res["ferr"], res["chi2"],
                res["conv"], res["stall"], res["nst"], res_perr=perr,
                res_pdchi2=res.get("stall_dchi2"), res_smetrics=sm, res_swin=sw,
                res_femarg=fm, res_div=res.get("diverged"), res_ngood=res["n_good"],
                n_bands=nb))
    if include_giants:
        gids = np.concatenate([np.asarray(buckets.giant_ids, np.int64),
                               np.asarray(capped, np.int64)])
        rows += giant_seed_rows(gids, island_set, pos_x=x, pos_y=y, logre=lr, ee1=e1, ee2=e2,
                                sersic_n=sn, n_bands=nb)
    return IslandFitResult.from_rows(bandnames, rows)


class IslandBatchGraph:
    """The launch sequence of a prepared batch captured once into a HIP
    graph and replayed: every replay refits the batch from the seeds in its
    device buffers (the counters are cleared by a kernel node, and nothing
    in the sequence allocates or synchronises)."""

    def __init__(self, classes, psf_trees: dict, max_steps: int, dstack_dev, wstack_dev, *,
                 precision: str = "fp32", sersic_dev=None, prerender: bool = True):
        import torch
        self.batch = prepare_island_batch(classes, psf_trees, max_steps, dstack_dev,
                                          wstack_dev, precision=precision,
                                          sersic_dev=sersic_dev, split=True,
                                          prerender=prerender)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.batch.launch(torch.cuda.current_stream())

    def replay(self):
        self.graph.replay()

    def collect(self):
        return self.batch.collect()

