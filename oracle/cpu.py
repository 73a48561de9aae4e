"""ctypes wrapper of the CPU restatement ``libuph_cpu.so`` (TEST
INFRASTRUCTURE: the CPU baseline that bench.py times beside the GPU path).

``fit_islands`` runs the per-island VarPro-LM (float32 pixels, float64
solves) over a list of islands with OpenMP; ``band_solve`` runs grid-band
linear solves.  Built by ``oracle/build_cpu.py``; its results are checked
against the float64 NumPy oracle in tests/test_oracle_cpu.py.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import core
from .island import LMOptions

HERE = Path(__file__).resolve().parent
LIBS = {"fp32": HERE / "libuph_cpu.so", "fp64": HERE / "libuph_cpu64.so"}
_libs = {}


def lib(precision: str = "fp32"):
    """The float32 baseline, or the float64 build of the same source."""
    if precision not in _libs:
        path = LIBS[precision]
        if not path.exists():
            raise RuntimeError(f"{path} is missing: run `python oracle/build_cpu.py`")
        L = C.CDLL(str(path))
        L.uphc_num_threads.restype = C.c_int
        L.uphc_fit_islands.restype = C.c_int
        L.uphc_band_solve.restype = C.c_int
        L.uphc_render_stamps.restype = C.c_int
        _libs[precision] = L
    return _libs[precision]


def num_threads() -> int:
    return int(lib().uphc_num_threads())


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def fit_islands(items, psfs, opts: LMOptions | None = None, n_threads: int = 0,
                precision: str = "fp32"):
    """Fit islands.  ``items``: dicts with data, w (nb, B, B), x0b, y0b,
    st, sn (n_frozen), fn (free_n), mmask (N,), keep (N, 6), p0, lower,
    upper (6N,), T (stamp tier).  ``psfs``: per band dict(coefs (S, S, nt),
    x_origin, x_scale, y_origin, y_scale).  Returns per island dict(p (6N,),
    fluxes, ferr (nb, N), nst, conv, stall, sec: the island's wall time on its
    thread)."""
    o = opts or LMOptions()
    n = len(items)
    if n == 0:
        return []
    nb = len(psfs)
    nt = int(np.asarray(psfs[0]["coefs"]).shape[-1])
    Bs = np.array([it["data"].shape[-1] for it in items], np.int32)
    Ts = np.array([int(it["T"]) for it in items], np.int32)
    Ns = np.array([np.asarray(it["st"]).size for it in items], np.int32)
    pix = np.concatenate([[0], np.cumsum(nb * Bs.astype(np.int64) ** 2)]).astype(np.int64)
    mem = np.concatenate([[0], np.cumsum(Ns)]).astype(np.int32)
    data = np.concatenate([np.asarray(it["data"], np.float32).ravel() for it in items])
    w = np.concatenate([np.asarray(it["w"], np.float32).ravel() for it in items])
    cat = lambda k, dt: np.ascontiguousarray(np.concatenate(
        [np.asarray(it[k]).ravel() for it in items]).astype(dt))
    stype, nfz = cat("st", np.int32), cat("sn", np.float32)
    fn, mm, keep = cat("fn", np.uint8), cat("mmask", np.uint8), cat("keep", np.uint8)
    p0, lo, hi = cat("p0", np.float32), cat("lower", np.float32), cat("upper", np.float32)
    x0b = np.array([it["x0b"] for it in items], np.float32)
    y0b = np.array([it["y0b"] for it in items], np.float32)
    tiers = sorted(set(Ts.tolist()))
    blocks, offs, off = [], [], 0
    for T in tiers:
        c = np.stack([np.moveaxis(core.pad_psf_coefs(np.asarray(p["coefs"], np.float64), T), -1, 0)
                      for p in psfs]).astype(np.float32)
        blocks.append(c.ravel())
        offs.append(off)
        off += c.size
    coefs = np.concatenate(blocks)
    geo = np.array([(p["x_origin"], p["x_scale"], p["y_origin"], p["y_scale"]) for p in psfs],
                   np.float32)
    tb = core.sersic_table()
    kn = np.ascontiguousarray(tb["n_knots"], np.float64)
    amps = np.ascontiguousarray(tb["amps"], np.float64)
    lv = np.ascontiguousarray(tb["log_vars"], np.float64)
    cr = np.ascontiguousarray(tb["core_amp"], np.float64)
    opt = np.array([o.reg, o.pos_step_max, o.shape_step_max, o.un_step_max, o.ftol_rel,
                    o.step_tol, o.shape_step_tol], np.float32)
    p_out = np.zeros(6 * int(mem[-1]), np.float32)
    f_out = np.zeros(nb * int(mem[-1]), np.float32)
    e_out = np.zeros(nb * int(mem[-1]), np.float32)
    nst = np.zeros(n, np.int32)
    conv = np.zeros(n, np.uint8)
    stall = np.zeros(n, np.uint8)
    sec = np.zeros(n, np.float64)
    tT = np.array(tiers, np.int32)
    tO = np.array(offs, np.int64)
    st = lib(precision).uphc_fit_islands(
        C.c_int(n), C.c_int(nb), C.c_int(nt), _p(Bs), _p(Ts), _p(Ns), _p(pix), _p(mem), _p(data),
        _p(w), _p(x0b), _p(y0b), _p(stype), _p(nfz), _p(fn), _p(mm), _p(keep), _p(p0), _p(lo),
        _p(hi), C.c_int(len(tiers)), _p(tT), _p(coefs), _p(tO), _p(geo), C.c_int(kn.size),
        C.c_int(amps.shape[1]), _p(kn), _p(amps), _p(lv), _p(cr), C.c_double(tb["n_min"]),
        C.c_double(tb["n_max"]), _p(opt), C.c_int(o.max_steps), C.c_int(n_threads), _p(p_out),
        _p(f_out), _p(e_out), _p(nst), _p(conv), _p(stall), _p(sec))
    if st != 0:
        raise RuntimeError("uphc_fit_islands failed (a stamp tier without PSF coefficients)")
    out = []
    for i in range(n):
        a, b = int(mem[i]), int(mem[i + 1])
        out.append(dict(p=p_out[6 * a:6 * b].copy(),
                        fluxes=f_out[nb * a:nb * b].reshape(nb, b - a).copy(),
                        ferr=e_out[nb * a:nb * b].reshape(nb, b - a).copy(),
                        nst=int(nst[i]), conv=bool(conv[i]), stall=bool(stall[i]),
                        sec=float(sec[i])))
    return out


def _table():
    tb = core.sersic_table()
    return (np.ascontiguousarray(tb["n_knots"], np.float64),
            np.ascontiguousarray(tb["amps"], np.float64),
            np.ascontiguousarray(tb["log_vars"], np.float64),
            np.ascontiguousarray(tb["core_amp"], np.float64), tb["n_min"], tb["n_max"])


def render_stamps(psfs, T, band, x, y, gx0, gy0, stype, logre, ee1, ee2, sersic_n,
                  renorm=False, n_threads: int = 0, precision: str = "fp32"):
    """(n, T, T) unit-flux stamps of the linear solve (PSF padded to T)."""
    n = int(np.asarray(x).size)
    nt = int(np.asarray(psfs[0]["coefs"]).shape[-1])
    coefs = np.ascontiguousarray(np.stack(
        [np.moveaxis(core.pad_psf_coefs(np.asarray(p["coefs"], np.float64), T), -1, 0)
         for p in psfs]), dtype=np.float32)
    geo = np.array([(p["x_origin"], p["x_scale"], p["y_origin"], p["y_scale"]) for p in psfs],
                   np.float32)
    a = lambda v, dt: np.ascontiguousarray(np.broadcast_to(np.asarray(v), (n,)).astype(dt))
    args = [a(band, np.int32), a(x, np.float32), a(y, np.float32), a(gx0, np.float32),
            a(gy0, np.float32), a(stype, np.int32), a(logre, np.float32), a(ee1, np.float32),
            a(ee2, np.float32), a(sersic_n, np.float32)]
    kn, amps, lv, cr, nmin, nmax = _table()
    out = np.zeros((n, T, T), np.float32)
    lib(precision).uphc_render_stamps(
        C.c_int(n), C.c_int(T), C.c_int(nt), _p(coefs), _p(geo), *[_p(v) for v in args],
        C.c_int(int(renorm)), C.c_int(kn.size), C.c_int(amps.shape[1]), _p(kn), _p(amps),
        _p(lv), _p(cr), C.c_double(nmin), C.c_double(nmax), C.c_int(n_threads), _p(out))
    return out


def band_solve(problems, bg_geo, k_max: int = 32, reg: float = 1e-9, n_threads: int = 0,
               precision: str = "fp32"):
    """Grid-band solves.  ``problems``: dicts with data, invvar (h, w) f32,
    good (h, w) bool, stamps (n, T, T), ix, iy (n,) window-local top-left,
    nbr (n, k_max) local indices (-1 none), gx0, gy0 (the window's full-image
    origin); ``bg_geo`` = (xo, xs, yo, ys) of the plane background.  Returns
    per problem (flux, var, bg (3,), bg_var (3,))."""
    P = len(problems)
    if P == 0:
        return []
    T = int(problems[0]["stamps"].shape[-1])
    ns = np.array([pr["stamps"].shape[0] for pr in problems], np.int64)
    so = np.concatenate([[0], np.cumsum(ns)]).astype(np.int64)
    hs = np.array([pr["data"].shape[0] for pr in problems], np.int32)
    ws = np.array([pr["data"].shape[1] for pr in problems], np.int32)
    po = np.concatenate([[0], np.cumsum(hs.astype(np.int64) * ws)]).astype(np.int64)
    cat = lambda k, dt: np.ascontiguousarray(np.concatenate(
        [np.asarray(pr[k]).ravel() for pr in problems]).astype(dt))
    data, iv, gd = cat("data", np.float32), cat("invvar", np.float32), cat("good", np.uint8)
    stamps = cat("stamps", np.float32)
    ix, iy, nbr = cat("ix", np.int32), cat("iy", np.int32), cat("nbr", np.int32)
    gx0 = np.array([pr["gx0"] for pr in problems], np.float32)
    gy0 = np.array([pr["gy0"] for pr in problems], np.float32)
    n = int(so[-1])
    flux = np.zeros(n, np.float32)
    var = np.zeros(n, np.float32)
    bg = np.zeros(3 * P, np.float32)
    bgv = np.zeros(3 * P, np.float32)
    xo, xs, yo, ys = (float(v) for v in bg_geo)
    lib(precision).uphc_band_solve(
        C.c_int(P), C.c_int(T), C.c_int(k_max), _p(so), _p(po), _p(hs), _p(ws), _p(gx0),
        _p(gy0), _p(data), _p(iv), _p(gd), _p(stamps), _p(ix), _p(iy), _p(nbr), C.c_double(xo),
        C.c_double(xs), C.c_double(yo), C.c_double(ys), C.c_double(reg), C.c_int(n_threads),
        _p(flux), _p(var), _p(bg), _p(bgv))
    return [(flux[so[q]:so[q + 1]], var[so[q]:so[q + 1]], bg[3 * q:3 * q + 3],
             bgv[3 * q:3 * q + 3]) for q in range(P)]
