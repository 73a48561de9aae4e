"""Seeded synthetic scenes for tests and benchmarks.

Conventions follow the reference's synthetic-truth generator
(``tract7dt/tests/synth``): pixel centres on integer coordinates,
``data[y, x]``, 55-px PSF boxes.  Truth is rendered independently of the
fit's model path: stars as pixel-integrated Gaussian mixtures (erf),
galaxies as the Sersic mixture convolved analytically with the PSF
Gaussians (plus a 1/12 px^2 pixel-box variance).  The fitted PSF model is a
``PsfPolynomial`` obtained by per-pixel least squares over a grid of field
positions.

Scene recipes (seeds fixed per configuration):
  * ``scene_c1``: 1 band, 256^2, sky 100, read noise 5, constant 3-Gaussian
    PSF (sigma 1, 2, 4; weights 0.7, 0.25, 0.05), 40 STAR + 10 EXP.
  * ``scene_c2``: 20 bands sharing 2048^2, sky 50-500, PSF sigma/weights
    varying +-10 % as quadratics in (X, Y), fitted at order 4 (15 terms);
    5000 sources 70 % STAR / 15 % EXP / 10 % DEV / 5 % SERSIC, clustered
    (Thomas process); sizes are parameters so tests can shrink it.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.special import erf

from uniphot_amd.psf_polynomial import PsfPolynomial, terms_for_order
from uniphot_amd.sersic_table import mog_params_at
from uniphot_amd.source_model import (SOURCE_DEV, SOURCE_EXP, SOURCE_SERSIC, SOURCE_STAR,
                           ellipse_soft_from_re_ab_phi_np)

SEEDS = dict(c1=20261201, c2=20261202, c3=20261203, c4=20261204, c5=20261205)


@dataclass
class Scene:
    data: np.ndarray          # (nb, H, W) float32, sky baseline removed
    invvar: np.ndarray        # (nb, H, W) float32
    good: np.ndarray          # (nb, H, W) bool
    psfs: list                # per band PsfPolynomial
    # catalog seeds (what the fit starts from)
    x: np.ndarray
    y: np.ndarray
    stype: np.ndarray
    logre: np.ndarray
    ee1: np.ndarray
    ee2: np.ndarray
    sersic_n: np.ndarray
    # truth
    true_x: np.ndarray
    true_y: np.ndarray
    true_flux: np.ndarray     # (n, nb)
    meta: dict = field(default_factory=dict)

    @property
    def n_bands(self) -> int:
        return self.data.shape[0]


def _erf_profile(centres, sig, n_pix):
    """(n, n_pix) pixel-integrated 1-D Gaussian at ``centres`` (pixel
    index space, integer centres)."""
    edges = np.arange(n_pix + 1, dtype=np.float64) - 0.5
    z = (edges[None, :] - centres[:, None]) / (np.sqrt(2.0) * sig[:, None])
    c = 0.5 * erf(z)
    return c[:, 1:] - c[:, :-1]


def gaussian_mixture_stamp(T, sigmas, weights, dx=0.0, dy=0.0):
    """Pixel-integrated Gaussian mixture on a T x T box centred at
    (T//2 + dx, T//2 + dy), normalised to unit box sum."""
    c = T // 2
    st = np.zeros((T, T))
    for s, w in zip(sigmas, weights):
        px = _erf_profile(np.array([c + dx]), np.array([s]), T)[0]
        py = _erf_profile(np.array([c + dy]), np.array([s]), T)[0]
        st += w * np.outer(py, px)
    return st / st.sum()


class MixturePsfField:
    """Spatially varying 3-Gaussian PSF: sigma_k(X, Y) = sigma_k0 (1 + a q_k),
    w_k likewise (renormalised), q_k quadratics bounded by 1 on [-1, 1]^2."""

    def __init__(self, H, W, sigmas, weights, amp=0.0, rng=None):
        self.H, self.W = H, W
        self.s0 = np.asarray(sigmas, np.float64)
        self.w0 = np.asarray(weights, np.float64)
        self.amp = float(amp)
        K = self.s0.size
        if amp and rng is not None:
            self.qs = rng.uniform(-1, 1, (K, 6))
            self.qw = rng.uniform(-1, 1, (K, 6))
            self.qs /= np.abs(self.qs).sum(axis=1, keepdims=True)
            self.qw /= np.abs(self.qw).sum(axis=1, keepdims=True)
        else:
            self.qs = np.zeros((K, 6))
            self.qw = np.zeros((K, 6))
        self.xo, self.xs = 0.5 * (W - 1), max(0.5 * W, 1.0)
        self.yo, self.ys = 0.5 * (H - 1), max(0.5 * H, 1.0)

    def params(self, x, y):
        X = (np.asarray(x, np.float64) - self.xo) / self.xs
        Y = (np.asarray(y, np.float64) - self.yo) / self.ys
        q = np.stack([np.ones_like(X), X, Y, X * X, X * Y, Y * Y], axis=-1)
        s = self.s0 * (1.0 + self.amp * (q @ self.qs.T))
        w = self.w0 * (1.0 + self.amp * (q @ self.qw.T))
        return s, w / w.sum(axis=-1, keepdims=True)

    def fit_polynomial(self, T=55, order=4, n_grid=21) -> PsfPolynomial:
        terms = terms_for_order(order)
        if order == 0 or self.amp == 0.0:
            s, w = self.params(self.xo, self.yo)
            st = gaussian_mixture_stamp(T, s, w)
            coefs = st[:, :, None]
            if order > 0:
                coefs = np.concatenate([coefs, np.zeros((T, T, len(terms) - 1))], axis=2)
            return PsfPolynomial(coefs.astype(np.float32), self.xo, self.xs, self.yo,
                                 self.ys, T, T)
        g = np.linspace(-1, 1, n_grid)
        GX, GY = np.meshgrid(g, g)
        X, Y = GX.ravel(), GY.ravel()
        design = np.stack([X ** m * Y ** n for (m, n) in terms], axis=1)
        stamps = []
        for xx, yy in zip(X, Y):
            s, w = self.params(self.xo + xx * self.xs, self.yo + yy * self.ys)
            stamps.append(gaussian_mixture_stamp(T, s, w).ravel())
        sol, *_ = np.linalg.lstsq(design, np.array(stamps), rcond=None)
        coefs = sol.T.reshape(T, T, len(terms))
        return PsfPolynomial(coefs.astype(np.float32), self.xo, self.xs, self.yo, self.ys, T, T)


def _cov_terms(logre, ee1, ee2):
    re = np.exp(logre)
    ee = np.hypot(ee1, ee2)
    ee1 = np.where(ee == 0, 1e-4, ee1)
    ee = np.hypot(ee1, ee2)
    e = 1.0 - np.exp(-ee)
    inv_ba = np.minimum((1 + e) / np.maximum(1 - e, 1e-4), 1000.0)
    th = 0.5 * np.arctan2(ee2, ee1)
    ct, st = np.cos(th), np.sin(th)
    a, b = re * ct / inv_ba, re * st
    c, d = -re * st / inv_ba, re * ct
    return a * a + b * b, c * c + d * d, a * c + b * d


def render_truth(H, W, psf_field: MixturePsfField, x, y, stype, logre, ee1, ee2, sn, flux,
                 box=27):
    """Noise-free truth image for one band (float64)."""
    img = np.zeros((H, W))
    s_all, w_all = psf_field.params(x, y)
    n = x.size
    R = box
    ix = np.rint(x).astype(np.int64)
    iy = np.rint(y).astype(np.int64)
    # stars: separable erf profiles
    for i in np.where(stype == SOURCE_STAR)[0]:
        cx, cy = x[i] - (ix[i] - R), y[i] - (iy[i] - R)
        st = np.zeros((2 * R + 1, 2 * R + 1))
        for s, w in zip(s_all[i], w_all[i]):
            px = _erf_profile(np.array([cx]), np.array([s]), 2 * R + 1)[0]
            py = _erf_profile(np.array([cy]), np.array([s]), 2 * R + 1)[0]
            st += w * np.outer(py, px)
        _add_box(img, st * flux[i], ix[i] - R, iy[i] - R)
    yy, xx = np.mgrid[-R:R + 1, -R:R + 1].astype(np.float64)
    for i in np.where(stype != SOURCE_STAR)[0]:
        a, v, core = mog_params_at(float(sn[i]))
        cxx, cyy, cxy = _cov_terms(float(logre[i]), float(ee1[i]), float(ee2[i]))
        ddx = xx - (x[i] - ix[i])
        ddy = yy - (y[i] - iy[i])
        st = np.zeros_like(xx)
        for s, w in zip(s_all[i], w_all[i]):
            s2 = s * s + 1.0 / 12.0
            if core:
                st += core * w * np.exp(-0.5 * (ddx ** 2 + ddy ** 2) / s2) / (2 * np.pi * s2)
            for ak, vk in zip(a, v):
                if ak == 0.0:
                    continue
                c00, c11, c01 = vk * cxx + s2, vk * cyy + s2, vk * cxy
                det = c00 * c11 - c01 * c01
                q = (c11 * ddx * ddx - 2 * c01 * ddx * ddy + c00 * ddy * ddy) / det
                st += ak * w * np.exp(-0.5 * q) / (2 * np.pi * np.sqrt(det))
        _add_box(img, st * flux[i], ix[i] - R, iy[i] - R)
    return img


def _add_box(img, st, x0, y0):
    H, W = img.shape
    h, w = st.shape
    ya, yb = max(0, y0), min(H, y0 + h)
    xa, xb = max(0, x0), min(W, x0 + w)
    if ya < yb and xa < xb:
        img[ya:yb, xa:xb] += st[ya - y0:yb - y0, xa - x0:xb - x0]


def _finish(rng, truths, skies, read_noise=5.0):
    data, invvar = [], []
    for t, sky in zip(truths, skies):
        d = rng.poisson(np.maximum(t + sky, 0.0)).astype(np.float64)
        d += rng.normal(0.0, read_noise, t.shape) - sky
        data.append(d.astype(np.float32))
        invvar.append(np.full(t.shape, 1.0 / (sky + read_noise ** 2), np.float32))
    data = np.stack(data)
    invvar = np.stack(invvar)
    return data, invvar, np.ones(data.shape, bool)


def _seed_catalog(rng, tx, ty, st, lr, e1, e2, pos_sigma=0.2, shape_frac=0.1):
    x = (tx + rng.normal(0, pos_sigma, tx.size)).astype(np.float32)
    y = (ty + rng.normal(0, pos_sigma, ty.size)).astype(np.float32)
    gal = st != SOURCE_STAR
    lr_s = np.where(gal, lr + np.log1p(rng.normal(0, shape_frac, lr.size).clip(-0.5, 0.5)),
                    np.log(3.0)).astype(np.float32)
    e1_s = np.where(gal, e1 * (1 + rng.normal(0, shape_frac, lr.size)), 0.0).astype(np.float32)
    e2_s = np.where(gal, e2 * (1 + rng.normal(0, shape_frac, lr.size)), 0.0).astype(np.float32)
    return x, y, lr_s, e1_s, e2_s


def _galaxy_shapes(rng, n):
    re = rng.uniform(1.0, 4.0, n)
    q = rng.uniform(0.4, 1.0, n)
    pa = rng.uniform(0.0, 180.0, n)
    return ellipse_soft_from_re_ab_phi_np(re, q, pa)


def scene_c1(seed=SEEDS["c1"], H=256, W=256, n_star=40, n_exp=10) -> Scene:
    rng = np.random.default_rng(seed)
    pf = MixturePsfField(H, W, (1.0, 2.0, 4.0), (0.7, 0.25, 0.05))
    psf = pf.fit_polynomial(55, order=0)
    n = n_star + n_exp
    tx = rng.uniform(0, W - 1, n)
    ty = rng.uniform(0, H - 1, n)
    st = np.array([SOURCE_STAR] * n_star + [SOURCE_EXP] * n_exp, np.int32)
    flux = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n))[:, None]
    lr, e1, e2 = _galaxy_shapes(rng, n)
    lr = np.where(st == SOURCE_STAR, 0.0, lr).astype(np.float32)
    e1 = np.where(st == SOURCE_STAR, 0.0, e1).astype(np.float32)
    e2 = np.where(st == SOURCE_STAR, 0.0, e2).astype(np.float32)
    sn = np.ones(n, np.float32)
    truth = render_truth(H, W, pf, tx, ty, st, lr, e1, e2, sn, flux[:, 0])
    data, invvar, good = _finish(rng, [truth], [100.0])
    x, y, lr_s, e1_s, e2_s = _seed_catalog(rng, tx, ty, st, lr, e1, e2)
    return Scene(data, invvar, good, [psf], x, y, st, lr_s, e1_s, e2_s, sn, tx, ty, flux,
                 meta=dict(config="c1", seed=seed, sky=[100.0], link_radius=4.0,
                           grid=(2, 2), halo=30))


def scene_c2(seed=SEEDS["c2"], H=2048, W=2048, n_bands=20, n_src=5000, star_only=False,
             parent_frac=0.3, sigma_c=6.0, order=4) -> Scene:
    rng = np.random.default_rng(seed)
    skies = rng.uniform(50.0, 500.0, n_bands)
    fields = [MixturePsfField(H, W, (1.0, 2.0, 4.0), (0.7, 0.25, 0.05), amp=0.1, rng=rng)
              for _ in range(n_bands)]
    psfs = [f.fit_polynomial(55, order=order) for f in fields]
    # Thomas process: parents uniform, 1 + Poisson children around them
    n_par = max(1, int(round(n_src * parent_frac * 1.15)))  # margin for edge losses
    px = rng.uniform(0, W - 1, n_par)
    py = rng.uniform(0, H - 1, n_par)
    kids = 1 + rng.poisson(1.0 / parent_frac - 1.0, n_par)
    xs, ys = [], []
    for i in range(n_par):
        xs.append(px[i] + rng.normal(0, sigma_c, kids[i]) * (np.arange(kids[i]) > 0))
        ys.append(py[i] + rng.normal(0, sigma_c, kids[i]) * (np.arange(kids[i]) > 0))
    tx = np.concatenate(xs)
    ty = np.concatenate(ys)
    ok = (tx >= 0) & (tx <= W - 1) & (ty >= 0) & (ty <= H - 1)
    tx, ty = tx[ok][:n_src], ty[ok][:n_src]
    n = tx.size
    if star_only:
        st = np.zeros(n, np.int32)
    else:
        u = rng.uniform(0, 1, n)
        st = np.where(u < 0.70, SOURCE_STAR, np.where(u < 0.85, SOURCE_EXP, np.where(
            u < 0.95, SOURCE_DEV, SOURCE_SERSIC))).astype(np.int32)
    sn = np.where(st == SOURCE_DEV, 4.0, np.where(st == SOURCE_SERSIC,
                                                   rng.uniform(0.5, 6.0, n), 1.0))
    sn = sn.astype(np.float32)
    base = np.exp(rng.uniform(np.log(1e2), np.log(1e6), n))
    flux = base[:, None] * rng.uniform(0.5, 2.0, (n, n_bands))
    lr, e1, e2 = _galaxy_shapes(rng, n)
    gal = st != SOURCE_STAR
    lr = np.where(gal, lr, 0.0).astype(np.float32)
    e1 = np.where(gal, e1, 0.0).astype(np.float32)
    e2 = np.where(gal, e2, 0.0).astype(np.float32)
    truths = [render_truth(H, W, fields[b], tx, ty, st, lr, e1, e2, sn, flux[:, b])
              for b in range(n_bands)]
    data, invvar, good = _finish(rng, truths, skies)
    x, y, lr_s, e1_s, e2_s = _seed_catalog(rng, tx, ty, st, lr, e1, e2)
    return Scene(data, invvar, good, psfs, x, y, st, lr_s, e1_s, e2_s, sn, tx, ty, flux,
                 meta=dict(config="c2", seed=seed, sky=list(skies), grid=(H // 256, W // 256),
                           halo=30))
