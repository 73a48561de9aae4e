"""Field-tile requests for the serving workload: square tiles of side 256 -
4096 px with 1 - 40 bands, sources at 1e-3 per px^2 with the 20-band field
scene's mix, clustering, fluxes and shapes, rendered on the GPU.

The PSF of band b is one spatially varying 3-Gaussian mixture in normalised
field coordinates (X, Y) in [-1, 1]^2, fitted once to a 15-term polynomial
(``PsfBank``); a tile sees it in its own frame, so a server keeps the
coefficients resident and only the frame (origin, scale) changes per tile.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass

import numpy as np

from uniphot_amd.psf_polynomial import PsfPolynomial
from uniphot_amd.source_model import SOURCE_DEV, SOURCE_EXP, SOURCE_SERSIC, SOURCE_STAR
from .field import SEEDS, MixturePsfField, _cov_terms, _galaxy_shapes, _seed_catalog
from .cutouts import galaxy_mixtures, render_profiles


class PsfBank:
    """Per-band PSF fields and their polynomial fits, frame-independent."""

    def __init__(self, seed=SEEDS["c5"], n_bands=40, T=55, order=4):
        rng = np.random.default_rng(seed)
        self.fields = [MixturePsfField(2048, 2048, (1.0, 2.0, 4.0), (0.7, 0.25, 0.05), amp=0.1,
                                       rng=rng) for _ in range(n_bands)]
        self.polys = [f.fit_polynomial(T, order=order) for f in self.fields]
        self.n_bands = n_bands

    def field(self, b, H, W) -> MixturePsfField:
        f = copy.copy(self.fields[b])
        f.H, f.W = H, W
        f.xo, f.xs = 0.5 * (W - 1), max(0.5 * W, 1.0)
        f.yo, f.ys = 0.5 * (H - 1), max(0.5 * H, 1.0)
        return f

    def psf(self, b, H, W) -> PsfPolynomial:
        p = self.polys[b]
        return PsfPolynomial(p.coefs, 0.5 * (W - 1), max(0.5 * W, 1.0), 0.5 * (H - 1),
                             max(0.5 * H, 1.0), p.H_psf, p.W_psf)


@dataclass
class TileRequest:
    rid: int
    arrival: float            # seconds after the stream starts
    side: int
    n_bands: int
    data: object              # torch (nb, side, side) float32 on the device
    invvar: object
    good: object              # torch uint8
    psfs: list                # PsfPolynomial per band, in the tile frame
    x: np.ndarray
    y: np.ndarray
    stype: np.ndarray
    logre: np.ndarray
    ee1: np.ndarray
    ee2: np.ndarray
    sersic_n: np.ndarray
    true_flux: np.ndarray

    @property
    def n_sources(self) -> int:
        return int(self.x.size)


def render_tile(H, W, fields, tx, ty, st, lr, e1, e2, sn, flux, device, box=27, chunk=400):
    """(nb, H, W) float64 truth: every source's (2 box + 1)^2 profile added
    into the image (overlaps accumulate)."""
    import torch
    nb = len(fields)
    f64 = dict(dtype=torch.float64, device=device)
    R = box
    S = 2 * R + 1
    img = torch.zeros((nb, H + 2 * S, W + 2 * S), **f64)
    n = tx.size
    if n == 0:
        return img[:, S:S + H, S:S + W]
    sig = np.stack([f.params(tx, ty)[0] for f in fields], axis=1)
    wgt = np.stack([f.params(tx, ty)[1] for f in fields], axis=1)
    cov = np.stack(_cov_terms(lr.astype(np.float64), e1.astype(np.float64),
                              e2.astype(np.float64)), axis=1)
    ga, gv, gcore = galaxy_mixtures(st, sn)
    ix = np.rint(tx).astype(np.int64) - R
    iy = np.rint(ty).astype(np.int64) - R
    off = torch.arange(S, device=device)
    Wp = W + 2 * S
    flat = img.view(nb, -1)
    for c0 in range(0, n, chunk):
        m = np.arange(c0, min(n, c0 + chunk))
        lx = torch.tensor(tx[m] - ix[m], **f64)
        ly = torch.tensor(ty[m] - iy[m], **f64)
        amp = torch.tensor(flux[m][:, :, None] * wgt[m], **f64)
        prof = render_profiles(lx, ly, S, torch.tensor(sig[m], **f64), amp,
                               torch.tensor(st[m] == SOURCE_STAR, device=device),
                               torch.tensor(cov[m], **f64), torch.tensor(ga[m], **f64),
                               torch.tensor(gv[m], **f64), torch.tensor(gcore[m], **f64))
        rows = torch.as_tensor(iy[m] + S, device=device)[:, None] + off[None, :]
        cols = torch.as_tensor(ix[m] + S, device=device)[:, None] + off[None, :]
        idx = (rows[:, :, None] * Wp + cols[:, None, :]).reshape(-1)
        flat.index_add_(1, idx, prof.permute(1, 0, 2, 3).reshape(nb, -1))
    return img[:, S:S + H, S:S + W]


def make_tile(rng, rid, arrival, side, n_bands, bank: PsfBank, device="cuda", density=1e-3,
              parent_frac=0.3, sigma_c=6.0, read_noise=5.0) -> TileRequest:
    import torch
    H = W = int(side)
    n_target = int(rng.poisson(density * H * W))
    n_par = max(1, int(round(n_target * parent_frac * 1.15)))
    px = rng.uniform(0, W - 1, n_par)
    py = rng.uniform(0, H - 1, n_par)
    kids = 1 + rng.poisson(1.0 / parent_frac - 1.0, n_par)
    first = np.repeat(np.cumsum(kids) - kids, kids)
    is_par = np.arange(kids.sum()) == first
    tx = np.repeat(px, kids) + rng.normal(0, sigma_c, kids.sum()) * ~is_par
    ty = np.repeat(py, kids) + rng.normal(0, sigma_c, kids.sum()) * ~is_par
    ok = (tx >= 0) & (tx <= W - 1) & (ty >= 0) & (ty <= H - 1)
    tx, ty = tx[ok][:n_target], ty[ok][:n_target]
    n = tx.size
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
    bands = rng.choice(bank.n_bands, n_bands, replace=False)
    fields = [bank.field(int(b), H, W) for b in bands]
    truth = render_tile(H, W, fields, tx, ty, st, lr, e1, e2, sn, flux, device)
    skies = torch.tensor(rng.uniform(50.0, 500.0, n_bands), dtype=torch.float64, device=device)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(rng.integers(1 << 62)))
    lam = torch.clamp(truth + skies[:, None, None], min=0.0)
    d = torch.poisson(lam, generator=gen)
    d += torch.randn(d.shape, generator=gen, dtype=torch.float64, device=device) * read_noise
    d -= skies[:, None, None]
    iv = (1.0 / (skies + read_noise ** 2))[:, None, None].expand(n_bands, H, W)
    x, y, lr_s, e1_s, e2_s = _seed_catalog(rng, tx, ty, st, lr, e1, e2)
    return TileRequest(rid, float(arrival), H, n_bands, d.float().contiguous(),
                       iv.float().contiguous(), torch.ones((n_bands, H, W), dtype=torch.uint8,
                                                           device=device),
                       [bank.psf(int(b), H, W) for b in bands], x, y, st, lr_s, e1_s, e2_s, sn,
                       flux)


def tile_stream(seed=SEEDS["c5"], n_req=16, rate=50.0, sides=(256, 512, 1024, 2048, 4096),
                max_bands=40, bank: PsfBank | None = None, device="cuda"):
    """Requests with Poisson arrivals (``rate`` per second), side uniform
    in ``sides`` and band count uniform in 1..max_bands."""
    rng = np.random.default_rng(seed)
    bank = bank or PsfBank(seed, n_bands=max_bands)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    out = []
    for i in range(n_req):
        side = int(rng.choice(sides))
        nb = int(rng.integers(1, max_bands + 1))
        out.append(make_tile(rng, i, t[i], side, nb, bank, device))
    return out, bank
