"""PSFEx-style polynomial PSF model (host side) and its device layout.

Mirrors ``tract7dt/fit/grid_gpu/psf_polynomial.py:55-279`` (the
``PsfPolynomial`` dataclass, term order, numpy evaluation) and the island
PSF tree of ``island_solver.py:1851-1871`` (zero-padded to a stamp tier,
stacked over bands).  The device layout is planar, (n_bands, n_terms, T, T)
float32, so that one wavefront reads one term plane contiguously.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass

import numpy as np


def terms_for_order(order: int) -> tuple:
    """(m, n) exponent pairs of X^m Y^n: low total degree first, within a
    degree decreasing power of X (bg_polynomial.py:40-100)."""
    if not 0 <= order <= 5:
        raise ValueError(f"polynomial_order must be 0..5; got {order}")
    return tuple((m, d - m) for d in range(order + 1) for m in range(d, -1, -1))


def n_terms_for_order(order: int) -> int:
    return len(terms_for_order(order))


ORDER_FOR_N_TERMS = {1: 0, 3: 1, 6: 2, 10: 3, 15: 4, 21: 5}


@dataclass
class PsfPolynomial:
    """Per-band PSF: coefs (H_psf, W_psf, n_terms); X = (x - x_origin) /
    x_scale, Y likewise (psf_polynomial.py:55-95)."""
    coefs: np.ndarray
    x_origin: float
    x_scale: float
    y_origin: float
    y_scale: float
    H_psf: int
    W_psf: int

    @property
    def n_terms(self) -> int:
        return int(self.coefs.shape[-1])


def evaluate_psf_at_np(psf: PsfPolynomial, x: float, y: float) -> np.ndarray:
    """float32 stamp at (x, y) (psf_polynomial.py:200-226)."""
    X = (float(x) - psf.x_origin) / psf.x_scale
    Y = (float(y) - psf.y_origin) / psf.y_scale
    terms = terms_for_order(ORDER_FOR_N_TERMS[psf.n_terms])
    basis = np.array([X ** m * Y ** n for (m, n) in terms], dtype=np.float32)
    return np.tensordot(psf.coefs.astype(np.float32), basis, axes=([-1], [-1])).astype(np.float32)


def evaluate_psf_at_batch_np(psf: PsfPolynomial, xs, ys) -> np.ndarray:
    """float32 stamps (n, H, W) at n positions in one product: basis in
    float64 rounded to float32, coefficients (H W, n_t) times basis (n_t, n)
    (psf_polynomial.py:236-270)."""
    X = (np.asarray(xs, np.float64) - psf.x_origin) / psf.x_scale
    Y = (np.asarray(ys, np.float64) - psf.y_origin) / psf.y_scale
    terms = terms_for_order(ORDER_FOR_N_TERMS[psf.n_terms])
    basis = np.stack([(X ** m * Y ** n).astype(np.float32) for (m, n) in terms])
    H, W = psf.coefs.shape[:2]
    flat = psf.coefs.reshape(H * W, psf.n_terms).astype(np.float32, copy=False)
    return (flat @ basis).T.reshape(-1, H, W).astype(np.float32, copy=False)


def pad_psf(psf: PsfPolynomial, tier: int) -> PsfPolynomial:
    """Zero-pad to a larger odd tier keeping the centre pixel at tier//2
    (island_solver.py:193-206, runner.py:327-350)."""
    H, W = int(psf.H_psf), int(psf.W_psf)
    if tier <= H and tier <= W:
        return psf
    c = np.zeros((tier, tier) + psf.coefs.shape[2:], dtype=psf.coefs.dtype)
    py, px = (tier - H) // 2, (tier - W) // 2
    c[py:py + H, px:px + W] = psf.coefs
    return dataclasses.replace(psf, coefs=c, H_psf=tier, W_psf=tier)


def psf_device_arrays(psfs, tier: int):
    """(coefs (nb, nt, T, T) float32 planar, geo (nb, 4) float32) host arrays
    for the kernels, padded to ``tier``."""
    nts = {p.n_terms for p in psfs}
    if len(nts) != 1:
        raise ValueError(f"bands disagree on n_terms: {sorted(nts)}")
    out = np.zeros((len(psfs), nts.pop(), tier, tier), dtype=np.float32)
    geo = np.zeros((len(psfs), 4), dtype=np.float32)
    for b, p in enumerate(psfs):
        q = pad_psf(p, tier)
        if q.H_psf != tier or q.W_psf != tier:
            raise ValueError(f"PSF stamp {q.H_psf}x{q.W_psf} larger than tier {tier}")
        out[b] = np.moveaxis(q.coefs.astype(np.float32), -1, 0)
        geo[b] = (q.x_origin, q.x_scale, q.y_origin, q.y_scale)
    return out, geo
