"""Sersic-profile mixture-of-Gaussians table (host side) and its device block.

The table format and lookup follow ``tract7dt/fit/grid_gpu/sersic_mog.py``
(:48-100): knots in n with each segment-boundary knot duplicated, 15
Gaussian components (amplitude, log variance in Re^2 units) plus a delta
core per knot, ``searchsorted(side="right") - 1`` and linear interpolation
inside a segment.  Two families, spliced at n = 0.8 as in the reference
(``sersic_mog_gen.py:45-52``):
  * n in [0.29, 0.8]: The Tractor's fits, which exist only as data; they
    are read from ``data/sersic_tractor_rows.npz`` (see ``data/NOTICE``);
  * n in [0.8, 6.3]: generated here from the published Shajib (2019)
    decomposition of an untruncated Sersic profile (the method of
    ``sersic_mog_gen.py:61-111``) on the 0.005 knot grid; they equal the
    reference table's rows to ~3e-14.
The assembled table has the reference's 1209 knots.

Device block layout (float64, see uph_galaxy.hpp):
  [0] n_knots, [1] n_min, [2] n_max, [3] K, then knots (n_knots),
  amps (n_knots, K), log_vars (n_knots, K), core (n_knots).
"""
from __future__ import annotations

from functools import lru_cache
from pathlib import Path

import numpy as np
from scipy.special import comb, gammaln

K_MAX = 15
SPLICE_N, N_MAX = 0.8, 6.3
DN = 0.005
LOW_ROWS = Path(__file__).resolve().parent / "data" / "sersic_tractor_rows.npz"


@lru_cache(maxsize=None)
def _euler_nodes(P: int = 10):
    """Nodes and weights of the Euler-accelerated inverse Laplace sum
    (Abate & Whitt 2006) for a decomposition in Gaussians."""
    k = np.arange(2 * P + 1)
    nodes = np.sqrt(2.0 * P * np.log(10.0) / 3.0 + 2.0j * np.pi * k)
    xi = np.ones(2 * P + 1)
    xi[0] = 0.5
    xi[2 * P] = 0.5 ** P
    for j in range(P - 1, 0, -1):
        xi[P + j] = xi[P + j + 1] + comb(P, P - j) * 0.5 ** P
    weights = (-1.0) ** k * xi * 10.0 ** (P / 3.0) * 2.0 * np.sqrt(2.0 * np.pi)
    return weights, nodes


def _bn(n):
    n = np.asarray(n, np.float64)
    t1 = 2.0 * n - 1.0 / 3.0 + 4.0 / (405.0 * n) + 46.0 / (25515.0 * n ** 2)
    t2 = (0.00018073182821237496 / (n + np.exp(n)) + 3.7026973571904255e-5) / (
        (-0.09149183119702775 - n) ** 2)
    t3 = (np.log(n) + (n * 2.6248718397705195 - 0.9727511612512357)
          / (n ** -4 + 94.78011643586419)) / (n - 0.006044236674273689)
    return t1 + t2 * t3


def shajib_components(n: float, K: int = K_MAX, P: int = 10):
    """(amplitudes, variances in Re^2) of the K-Gaussian expansion of the
    unit-flux Sersic profile of index n; widest sigma ramps from 2.5 Re at
    n = 0.8 to 15 Re at n = 2."""
    wts, nodes = _euler_nodes(P)
    s_hi = min(max(2.5 + 12.5 * (n - 0.8) / 1.2, 2.5), 15.0)
    sig = np.logspace(-2.0, np.log10(s_hi), K)
    b = _bn(n)
    ie = b ** (2 * n) / (2.0 * np.pi * n * np.exp(b + gammaln(2 * n)))
    r = np.outer(sig, nodes)
    prof = ie * np.exp(-b * (r ** (1.0 / n) - 1.0))
    f = (wts * prof).real.sum(axis=1)
    step = np.abs(np.mean(np.diff(np.log(sig))))
    amp = f * step / np.sqrt(2.0 * np.pi)
    amp[[0, -1]] *= 0.5
    return amp * 2.0 * np.pi * sig ** 2, sig ** 2


@lru_cache(maxsize=1)
def table() -> dict:
    low = np.load(LOW_ROWS, allow_pickle=False)
    if float(low["n_knots"][-1]) != SPLICE_N or low["amps"].shape[1] != K_MAX:
        raise ValueError(f"{LOW_ROWS}: unexpected layout")
    m = int(round((N_MAX - SPLICE_N) / DN)) + 1
    kn = np.linspace(SPLICE_N, N_MAX, m)
    rows = [shajib_components(float(n)) for n in kn]
    amps = np.array([a for a, _ in rows])
    lvs = np.log(np.array([v for _, v in rows]))
    return dict(n_knots=np.concatenate([low["n_knots"], kn]),
                amps=np.concatenate([low["amps"], amps]),
                log_vars=np.concatenate([low["log_vars"], lvs]),
                core_amp=np.concatenate([low["core_amp"], np.zeros(m)]),
                n_min=float(low["n_min"]), n_max=N_MAX)


def mog_params_at(n: float):
    """n -> (amps (K,), vars (K,), core) in float64."""
    t = table()
    kn = t["n_knots"]
    nn = min(max(float(n), t["n_min"]), t["n_max"])
    i = min(max(int(np.searchsorted(kn, nn, side="right")) - 1, 0), kn.size - 2)
    n0, n1 = kn[i], kn[i + 1]
    w = (nn - n0) / (n1 - n0) if n1 > n0 else 0.0
    a = (1 - w) * t["amps"][i] + w * t["amps"][i + 1]
    lv = (1 - w) * t["log_vars"][i] + w * t["log_vars"][i + 1]
    c = (1 - w) * t["core_amp"][i] + w * t["core_amp"][i + 1]
    return a, np.exp(lv), float(c)


def device_block() -> np.ndarray:
    """Flat float64 block of the table for the kernels."""
    t = table()
    nk = t["n_knots"].size
    head = np.array([nk, t["n_min"], t["n_max"], K_MAX], np.float64)
    return np.concatenate([head, t["n_knots"], t["amps"].ravel(), t["log_vars"].ravel(),
                           t["core_amp"]]).astype(np.float64)


_DEVICE_BLOCKS = {}


def device_table(device="cuda"):
    """The float64 table block resident on ``device`` (uploaded once)."""
    import torch
    key = str(device)
    if key not in _DEVICE_BLOCKS:
        _DEVICE_BLOCKS[key] = torch.from_numpy(device_block()).to(device)
    return _DEVICE_BLOCKS[key]
