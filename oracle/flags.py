"""FP64 restatement of the band-flag PSF diagnostics (TEST INFRASTRUCTURE).

Reference: ``tract7dt/fit/grid_gpu/band_flags.py`` -- ``psf_lobe_zone``
(:115-137: the band's PSF model stamp at the position has no positive peak,
or its minimum is below -lobe_frac x peak; non-finite positions are never
flagged) and ``psf_uncovered_fraction`` (:140-196: the positive part of the
stamp, centred on the nearest pixel, summed over no-data pixels -- pixels
beyond the image count as no data -- over its total; 0 for a non-finite
position or a stamp touching no such pixel, 1 for a position off the image
or a stamp without positive part).  The stamp is ``evaluate_psf_at_batch_np``
(``psf_polynomial.py:229-270``) evaluated here in float64, per source, by
brute force over the stamp (no summed-area shortcut).

Parity pins (tests/test_band_flags.py), the reference's own known answers
(``tests/test_partial_coverage.py``): a fully covered interior is 0; a
constant Gaussian PSF on a footprint edge matches a direct window sum; a
source on the edge is about half uncovered; the image border counts as no
data; NaN gives 0 and an off-image position 1.  Lobe: a Gaussian has no
lobe; a ring of -0.3 % of the peak is a lobe and one of -0.1 % is not; a
non-positive stamp is one.
"""
from __future__ import annotations

import numpy as np

from .core import eval_psf


def psf_stamp(psf: dict, x: float, y: float) -> np.ndarray:
    return eval_psf(psf, float(x), float(y))


def lobe_zone(psf: dict, xs, ys, lobe_frac: float = 0.002) -> np.ndarray:
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    out = np.zeros(xs.shape, bool)
    for i, (x, y) in enumerate(zip(xs, ys)):
        if not (np.isfinite(x) and np.isfinite(y)):
            continue
        st = psf_stamp(psf, x, y)
        pk, mn = st.max(), st.min()
        out[i] = pk <= 0.0 or mn < -lobe_frac * pk
    return out


def uncovered_fraction(psf: dict, nodata: np.ndarray, xs, ys) -> np.ndarray:
    nodata = np.asarray(nodata, bool)
    H, W = nodata.shape
    S = int(np.asarray(psf["coefs"]).shape[0])
    h = S // 2
    xs = np.asarray(xs, np.float64)
    ys = np.asarray(ys, np.float64)
    out = np.zeros(xs.shape, np.float64)
    for i, (x, y) in enumerate(zip(xs, ys)):
        if not (np.isfinite(x) and np.isfinite(y)):
            continue
        ix, iy = int(np.rint(x)), int(np.rint(y))
        if not (0 <= ix < W and 0 <= iy < H):
            out[i] = 1.0
            continue
        miss = np.ones((S, S), bool)
        r0, r1 = max(iy - h, 0), min(iy - h + S, H)
        c0, c1 = max(ix - h, 0), min(ix - h + S, W)
        miss[r0 - (iy - h):r1 - (iy - h), c0 - (ix - h):c1 - (ix - h)] = nodata[r0:r1, c0:c1]
        if not miss.any():
            continue
        pos = np.clip(psf_stamp(psf, x, y), 0.0, None)
        tot = pos.sum()
        out[i] = float((pos * miss).sum() / tot) if tot > 0 else 1.0
    return out
