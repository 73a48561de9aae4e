"""Source types and morphology parameterisation (host side).

Follows ``tract7dt/fit/grid_gpu/source_model.py``: type codes (:24-27),
the Sersic-n logit bijection (:38-54), EllipseESoft conversions
(:104-138) and the catalog-to-parameter rules of ``catalog_source_params``
(:141-215), and the stamp renderers ``render_source_stamps`` /
``render_star_stamps`` / ``render_source_stamp_at`` (:264-390) over the
device kernels.
"""
from __future__ import annotations

import numpy as np

SOURCE_STAR, SOURCE_EXP, SOURCE_DEV, SOURCE_SERSIC = 0, 1, 2, 3
DEFAULT_N_DEV = 4.0
DEFAULT_N_SERSIC = 2.0
SERSIC_N_MIN = 0.29
SERSIC_N_MAX = 6.3

_TYPE_CODES = {"": SOURCE_STAR, "STAR": SOURCE_STAR, "POINTSOURCE": SOURCE_STAR,
               "POINT": SOURCE_STAR, "PSF": SOURCE_STAR, "EXP": SOURCE_EXP,
               "DEV": SOURCE_DEV, "SERSIC": SOURCE_SERSIC, "SER": SOURCE_SERSIC}


def sersic_n_to_u(n):
    t = (np.clip(n, SERSIC_N_MIN + 1e-4, SERSIC_N_MAX - 1e-4) - SERSIC_N_MIN) / (
        SERSIC_N_MAX - SERSIC_N_MIN)
    return np.log(t / (1.0 - t))


def sersic_u_to_n(u):
    return SERSIC_N_MIN + (SERSIC_N_MAX - SERSIC_N_MIN) / (1.0 + np.exp(-u))


def ellipse_soft_from_re_ab_phi_np(re_pix, ab, phi_deg):
    """(Re px, axis ratio, phi deg) -> (logre, ee1, ee2) float32."""
    re = np.clip(np.asarray(re_pix, np.float64), 0.1, 1000.0)
    ba = np.clip(np.asarray(ab, np.float64), 0.1, 1.0)
    inv_ba = 1.0 / ba
    e = np.clip((inv_ba - 1.0) / (inv_ba + 1.0), 0.0, 0.999999)
    ee = -np.log(np.maximum(1.0 - e, 1e-12))
    ang = np.deg2rad(-2.0 * np.asarray(phi_deg, np.float64))
    return (np.log(re).astype(np.float32), (ee * np.cos(ang)).astype(np.float32),
            (ee * np.sin(ang)).astype(np.float32))


def ellipse_soft_to_re_ab_phi_np(logre, ee1, ee2):
    re = np.exp(np.asarray(logre, np.float64))
    ee = np.hypot(np.asarray(ee1, np.float64), np.asarray(ee2, np.float64))
    e = 1.0 - np.exp(-ee)
    ab = (1.0 - e) / np.maximum(1.0 + e, 1e-12)
    phi = np.mod(-np.rad2deg(np.arctan2(ee2, ee1) / 2.0), 180.0)
    return re.astype(np.float32), ab.astype(np.float32), phi.astype(np.float32)


def type_codes(type_strings) -> np.ndarray:
    return np.array([_TYPE_CODES.get(str(t).strip().upper(), SOURCE_STAR)
                     for t in type_strings], dtype=np.int32)


def _catalog_columns(catalog):
    """TYPE, RE, ELL, THETA, SERSIC_N of a DataFrame-like catalogue, the
    column names matched without case (None where absent)."""
    names = {str(c).lower(): c for c in catalog.columns}

    def col(k, numeric=True):
        c = names.get(k)
        if c is None:
            return None
        v = np.asarray(catalog[c])
        if not numeric:
            return v.astype(str)
        import pandas as pd
        return pd.to_numeric(catalog[c], errors="coerce").to_numpy(dtype=float)
    typ = col("type", numeric=False)
    return (np.full(len(catalog), "STAR", dtype=object) if typ is None else typ,
            col("re"), col("ell"), col("theta"), col("sersic_n"))


def catalog_source_params(types, re=None, ell=None, theta=None, sersic_n=None,
                          re_fallback_pix: float = 3.0):
    """Catalog morphology -> (source_type, logre, ee1, ee2, sersic_n)
    (source_model.py:141-215).  ``types`` is a DataFrame-like catalogue
    (columns TYPE, RE, ELL, THETA, SERSIC_N in any case) or an array of
    TYPE strings / type codes with the other columns as arrays.  Re NaN ->
    ``re_fallback_pix``, clipped to [0.3, 100]; ab = 1 - ELL (NaN -> 0)
    clipped to [0.1, 1], 0.8 without an ELL column; phi from THETA (NaN ->
    0); n = 4 for DEV and 2 for SERSIC, a finite SERSIC_N used for SERSIC
    rows only, clipped to [0.29, 6.3]."""
    if hasattr(types, "columns"):
        types, re, ell, theta, sersic_n = _catalog_columns(types)
    st = type_codes(types) if not np.issubdtype(np.asarray(types).dtype, np.integer) \
        else np.asarray(types, np.int32)
    n = st.size
    sn = np.full(n, 1.0)
    sn[st == SOURCE_DEV] = DEFAULT_N_DEV
    sn[st == SOURCE_SERSIC] = DEFAULT_N_SERSIC
    if sersic_n is not None:
        raw = np.asarray(sersic_n, np.float64)
        use = np.isfinite(raw) & (st == SOURCE_SERSIC)
        sn[use] = raw[use]
    sn = np.clip(sn, SERSIC_N_MIN, SERSIC_N_MAX)
    r = np.full(n, float(re_fallback_pix)) if re is None else np.where(
        np.isfinite(re), re, float(re_fallback_pix))
    r = np.clip(r, 0.3, 100.0)
    ab = np.full(n, 0.8) if ell is None else np.clip(1.0 - np.nan_to_num(ell, nan=0.0), 0.1, 1.0)
    if theta is None:
        phi = np.zeros(n)
    else:
        t = np.nan_to_num(np.asarray(theta, np.float64), nan=0.0)
        phi = -np.where(t >= 0, 90.0 - t, -90.0 - t)
    lr, e1, e2 = ellipse_soft_from_re_ab_phi_np(r, ab, phi)
    return st, lr, e1, e2, sn.astype(np.float32)


def _render(psf, pos_x, pos_y, grid_x0, grid_y0, source_type, logre, ee1, ee2, sersic_n,
            render_backend, stream=None):
    from .forced_phot_sparse import render_unit_stamps
    from .psf_polynomial import psf_device_arrays
    import torch
    st = np.asarray(source_type, np.int32)
    if render_backend not in ("direct", "fft"):
        raise ValueError(f"unknown render_backend: {render_backend!r}")
    if render_backend == "direct" and np.any((st == SOURCE_DEV) | (st == SOURCE_SERSIC)):
        raise ValueError("render_backend='direct' cannot render DEV/SERSIC sources; use 'fft'")
    T = int(psf.H_psf)
    c, g = psf_device_arrays([psf], T)
    n = np.asarray(pos_x).size
    dev = torch.device("cuda")
    if render_backend == "direct":
        return render_direct_stamps(torch.from_numpy(c).to(dev), torch.from_numpy(g).to(dev), T,
                                    c.shape[1], np.zeros(n, np.int32), pos_x, pos_y,
                                    np.full(n, grid_x0), np.full(n, grid_y0), st, logre, ee1,
                                    ee2, stream=stream)
    return render_unit_stamps(torch.from_numpy(c).to(dev), torch.from_numpy(g).to(dev), T,
                              c.shape[1], np.zeros(n, np.int32), pos_x, pos_y,
                              np.full(n, grid_x0), np.full(n, grid_y0), st,
                              logre, ee1, ee2,
                              np.ones(n) if sersic_n is None else sersic_n, stream=stream)


def render_direct_stamps(psf_coefs_dev, psf_geo_dev, T: int, n_terms: int, band, pos_x, pos_y,
                         origin_x, origin_y, stype, logre, ee1, ee2, *, stream=None):
    """(n, T, T) unit-flux stamps of the direct backend on the device
    (``uph_render_stamps_direct``): the PSF at the position for stars; for
    EXP the pixel-centre 6-Gaussian profile convolved directly with the PSF
    stamp and renormalised to unit sum (source_model.py:218-256, :298-303);
    Lanczos-3 shift by the remainder of (pos - origin)."""
    import torch
    from . import _lib
    dev = psf_coefs_dev.device
    L = _lib.lib()
    P = _lib.ptr
    n = int(np.asarray(pos_x).size)
    out = torch.empty((max(n, 1), T, T), dtype=torch.float32, device=dev)
    if n == 0:
        return out[:0]
    st = np.asarray(stype, np.int32)
    if np.any((st == SOURCE_DEV) | (st == SOURCE_SERSIC)):
        raise ValueError("render_backend='direct' cannot render DEV/SERSIC sources; use 'fft'")
    up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)
    k = max(1, min(n, 1024))
    sb = L.uph_render_direct_scratch_bytes(T, k)
    scratch = torch.empty(max(sb, 1), dtype=torch.uint8, device=dev) if sb else None
    args = [up(band, np.int32), up(pos_x, np.float32), up(pos_y, np.float32),
            up(origin_x, np.float32), up(origin_y, np.float32), up(st, np.int32),
            up(logre, np.float32), up(ee1, np.float32), up(ee2, np.float32)]
    _lib.check(L.uph_render_stamps_direct(n, T, n_terms, P(psf_coefs_dev), P(psf_geo_dev),
                                          *(P(a) for a in args), P(out), P(scratch), sb,
                                          _lib.stream_ptr(stream)))
    return out


def render_source_stamps(psf_jax, pos_x, pos_y, grid_x0: int, grid_y0: int, source_type,
                         logre, ee1, ee2, sersic_n=None, render_backend: str = "direct"):
    """Unit-flux stamps (n, T, T) on the device for mixed star/galaxy
    sources (source_model.py:367-390): the PSF polynomial at the full-image
    position, EXP/DEV/SERSIC through the Fourier mixture convolution
    (``render_backend="fft"``, not renormalised) or EXP through the direct
    pixel-space convolution (``render_backend="direct"``, unit box sum),
    Lanczos-3 shift by the remainder of (pos - grid origin).  ``psf_jax`` is
    a PsfPolynomial."""
    return _render(psf_jax, pos_x, pos_y, grid_x0, grid_y0, source_type, logre, ee1, ee2,
                   sersic_n, render_backend)


def render_star_stamps(psf_jax, pos_x, pos_y, grid_x0: int, grid_y0: int):
    """Unit-flux stamps of an all-star source set (source_model.py:331-364),
    equal to the STAR rows of :func:`render_source_stamps`."""
    n = np.asarray(pos_x).size
    z = np.zeros(n, np.float32)
    return _render(psf_jax, pos_x, pos_y, grid_x0, grid_y0, np.zeros(n, np.int32), z, z, z,
                   None, "fft")


def render_source_stamp_at(psf_jax, x, y, grid_x0: int, grid_y0: int, source_type, logre, ee1,
                           ee2, sersic_n=None, render_backend: str = "direct"):
    """One source's unit-flux stamp (T, T) (source_model.py:264-328)."""
    return render_source_stamps(psf_jax, np.array([x]), np.array([y]), grid_x0, grid_y0,
                                np.array([source_type]), np.array([logre]), np.array([ee1]),
                                np.array([ee2]), None if sersic_n is None else
                                np.array([sersic_n]), render_backend)[0]
