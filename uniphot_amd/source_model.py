"""Source types and morphology parameterisation (host side).

Follows ``tract7dt/fit/grid_gpu/source_model.py``: type codes (:24-27),
the Sersic-n logit bijection (:38-54), EllipseESoft conversions
(:104-138) and the catalog-to-parameter rules of ``catalog_source_params``
(:141-215), here taking arrays instead of a DataFrame.
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


def catalog_source_params(types, re=None, ell=None, theta=None, sersic_n=None,
                          re_fallback_pix: float = 3.0):
    """Catalog morphology -> (source_type, logre, ee1, ee2, sersic_n).
    Re clipped to [0.3, 100] (fallback 3 px), ab = 1 - ELL in [0.1, 1]
    (default 0.8), phi from THETA, DEV n = 4, SERSIC n = 2 unless given."""
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
