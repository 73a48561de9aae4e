"""ctypes binding of ``libuniphot_amd.so`` (the C ABI in include/uniphot_amd.h).

The library is built in-tree by ``python -m uniphot_amd.build``.  Loading
fails loudly when it is missing: there is no CPU fallback on the product
path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libuniphot_amd.so"

UPH_OK, UPH_ERR_INVALID_ARGUMENT, UPH_ERR_OOM, UPH_ERR_INTERNAL = 0, 1, 2, 3
UPH_PRECISION_FP32, UPH_PRECISION_FP64 = 0, 1

# Every exported entry point declared in include/uniphot_amd.h.
EXPORTS = (
    "uph_version", "uph_last_error", "uph_island_scratch_per_block", "uph_island_fit",
    "uph_island_fit_batch", "uph_island_split_counters", "uph_island_fit_split",
    "uph_prerender_scratch_bytes", "uph_prerender_spectra_floats",
    "uph_prerender64_scratch_bytes", "uph_prerender64_spectra_doubles",
    "uph_render_scratch_bytes", "uph_render_stamps", "uph_band_scratch_bytes", "uph_band_solve",
    "uph_wing_images", "uph_blank_work_floats", "uph_blank_response", "uph_stamp_sum",
    "uph_gather_windows", "uph_subtract_bg_planes", "uph_band_solve_noalloc",
    "uph_band_psf_flags", "uph_render_direct_scratch_bytes", "uph_render_stamps_direct",
    "uph_confusion_work_floats", "uph_confusion", "uph_fcorr_load", "uph_fcorr_select_work_words",
    "uph_fcorr_select", "uph_fcorr_clean", "uph_fcorr_tile_stats", "uph_fcorr_mask",
    "uph_fcorr_nlag", "uph_fcorr_autocov", "uph_psf_cutouts", "uph_psf_centroids",
    "uph_psf_stamps", "uph_psf_polyfit", "uph_psf_chi2", "uph_psf_pixel_median",
)
UPH_BAND_NOALLOC_BYTES = 64 << 20


class UphError(RuntimeError):
    """A non-OK status from the C ABI; ``code`` is the status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"uniphot_amd status {code}: {msg}")
        self.code = code


class UphOOM(UphError):
    """UPH_ERR_OOM: the caller may halve the batch and retry."""


class LMOptions(C.Structure):
    _fields_ = [("reg", C.c_float), ("pos_step_max", C.c_float),
                ("shape_step_max", C.c_float), ("un_step_max", C.c_float),
                ("ftol_rel", C.c_float), ("step_tol", C.c_float),
                ("shape_step_tol", C.c_float), ("max_steps", C.c_int), ("resume", C.c_int)]


class IslandClass(C.Structure):
    _fields_ = [("n_isl", C.c_int), ("N", C.c_int), ("B", C.c_int), ("T", C.c_int),
                ("n_terms", C.c_int), ("n_bands", C.c_int), ("n_live_max", C.c_int),
                ("Hp", C.c_int), ("Wp", C.c_int), ("mixed", C.c_int)]


_P = C.c_void_p


class IslandBatch(C.Structure):
    _fields_ = [("n_isl", C.c_int), ("n_bands", C.c_int), ("n_terms", C.c_int),
                ("Hp", C.c_int), ("Wp", C.c_int), ("max_B", C.c_int), ("mixed", C.c_int),
                ("n_tiers", C.c_int), ("tier_T", C.c_int * 3), ("psf_coefs", _P * 3),
                ("shape", _P), ("mem_off", _P), ("flux_off", _P)]


class IslandInputs(C.Structure):
    _fields_ = [(n, _P) for n in (
        "dstack", "wstack", "psf_coefs", "psf_geo", "sersic", "ox", "oy", "x0b", "y0b",
        "stype", "n_frozen", "free_n", "mmask", "keep", "p0", "lower", "upper", "order")]


class IslandOutputs(C.Structure):
    _fields_ = [(n, _P) for n in (
        "p", "fluxes", "ferr", "chi2", "lam", "conv", "stall", "nst", "consec", "nrej",
        "n_good")]


class IslandSplit(C.Structure):
    _fields_ = [("n_items", C.c_int), ("item_isl", _P), ("item_band", _P), ("hg_off", _P),
                ("hg", _P), ("hg_doubles", C.c_size_t), ("p_try", _P), ("counters", _P),
                ("eval_events", _P), ("n_pre", C.c_int), ("n_pre_ext", C.c_int),
                ("pre_isl", _P), ("pre_slot", _P), ("pre_off", _P), ("pre_spec", _P),
                ("pre_of_slot", _P), ("pbuf", _P), ("pbuf_elems", C.c_size_t), ("fspec", _P),
                ("fspec_elems", C.c_size_t), ("gdt", _P), ("gtap", _P), ("post", C.c_int),
                ("post_off", _P), ("post_buf", _P), ("post_doubles", C.c_size_t),
                ("post_var", _P), ("post_dvar", _P), ("post_metrics", _P),
                ("gram_bf16", C.c_int), ("gram_unsettled", _P), ("big_n", C.c_int),
                ("big_ntp", C.c_int), ("big_slot", _P), ("big_isl", _P), ("big_part", _P),
                ("post_kappa", _P), ("post_hg", _P), ("post_hg_off", _P)]


_lib = None


def lib() -> C.CDLL:
    """Load (once) and return the library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m uniphot_amd.build` "
            "(the fit path has no CPU fallback)")
    L = C.CDLL(str(LIB_PATH))
    L.uph_version.restype = C.c_int
    L.uph_last_error.restype = C.c_char_p
    L.uph_island_scratch_per_block.restype = C.c_size_t
    L.uph_island_scratch_per_block.argtypes = [C.c_int] * 5
    L.uph_island_fit.restype = C.c_int
    L.uph_island_fit.argtypes = [C.POINTER(IslandClass), C.POINTER(IslandInputs),
                                 C.POINTER(IslandOutputs), C.POINTER(LMOptions), C.c_int,
                                 _P, C.c_size_t, _P, _P]
    L.uph_island_fit_batch.restype = C.c_int
    L.uph_island_fit_batch.argtypes = [C.POINTER(IslandBatch), C.POINTER(IslandInputs),
                                       C.POINTER(IslandOutputs), C.POINTER(LMOptions), C.c_int,
                                       _P, C.c_size_t, C.c_size_t, _P, _P]
    L.uph_island_split_counters.restype = C.c_size_t
    L.uph_island_split_counters.argtypes = [C.c_int]
    L.uph_prerender_scratch_bytes.restype = C.c_size_t
    L.uph_prerender_scratch_bytes.argtypes = [C.c_int]
    L.uph_prerender_spectra_floats.restype = C.c_size_t
    L.uph_prerender_spectra_floats.argtypes = [C.c_int]
    for name in ("uph_prerender64_scratch_bytes", "uph_prerender64_spectra_doubles"):
        getattr(L, name).restype = C.c_size_t
        getattr(L, name).argtypes = [C.c_int]
    L.uph_island_fit_split.restype = C.c_int
    L.uph_island_fit_split.argtypes = [C.POINTER(IslandBatch), C.POINTER(IslandInputs),
                                       C.POINTER(IslandOutputs), C.POINTER(LMOptions), C.c_int,
                                       C.POINTER(IslandSplit), _P, C.c_size_t, C.c_size_t, _P]
    L.uph_render_scratch_bytes.restype = C.c_size_t
    L.uph_render_scratch_bytes.argtypes = [C.c_int, C.c_int]
    L.uph_render_stamps.restype = C.c_int
    L.uph_render_stamps.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P,
                                    _P, _P, _P, _P, _P, C.c_int, _P, _P, C.c_size_t, _P]
    L.uph_band_scratch_bytes.restype = C.c_size_t
    L.uph_band_scratch_bytes.argtypes = [C.c_int, C.c_int, C.c_int]
    L.uph_band_solve.restype = C.c_int
    L.uph_band_solve.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int,
                                 C.c_int, C.c_int, _P, _P, _P, C.c_double, C.c_double, C.c_double, C.c_double,
                                 _P, _P, _P, _P, C.c_float, _P, _P, _P, _P, _P, C.c_size_t, _P]
    L.uph_band_solve_noalloc.restype = C.c_int
    L.uph_band_solve_noalloc.argtypes = [C.c_int] * 5 + [_P, _P] + [C.c_int] * 4 + [_P] * 3 + \
        [C.c_double] * 4 + [_P] * 4 + [C.c_float] + [_P] * 4 + [_P]
    L.uph_blank_work_floats.restype = C.c_size_t
    L.uph_blank_work_floats.argtypes = [C.c_int, C.c_int]
    L.uph_blank_response.restype = C.c_int
    L.uph_blank_response.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int, _P, _P,
                                     _P, _P, _P, _P, _P, _P, _P, _P, C.c_double, C.c_double,
                                     C.c_double, C.c_double, _P, _P, C.c_size_t, C.c_int, _P, _P]
    L.uph_wing_images.restype = C.c_int
    L.uph_wing_images.argtypes = [C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P, _P,
                                  C.c_int, C.c_int, _P, _P]
    L.uph_stamp_sum.restype = C.c_int
    L.uph_stamp_sum.argtypes = [C.c_int] * 7 + [_P] * 7 + [C.c_float, C.c_int, C.c_int, _P, _P]
    L.uph_subtract_bg_planes.restype = C.c_int
    L.uph_subtract_bg_planes.argtypes = [C.c_int] * 4 + [_P, _P, C.c_int] + [C.c_double] * 4 + \
        [C.c_int, C.c_int, _P, _P]
    L.uph_gather_windows.restype = C.c_int
    L.uph_gather_windows.argtypes = [C.c_int, C.c_int, C.c_int, _P, C.c_int, C.c_int, _P, _P,
                                     C.c_int, C.c_int, _P, _P, _P]
    L.uph_band_psf_flags.restype = C.c_int
    L.uph_band_psf_flags.argtypes = [C.c_int] * 4 + [_P] * 5 + [C.c_int] * 2 + [_P] * 5 + [_P]
    L.uph_render_direct_scratch_bytes.restype = C.c_size_t
    L.uph_render_direct_scratch_bytes.argtypes = [C.c_int, C.c_int]
    L.uph_render_stamps_direct.restype = C.c_int
    L.uph_render_stamps_direct.argtypes = [C.c_int] * 3 + [_P] * 12 + [_P, C.c_size_t, _P]
    L.uph_confusion_work_floats.restype = C.c_size_t
    L.uph_confusion_work_floats.argtypes = [C.c_int, C.c_int]
    L.uph_confusion.restype = C.c_int
    L.uph_confusion.argtypes = ([C.c_int, C.c_int64, C.c_int, C.c_int, _P, _P, C.c_int, C.c_int]
                                + [_P] * 10 + [C.c_double] * 4 + [_P] * 9
                                + [C.c_size_t, C.c_int, _P, _P])
    L.uph_fcorr_load.restype = C.c_int
    L.uph_fcorr_load.argtypes = [C.c_int, C.c_int64] + [_P] * 11
    L.uph_fcorr_select_work_words.restype = C.c_size_t
    L.uph_fcorr_select_work_words.argtypes = [C.c_int]
    L.uph_fcorr_select.restype = C.c_int
    L.uph_fcorr_select.argtypes = [C.c_int, C.c_int64] + [_P] * 8 + [C.c_size_t, _P, _P]
    L.uph_fcorr_clean.restype = C.c_int
    L.uph_fcorr_clean.argtypes = [C.c_int, C.c_int64] + [_P] * 6 + [C.c_double, C.c_double,
                                                                    C.c_int, _P, _P]
    L.uph_fcorr_tile_stats.restype = C.c_int
    L.uph_fcorr_tile_stats.argtypes = [C.c_int, _P, _P, C.c_int] + [_P] * 6
    L.uph_fcorr_mask.restype = C.c_int
    L.uph_fcorr_mask.argtypes = [C.c_int, C.c_int64, _P, _P, _P, C.c_int] + [_P] * 5
    L.uph_fcorr_nlag.restype = C.c_int
    L.uph_fcorr_nlag.argtypes = [C.c_int]
    L.uph_fcorr_autocov.restype = C.c_int
    L.uph_fcorr_autocov.argtypes = [C.c_int] * 3 + [_P] * 5 + [C.c_size_t, _P, _P]
    L.uph_psf_cutouts.argtypes = [C.c_int] * 3 + [_P] * 4 + [C.c_int] + [C.c_double] * 3 + \
        [C.c_int] * 2 + [_P] * 4
    L.uph_psf_centroids.argtypes = [C.c_int, C.c_int, C.c_double, C.c_int, _P, _P, _P]
    L.uph_psf_stamps.argtypes = [C.c_int] * 3 + [_P] * 7 + [C.c_int] * 3 + [C.c_double] * 3 + \
        [C.c_int] * 2 + [_P] * 3
    L.uph_psf_polyfit.argtypes = [C.c_int] * 3 + [_P] * 5
    L.uph_psf_chi2.argtypes = [C.c_int] * 3 + [_P] * 5
    L.uph_psf_pixel_median.argtypes = [C.c_int] * 2 + [_P] * 4
    for f in ("uph_psf_cutouts", "uph_psf_centroids", "uph_psf_stamps", "uph_psf_polyfit",
              "uph_psf_chi2", "uph_psf_pixel_median"):
        getattr(L, f).restype = C.c_int
    _lib = L
    return L


def check(status: int) -> None:
    if status == UPH_OK:
        return
    msg = lib().uph_last_error().decode(errors="replace")
    if status == UPH_ERR_OOM:
        raise UphOOM(status, msg)
    raise UphError(status, msg)


def ptr(t) -> int | None:
    """Device pointer of a torch tensor (None for None)."""
    if t is None:
        return None
    return int(t.data_ptr())


def stream_ptr(stream=None) -> int | None:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)
