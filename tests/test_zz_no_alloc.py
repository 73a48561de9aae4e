"""The allocation-free band solve (uph_band_solve_noalloc) against the
allocating entry point (uph_band_solve) through the C ABI: two calls in a
row, the second with a smaller batch, each bit-identical to the allocating
call on the same inputs; an oversized batch is rejected before any launch."""
import numpy as np
import pytest

from _scenes import small_star_scene

pytestmark = pytest.mark.gpu


def _inputs(sc, n_src_keep, dev):
    """Device arrays of one grid-band problem per band over a window."""
    import torch
    from uniphot_amd import forced_phot_sparse as fps
    from uniphot_amd.psf_polynomial import psf_device_arrays
    nb, H, W = sc.data.shape
    x0, y0, h, w = 10, 5, 140, 150
    inside = np.flatnonzero((sc.x >= x0) & (sc.x < x0 + w) & (sc.y >= y0) & (sc.y < y0 + h))
    inside = inside[:n_src_keep]
    xs, ys = sc.x[inside], sc.y[inside]
    nbr, _ = fps.build_neighbor_list(xs, ys, 55, 55)
    c, g = psf_device_arrays(sc.psfs, 55)
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)   # noqa: E731
    n = xs.size
    stamps = fps.render_unit_stamps(up(c), up(g), 55, c.shape[1], np.repeat(np.arange(nb), n),
                                    np.tile(xs, nb), np.tile(ys, nb), np.full(nb * n, x0),
                                    np.full(nb * n, y0), np.zeros(nb * n), np.zeros(nb * n),
                                    np.zeros(nb * n), np.zeros(nb * n), np.ones(nb * n))
    ix, iy = fps.stamp_top_left(xs, ys, x0, y0, 55, 55)
    return dict(
        n_prob=nb, max_src=n, off=up(np.arange(nb + 1, dtype=np.int64) * n),
        win=up(np.array([[b, x0, y0, h, w] for b in range(nb)], np.int32)),
        H=H, W=W, data=up(sc.data.astype(np.float32)),
        invvar=up(np.where(sc.good, sc.invvar, 0).astype(np.float32)),
        good=up(sc.good.astype(np.uint8)), stamps=stamps,
        ix=up(np.tile(ix, nb).astype(np.int32)), iy=up(np.tile(iy, nb).astype(np.int32)),
        nbr=up(np.tile(nbr[:, :32], (nb, 1)).astype(np.int32)),
        geo=(0.5 * (W - 1), 0.5 * W, 0.5 * (H - 1), 0.5 * H))


def _outputs(a, dev):
    import torch
    n = a["n_prob"] * a["max_src"]
    f = dict(dtype=torch.float32, device=dev)
    return [torch.full((n,), float("nan"), **f), torch.full((n,), float("nan"), **f),
            torch.full((a["n_prob"], 3), float("nan"), **f),
            torch.full((a["n_prob"], 3), float("nan"), **f)]


def _solve(a, dev, no_alloc):
    import torch
    from uniphot_amd import _lib
    L, P = _lib.lib(), _lib.ptr
    outs = _outputs(a, dev)
    common = (P(a["off"]), P(a["win"]), a["H"], a["W"], 0, 0, P(a["data"]), P(a["invvar"]),
              P(a["good"]), *a["geo"], P(a["stamps"]), P(a["ix"]), P(a["iy"]), P(a["nbr"]),
              1e-9, *(P(o) for o in outs))
    if no_alloc:
        _lib.check(L.uph_band_solve_noalloc(a["n_prob"], 55, 1, 32, a["max_src"], *common,
                                            _lib.stream_ptr()))
    else:
        sb = L.uph_band_scratch_bytes(a["n_prob"], a["max_src"], 1)
        scr = torch.empty(sb, dtype=torch.uint8, device=dev)
        _lib.check(L.uph_band_solve(a["n_prob"], 55, 1, 32, *common, P(scr), sb,
                                    _lib.stream_ptr()))
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in outs]


def test_band_solve_noalloc_equals_the_allocating_call(cuda):
    from uniphot_amd import _lib
    sc = small_star_scene(n_src=24)
    for n_keep in (24, 9):       # the second call is smaller
        a = _inputs(sc, n_keep, cuda)
        print("no-alloc band solve:", a["n_prob"], "problems of", a["max_src"], "sources",
              flush=True)
        ref = _solve(a, cuda, no_alloc=False)
        got = _solve(a, cuda, no_alloc=True)
        for r, g_ in zip(ref, got):
            assert np.isfinite(r).all()
            assert np.array_equal(r, g_)
    # a batch beyond the static workspace is refused on the host
    L = _lib.lib()
    st = L.uph_band_solve_noalloc(4000, 55, 1, 32, 200, None, None, 10, 10, 0, 0, None, None,
                                  None, 0.0, 1.0, 0.0, 1.0, None, None, None, None, 1e-9, None,
                                  None, None, None, None)
    assert st == _lib.UPH_ERR_INVALID_ARGUMENT
    assert b"static" in L.uph_last_error()
