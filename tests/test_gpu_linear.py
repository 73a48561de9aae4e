"""GPU parity of the stamp renderer and the grid-band linear solve against
the float64 oracle (oracle/linear.py)."""
import numpy as np
import pytest

from _scenes import small_star_scene
from oracle import core as ocore
from oracle import linear as olin

pytestmark = pytest.mark.gpu


def _psf_dict(p):
    return dict(coefs=p.coefs.astype(np.float64), x_origin=p.x_origin, x_scale=p.x_scale,
                y_origin=p.y_origin, y_scale=p.y_scale)


def test_render_stamps_match_oracle(cuda):
    import torch
    from uniphot_amd import _lib
    from uniphot_amd.psf_polynomial import psf_device_arrays
    sc = small_star_scene()
    c, g = psf_device_arrays(sc.psfs, 55)
    dev = cuda
    n = sc.x.size
    band = (np.arange(n) % 2).astype(np.int32)
    gx0 = np.full(n, 17.0, np.float32)
    gy0 = np.full(n, 3.0, np.float32)
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    out = torch.empty((n, 55, 55), dtype=torch.float32, device=dev)
    L = _lib.lib()
    P = _lib.ptr
    tc, tg, tb, tx, ty, tgx, tgy = map(up, (c, g, band, sc.x, sc.y, gx0, gy0))
    _lib.check(L.uph_render_stamps(n, 55, c.shape[1], P(tc), P(tg), None, P(tb), P(tx), P(ty),
                                   P(tgx), P(tgy), None, None, None, None, None, 1, P(out),
                                   None, 0, _lib.stream_ptr()))
    got = out.cpu().numpy()
    for i in range(n):
        ref, _ = ocore.member_stamp(_psf_dict(sc.psfs[band[i]]), float(sc.x[i]), float(sc.y[i]),
                                    17.0, 3.0, 0, renorm=True)
        assert np.max(np.abs(got[i] - ref)) < 2e-7 * np.max(np.abs(ref)) * 10


def test_band_solve_matches_oracle(cuda):
    import torch
    from uniphot_amd import forced_phot_sparse as fps
    from uniphot_amd.psf_polynomial import psf_device_arrays
    sc = small_star_scene(n_src=20)
    nb, H, W = sc.data.shape
    dev = cuda
    c, g = psf_device_arrays(sc.psfs, 55)
    # one grid = a sub-window of the image with all sources whose centre is inside
    x0, y0, h, w = 10, 5, 140, 150
    inside = (sc.x >= x0) & (sc.x < x0 + w) & (sc.y >= y0) & (sc.y < y0 + h)
    xs, ys = sc.x[inside], sc.y[inside]
    nbr, msk = fps.build_neighbor_list(xs, ys, 55, 55)
    probs = [fps.BandProblem(b, x0, y0, h, w, xs, ys, nbr) for b in range(nb)]
    bg_geo = (0.5 * (W - 1), 0.5 * W, 0.5 * (H - 1), 0.5 * H)
    up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    res = fps.solve_problems(probs, up(c), up(g), 55, c.shape[1], up(sc.data),
                             up(np.where(sc.good, sc.invvar, 0).astype(np.float32)),
                             up(sc.good.astype(np.uint8)), bg_order=1, bg_geo=bg_geo)
    for b in range(nb):
        psf = _psf_dict(sc.psfs[b])
        stamps = np.stack([ocore.member_stamp(psf, float(xx), float(yy), x0, y0, 0, renorm=True)[0]
                           for xx, yy in zip(xs, ys)])
        ix, iy = fps.stamp_top_left(xs, ys, x0, y0, 55, 55)
        basis = olin.bg_basis(h, w, x0, y0, *bg_geo, order=1)
        f, bgc, fv, bv = olin.solve_band(sc.data[b, y0:y0 + h, x0:x0 + w].astype(np.float64),
                                         sc.invvar[b, y0:y0 + h, x0:x0 + w].astype(np.float64),
                                         sc.good[b, y0:y0 + h, x0:x0 + w], stamps, ix, iy, nbr,
                                         msk, basis)
        r = res[b]
        sig = np.sqrt(fv)
        assert np.max(np.abs(r.fluxes - f) / np.maximum(np.abs(f), sig)) < 1e-5
        assert np.allclose(r.flux_var, fv, rtol=1e-5)
        assert np.allclose(r.bg_coefs, bgc, rtol=1e-4, atol=1e-4 * np.sqrt(bv).max())
        assert r.converged
