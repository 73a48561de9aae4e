"""Direct render backend (uph_render_stamps_direct) against the float64
oracle: STAR and EXP stamps at the LDS tier (55 px) and the scratch tier
(111 px), round and elongated profiles, and the facade's rules."""
import numpy as np
import pytest

from _scenes import small_mixed_scene
from oracle import core as ocore

pytestmark = pytest.mark.gpu


def _psf(p, T):
    return dict(coefs=ocore.pad_psf_coefs(p.coefs.astype(np.float64), T), x_origin=p.x_origin,
                x_scale=p.x_scale, y_origin=p.y_origin, y_scale=p.y_scale)


@pytest.mark.parametrize("T", [55, 111])
def test_direct_stamps_match_oracle(cuda, T):
    from uniphot_amd.psf_polynomial import pad_psf
    from uniphot_amd.source_model import render_source_stamps
    sc = small_mixed_scene(types=(0, 1, 1, 1))
    n = sc.x.size
    st = sc.stype.copy()
    st[st > 1] = 1                     # the direct backend renders STAR and EXP
    ee1, ee2 = sc.ee1.copy(), sc.ee2.copy()
    ee1[1], ee2[1] = 0.0, 0.0          # a round profile
    gx0, gy0 = 5.0, 9.0
    for b, p in enumerate(sc.psfs):
        got = render_source_stamps(pad_psf(p, T), sc.x, sc.y, gx0, gy0, st, sc.logre, ee1, ee2,
                                   render_backend="direct").cpu().numpy()
        for i in range(n):
            ref = ocore.direct_stamp(_psf(p, T), float(sc.x[i]), float(sc.y[i]), gx0, gy0,
                                     int(st[i]), float(sc.logre[i]), float(ee1[i]),
                                     float(ee2[i]))
            assert np.abs(got[i] - ref).max() < 2e-6 * np.abs(ref).max(), (b, i)
            assert abs(got[i].sum() - ref.sum()) < 1e-6
        # STAR rows are the FFT backend's STAR rows
        fft = render_source_stamps(pad_psf(p, T), sc.x, sc.y, gx0, gy0, st, sc.logre, ee1,
                                   ee2, render_backend="fft").cpu().numpy()
        assert np.array_equal(got[st == 0], fft[st == 0])


def test_direct_backend_rejects_dev_and_sersic(cuda):
    from uniphot_amd.source_model import render_source_stamps
    sc = small_mixed_scene(types=(0, 2))
    with pytest.raises(ValueError, match="DEV/SERSIC"):
        render_source_stamps(sc.psfs[0], sc.x, sc.y, 0, 0, sc.stype, sc.logre, sc.ee1, sc.ee2,
                             render_backend="direct")
