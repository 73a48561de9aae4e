"""Pins of the float64 island oracle that do not restate its formulas:
VarPro fluxes against numpy.linalg.lstsq, the analytic Gauss-Newton (H, g)
against central finite differences of the VarPro residual, and LM
behaviour (monotone chi2 over accepts, bounds, clips, frozen and padded
slots, the stall and step-size convergence prongs) on islands with known
answers."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import core
from oracle import island as oisl

from _scenes import mixed_classes, oracle_island, small_mixed_scene, small_star_scene, star_classes


def _psf_dict(p, T):
    return dict(coefs=core.pad_psf_coefs(p.coefs.astype(np.float64), T), x_origin=p.x_origin,
                x_scale=p.x_scale, y_origin=p.y_origin, y_scale=p.y_scale)


def test_varpro_fluxes_match_lstsq():
    rng = np.random.default_rng(3)
    N, B = 5, 24
    m = rng.uniform(0, 1, (N, B, B)) ** 3
    d = rng.normal(0, 1, (B, B)) + np.einsum("k,kij->ij", rng.uniform(1, 9, N), m)
    w = rng.uniform(0.2, 2.0, (B, B))
    w[3:7, 5:11] = 0.0                       # masked pixels carry no weight
    mmask = np.array([1, 1, 0, 1, 1], bool)  # slot 2 is padding: no model
    m[~mmask] = 0.0
    f, A64, _ = oisl.fluxes_band(m, d, w, mmask, 1e-9)
    sw = np.sqrt(w).ravel()
    X = (m[mmask].reshape(mmask.sum(), -1) * sw).T
    ref, *_ = np.linalg.lstsq(X, d.ravel() * sw, rcond=None)
    np.testing.assert_allclose(f[mmask], ref, rtol=1e-7, atol=1e-9)
    assert f[2] == 0.0
    assert A64[2, 2] == A64[2, 2] and A64[2, 2] >= 1.0  # padded slot: unit diagonal


def _varpro_residual(p, isl, o):
    out = []
    for b in range(isl.nb):
        m, _ = oisl.models_band(p, isl, b, False)
        f, _, _ = oisl.fluxes_band(m, isl.data[b], isl.w[b], isl.mmask, o.reg)
        out.append((np.sqrt(isl.w[b]) * (isl.data[b] - np.einsum("k,kij->ij", f, m))).ravel())
    return np.concatenate(out)


def _away_from_half(p, isl):
    """Shift positions so that x - x0 is not near a half-integer (the
    integer placement changes there, which finite differences straddle)."""
    P = p.reshape(isl.N, 6).copy()
    for k in range(isl.N):
        for j, o0 in ((0, isl.x0), (1, isl.y0)):
            fr = (P[k, j] - o0) % 1.0
            if abs(fr - 0.5) < 0.08:
                P[k, j] += 0.15
    return P.ravel()


@pytest.mark.parametrize("kind", ["star", "mixed"])
def test_gauss_newton_matches_finite_differences(kind):
    if kind == "star":
        sc = small_star_scene(seed=11, n_bands=2, n_src=6, order=1)
        plans, hosts, pad, d, w = star_classes(sc)
    else:
        sc = small_mixed_scene(seed=5, n_bands=2, n_src=6, order=1, types=(0, 1, 3))
        free = np.asarray(sc.stype == 3)
        plans, hosts, pad, d, w = mixed_classes(sc, free_n=free)
    # the island with the most live parameters
    best = max(((pi, r) for pi, pl in enumerate(plans) for r in range(pl.n_islands)),
               key=lambda t: int(hosts[t[0]]["keep"][t[1]].sum()))
    plan, host = plans[best[0]], hosts[best[0]]
    isl = oracle_island(plan, host, best[1], d, w, sc.psfs)
    o = oisl.LMOptions()
    p = _away_from_half(host["p0"][best[1]].ravel().astype(np.float64), isl)
    H, g, chi2, live = oisl.hg(p, isl, o)
    r0 = _varpro_residual(p, isl, o)
    assert chi2 == pytest.approx(float(r0 @ r0), rel=1e-12)
    J = np.zeros((len(live), r0.size))
    for i, (k, j) in enumerate(live):
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[6 * k + j] += h
        pm[6 * k + j] -= h
        J[i] = (_varpro_residual(pp, isl, o) - _varpro_residual(pm, isl, o)) / (2 * h)
    g_fd = J @ r0
    H_fd = J @ J.T
    scale_g = np.abs(g_fd).max()
    np.testing.assert_allclose(g, g_fd, rtol=2e-4, atol=2e-5 * scale_g)
    np.testing.assert_allclose(H, H_fd, rtol=2e-4, atol=2e-5 * np.abs(H_fd).max())
    if kind == "mixed":
        assert any(j >= 2 for (_, j) in live)   # shape parameters were exercised


def _noiseless_island(offset, seed=2):
    """One two-star island whose data is exactly the model at the truth
    (float64 oracle render, unit weights); seeds offset by ``offset`` px."""
    sc = small_star_scene(seed=seed, n_bands=2, n_src=6, order=1)
    plans, hosts, pad, d, w = star_classes(sc)
    pi, r = next((pi, r) for pi, pl in enumerate(plans) for r in range(pl.n_islands)
                 if hosts[pi]["mmask"][r].sum() == 2)
    plan, host = plans[pi], hosts[pi]
    isl = oracle_island(plan, host, r, d, w, sc.psfs)
    p_true = _away_from_half(host["p0"][r].ravel().astype(np.float64), isl)
    fl = np.array([[2e4, 5e3], [1e4, 8e3]])
    mm = isl.mmask
    for b in range(isl.nb):
        m, _ = oisl.models_band(p_true, isl, b, False)
        f = np.zeros(isl.N)
        f[np.where(mm)[0]] = fl[b]
        isl.data[b] = np.einsum("k,kij->ij", f, m)
        isl.w[b] = np.ones_like(isl.w[b])
    p0 = p_true.copy()
    P = p0.reshape(isl.N, 6)
    P[mm, 0] += offset
    P[mm, 1] -= offset
    return isl, p_true, p0


def test_lm_converges_on_step_prong_from_a_small_offset():
    isl, p_true, p0 = _noiseless_island(0.03)
    o = oisl.LMOptions()
    st = oisl.run_island(p0, isl, o, trace=True)
    h = st["history"]
    assert h[0]["accept"] and st["conv"] and not st["stall"]
    assert st["nst"] == 1                       # pos step 0.04 px < 0.1, shape step 0
    P = st["p"].reshape(isl.N, 6)[isl.mmask]
    T = p_true.reshape(isl.N, 6)[isl.mmask]
    assert np.abs(P[:, :2] - T[:, :2]).max() < 1e-3
    f, _ = oisl.final_fluxes(st["p"], isl, o)
    np.testing.assert_allclose(f[:, isl.mmask], [[2e4, 5e3], [1e4, 8e3]], rtol=1e-3)


def test_lm_stalls_after_four_rejects_when_steps_are_clipped_to_zero():
    isl, _, p0 = _noiseless_island(0.3)
    o = oisl.LMOptions(pos_step_max=0.0, shape_step_max=0.0, un_step_max=0.0)
    st = oisl.run_island(p0, isl, o, trace=True)
    h = st["history"]
    # p_try == p gives chi2_try == chi2, which the strict test rejects
    assert [s["accept"] for s in h] == [False] * 4
    assert [s["nrej"] for s in h] == [1, 2, 3, 4]
    np.testing.assert_allclose([s["lam"] for s in h], [1e-2, 1e-1, 1.0, 10.0])
    assert st["conv"] and st["stall"] and st["nst"] == 4
    np.testing.assert_array_equal(st["p"], np.float32(p0))  # float32 state, unchanged


def test_lm_invariants_on_a_noisy_mixed_island():
    sc = small_mixed_scene(seed=9, n_bands=2, n_src=8, order=1)
    plans, hosts, pad, d, w = mixed_classes(sc)
    pi, r = max(((pi, r) for pi, pl in enumerate(plans) for r in range(pl.n_islands)),
                key=lambda t: int(hosts[t[0]]["mmask"][t[1]].sum()))
    plan, host = plans[pi], hosts[pi]
    isl = oracle_island(plan, host, r, d, w, sc.psfs)
    o = oisl.LMOptions(max_steps=8)
    p0 = host["p0"][r].ravel().astype(np.float64)
    st = oisl.run_island(p0, isl, o, trace=True)
    cv = oisl.clip_vec(isl.N, o)
    prev_p, prev_chi2 = p0, oisl.chi2_of(p0, isl, o)
    keep = isl.keep.ravel()
    for s in st["history"]:
        assert s["chi2"] <= prev_chi2                     # never increases
        d_try = s["p_try"] - prev_p
        assert np.all(np.abs(d_try) <= cv + 1e-12)        # per-parameter clips
        assert np.all(s["p_try"] >= isl.lower - 1e-12) and np.all(s["p_try"] <= isl.upper + 1e-12)
        assert np.all(s["p"][~keep] == p0[~keep])         # frozen / padded / star shape
        if s["accept"]:
            assert s["chi2"] == s["chi2_try"] < prev_chi2
            np.testing.assert_array_equal(s["p"], s["p_try"])
        else:
            np.testing.assert_array_equal(s["p"], prev_p)
        prev_p, prev_chi2 = s["p"], s["chi2"]
    f, e = oisl.final_fluxes(st["p"], isl, o)
    assert np.all(f[:, ~isl.mmask] == 0.0)
    assert np.all(np.isfinite(e)) and np.all(e > 0)


def test_final_errors_invert_the_finite_difference_fisher():
    sc = small_star_scene(seed=4, n_bands=1, n_src=4, order=1)
    plans, hosts, pad, d, w = star_classes(sc)
    pi, r = 0, 0
    isl = oracle_island(plans[pi], hosts[pi], r, d, w, sc.psfs)
    o = oisl.LMOptions()
    p = _away_from_half(hosts[pi]["p0"][r].ravel().astype(np.float64), isl)
    ev = oisl.final_errors(p, isl, o)
    live = isl.live()
    r0 = _varpro_residual(p, isl, o)
    J = np.zeros((len(live), r0.size))
    for i, (k, j) in enumerate(live):
        pp, pm = p.copy(), p.copy()
        pp[6 * k + j] += 1e-5
        pm[6 * k + j] -= 1e-5
        J[i] = (_varpro_residual(pp, isl, o) - _varpro_residual(pm, isl, o)) / 2e-5
    C = np.linalg.inv(J @ J.T)
    for i, (k, j) in enumerate(live):
        assert ev[k, j] == pytest.approx(C[i, i], rel=1e-3)


def _fluxes_of(p, isl, o, b):
    m, _ = oisl.models_band(p, isl, b, False)
    f, _, _ = oisl.fluxes_band(m, isl.data[b], isl.w[b], isl.mmask, o.reg)
    return f


def test_flux_marginalisation_uses_finite_difference_flux_sensitivities():
    """final_flux_marg = diag(J Sigma J^T) with J the flux solve's
    sensitivity to the extended members' parameters: J from central finite
    differences of the fluxes, Sigma from the finite-difference Fisher
    matrix; star rows exactly 0."""
    sc = small_mixed_scene(seed=5, n_bands=2, n_src=6, order=1, types=(0, 1, 3))
    plans, hosts, pad, d, w = mixed_classes(sc, free_n=np.asarray(sc.stype == 3))
    best = max(((pi, r) for pi, pl in enumerate(plans) for r in range(pl.n_islands)
                if pl.kind == "mixed"), key=lambda t: int(hosts[t[0]]["keep"][t[1]].sum()))
    plan, host = plans[best[0]], hosts[best[0]]
    isl = oracle_island(plan, host, best[1], d, w, sc.psfs)
    o = oisl.LMOptions()
    p = _away_from_half(host["p0"][best[1]].ravel().astype(np.float64), isl)
    live = isl.live()
    ext = (isl.stype != core.SOURCE_STAR) & isl.mmask
    assert ext.any() and (isl.mmask & ~ext).any() or ext.any()
    r0 = _varpro_residual(p, isl, o)
    Jr = np.zeros((len(live), r0.size))
    Jf = np.zeros((isl.nb, isl.N, len(live)))
    for i, (k, j) in enumerate(live):
        pp, pm = p.copy(), p.copy()
        pp[6 * k + j] += 1e-5
        pm[6 * k + j] -= 1e-5
        Jr[i] = (_varpro_residual(pp, isl, o) - _varpro_residual(pm, isl, o)) / 2e-5
        for b in range(isl.nb):
            Jf[b, :, i] = (_fluxes_of(pp, isl, o, b) - _fluxes_of(pm, isl, o, b)) / 2e-5
    S = np.linalg.inv(Jr @ Jr.T)
    cols = np.array([ext[k] for (k, _) in live], np.float64)
    dv = oisl.final_flux_marg(p, isl, o)
    for b in range(isl.nb):
        J, _ = oisl.flux_sensitivities(p, isl, o, b)
        np.testing.assert_allclose(J, Jf[b], rtol=1e-4, atol=1e-6 * np.abs(Jf[b]).max())
        Jm = Jf[b] * cols[None]
        ref = np.where(ext, np.einsum("ka,ab,kb->k", Jm, S, Jm), 0.0)
        np.testing.assert_allclose(dv[b], ref, rtol=2e-3, atol=1e-9 * np.abs(ref).max())
        assert np.all(dv[b][~ext] == 0.0)


def test_source_metrics_closed_forms():
    """Isolated member: fracflux 0; data equal to the model: rchisq 0; one
    band: fracin = |sum m|, window flux f sum m; zero weight on a pixel set:
    fracmasked is the clipped unit-model sum over it."""
    sc = small_star_scene(seed=4, n_bands=1, n_src=4, order=1)
    plans, hosts, pad, d, w = star_classes(sc)
    pi, r = 0, 0
    isl = oracle_island(plans[pi], hosts[pi], r, d, w, sc.psfs)
    k0 = int(np.where(isl.mmask)[0][0])
    isl.mmask[:] = False                        # every other member a padded slot
    isl.mmask[k0] = True
    isl.keep &= isl.mmask[:, None]
    o = oisl.LMOptions()
    p = _away_from_half(hosts[pi]["p0"][r].ravel().astype(np.float64), isl)
    k = int(np.where(isl.mmask)[0][0])
    m, _ = oisl.models_band(p, isl, 0, False)
    isl.data[0] = 3e4 * m[k]
    isl.w[0] = np.full_like(isl.w[0], 0.01)
    met, fwin = oisl.source_metrics(p, isl, o)
    assert met[k, 1] == 0.0                     # fracflux of an isolated member
    # rchisq of an exact model: only the ridge of the flux solve remains,
    # f = b / (A + reg), R = 3e4 m reg / (A + reg)
    A = 0.01 * np.sum(m[k] ** 2)
    c = 3e4 * o.reg / (A + o.reg)
    assert met[k, 0] == pytest.approx(0.01 * c * c * np.sum(m[k] ** 3) / abs(m[k].sum()),
                                      rel=1e-6)
    assert fwin[0, k] == pytest.approx((3e4 - c) * m[k].sum(), rel=1e-9)
    assert met[k, 2] == pytest.approx(abs(m[k].sum()), rel=1e-12)   # fracin, one band
    isl.w[0][:, : isl.B // 2] = 0.0             # masked: the left half of the window
    met, _ = oisl.source_metrics(p, isl, o)
    expect = np.clip(m[k][:, : isl.B // 2].sum(), 0.0, 1.0)
    assert met[k, 3] == pytest.approx(expect, rel=1e-12)
    assert 0.0 < expect < 1.0


def _gauss_psf(T=15, sigma=1.3):
    r = np.arange(T) - T // 2
    g = np.exp(-0.5 * (r[:, None] ** 2 + r[None, :] ** 2) / sigma ** 2)
    z = g.sum()
    psf = dict(coefs=(g / z)[:, :, None], x_origin=0.0, x_scale=1.0, y_origin=0.0, y_scale=1.0)
    return psf, z


def _gauss_field(H, W, x, y, sigma, z):
    """Unit Gaussian sampled at the pixel centres of an H x W frame,
    normalised by the stamp's discrete sum z (independent of the renderer)."""
    yy, xx = np.mgrid[0:H, 0:W]
    return np.exp(-0.5 * ((xx - x) ** 2 + (yy - y) ** 2) / sigma ** 2) / z


def test_scene_subtract_removes_stars_and_keeps_extended_light():
    """Stars at integer positions (the Lanczos shift is then the identity):
    the stack minus the star scene is exactly the extended light, clipping
    at the stack edge included; the member add-back restores the island's
    own stars in its window and nothing else."""
    T, sig, pad = 15, 1.3, 8
    psf, z = _gauss_psf(T, sig)
    psfs = [psf, psf]
    nb, H, W = 2, 44, 50
    Hp, Wp = H + 2 * pad, W + 2 * pad
    rng = np.random.default_rng(3)
    x = np.array([5.0, 20.0, 23.0, -6.0, 30.0, 45.0])
    y = np.array([7.0, 20.0, 22.0, 10.0, 35.0, 40.0])
    st = np.array([0, 0, 0, 0, 1, 0])             # source 4 extended
    flux = rng.uniform(100, 1000, (6, nb))
    flux[5, 1] = 0.0                              # zero S1 flux: not subtracted in band 1
    gal = np.zeros((nb, Hp, Wp))
    gal[:, pad + 30:pad + 40, pad + 25:pad + 36] = rng.uniform(1, 5, (nb, 10, 11))
    stars = np.zeros((nb, Hp, Wp))
    for k in range(6):
        if st[k] == 0:
            ix, iy = int(x[k]) - T // 2 + pad, int(y[k]) - T // 2 + pad
            box = np.zeros((Hp, Wp))
            y0, x0 = max(iy, 0), max(ix, 0)
            y1, x1 = min(iy + T, Hp), min(ix + T, Wp)
            g = _gauss_field(Hp, Wp, x[k] + pad, y[k] + pad, sig, z)
            box[y0:y1, x0:x1] = g[y0:y1, x0:x1]
            for b in range(nb):
                stars[b] += flux[k, b] * box
    scene = oisl.scene_star_image((nb, Hp, Wp), pad, psfs, x, y, st, flux)
    left = stars - scene
    # source 5's band-1 light stays, everything else cancels
    ix5, iy5 = int(x[5]) - T // 2 + pad, int(y[5]) - T // 2 + pad
    left[1, max(iy5, 0):iy5 + T, max(ix5, 0):ix5 + T] = 0.0
    assert np.max(np.abs(left)) < 1e-12 * flux.max()
    assert np.max(np.abs((stars + gal - scene)[0] - gal[0])) < 1e-12 * flux.max()
    # island {1, 2, 4} in a 24 x 24 window at field origin (12, 14)
    mem = np.array([1, 2, 4])
    x0b, y0b, B = 12.0, 14.0, 24
    add = oisl.member_addback(psfs, x0b, y0b, B, x[mem], y[mem], st[mem], flux[mem])
    ox, oy = int(x0b) + pad, int(y0b) + pad
    cut = (stars + gal - scene)[:, oy:oy + B, ox:ox + B] + add
    own = np.zeros((nb, Hp, Wp))
    for k in (1, 2):
        g = _gauss_field(Hp, Wp, x[k] + pad, y[k] + pad, sig, z)
        for b in range(nb):
            own[b] += flux[k, b] * g
    want = (own + gal)[:, oy:oy + B, ox:ox + B]
    # the stars lie well inside the window, so their boxes are whole there
    assert np.max(np.abs(cut - want)) < 1e-9 * flux.max()
    # a larger stamp tier adds the same light (zero ring around the native box)
    ptier = [dict(psf, coefs=core.pad_psf_coefs(psf["coefs"], 25))] * nb
    add25 = oisl.member_addback(ptier, x0b, y0b, B, x[mem], y[mem], st[mem], flux[mem])
    assert np.max(np.abs(add25 - add)) < 1e-12 * flux.max()


def test_scene_subtract_shifts_stars_to_their_subpixel_position():
    """Non-integer positions: the subtracted star matches the Gaussian
    sampled at its true position to the Lanczos-3 interpolation error, and
    the mirrored shift (a sign slip) is far outside that."""
    T, sig, pad = 21, 1.6, 12
    psf, z = _gauss_psf(T, sig)
    Hp = Wp = 48 + 2 * pad
    x, y = np.array([20.3]), np.array([25.6])
    img = oisl.scene_star_image((1, Hp, Wp), pad, [psf], x, y, np.array([0]),
                                np.array([[1.0]]))[0]
    true = _gauss_field(Hp, Wp, x[0] + pad, y[0] + pad, sig, z)
    mirror = _gauss_field(Hp, Wp, 2 * np.rint(x[0]) - x[0] + pad,
                          2 * np.rint(y[0]) - y[0] + pad, sig, z)
    peak = true.max()
    err = np.max(np.abs(img - true))
    assert err < 1e-2 * peak, err / peak          # measured 6.2e-3 at sigma 1.6 px
    assert np.max(np.abs(img - mirror)) > 0.2 * peak


def test_kappa_scaled_closed_forms():
    """Unit-diagonal 2x2 [[1, r], [r, 1]]: kappa_1 = (1 + |r|) / (1 - |r|);
    a diagonal matrix of any scale: 1; invariant under diagonal scaling."""
    for r in (0.0, 0.3, -0.9, 0.999):
        H = np.array([[1.0, r], [r, 1.0]])
        assert oisl.kappa_scaled(H) == pytest.approx((1 + abs(r)) / (1 - abs(r)), rel=1e-9)
    assert oisl.kappa_scaled(np.diag([1e-6, 3.0, 1e8])) == pytest.approx(1.0)
    rng = np.random.default_rng(0)
    A = rng.normal(size=(5, 5))
    H = A @ A.T + np.eye(5)
    d = np.diag(rng.uniform(0.1, 10.0, 5))
    assert oisl.kappa_scaled(d @ H @ d) == pytest.approx(oisl.kappa_scaled(H), rel=1e-9)


def test_stall_probes_known_answers():
    """Clip-pinned steps: every rejected proposal is p itself, so the
    divergence probe fires and the plateau probe's best change is exactly 0.
    A noiseless island 0.3 px off: the proposals differ with the damping
    (no divergence) and the ladder's best rung recovers at least 99 % of
    chi2 (the Gauss-Newton step of a nearly quadratic problem)."""
    isl, _, p0 = _noiseless_island(0.3)
    pinned = oisl.LMOptions(pos_step_max=0.0, shape_step_max=0.0, un_step_max=0.0)
    st = oisl.run_island(p0, isl, pinned)
    assert st["stall"]
    assert oisl.diverged_probe(st["p"], st["lam"], isl, pinned)
    assert oisl.plateau_probe(st["p"], isl, pinned) == 0.0
    o = oisl.LMOptions()
    p = np.float32(p0)
    assert not oisl.diverged_probe(p, 10.0, isl, o)
    best = oisl.plateau_probe(p, isl, o)
    c0 = oisl.chi2_of(p, isl, o)
    assert best < -0.99 * c0          # >= 99 % of chi2 recovered by one rung
    # the best rung equals the brute-force minimum over the ladder
    h = oisl.hg(p, isl, o)
    brute = min(oisl.chi2_of(oisl.propose(p, lm, isl, o, h), isl, o) for lm in
                [1e-9] + [10.0 ** e for e in range(-8, 7)]) - c0
    assert best == pytest.approx(brute, rel=1e-6)


def test_bound_pinned_island_diverges():
    """Lower = upper = p: every proposal is clipped back to p."""
    isl, _, p0 = _noiseless_island(0.3)
    p = np.float32(p0)
    isl.lower = p.astype(np.float64)
    isl.upper = p.astype(np.float64)
    assert oisl.diverged_probe(p, 1e-2, isl, oisl.LMOptions())
