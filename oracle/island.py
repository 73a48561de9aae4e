"""FP64 restatement of the per-island VarPro-LM (TEST INFRASTRUCTURE).

Reference: ``tract7dt/fit/grid_gpu/island_solver.py`` -- ``models_band``
(:480-562), ``_fluxes_band`` / ``residual_band`` / ``chi2_of``
(:570-684), ``hg_scan`` (:812-865), legacy ``one_step`` (:921-1002),
``diverged_probe`` / ``plateau_probe`` (:1335-1465),
``final_fluxes`` (:1467-1482), ``final_errors`` (:1484-1547), ``final_flux_marg``
(:1549-1612), ``source_metrics`` (:1643-1715), and the
chunk loop of ``run_class_chunk`` (:2890-2977).  Host input layout:
``island_plan.class_host_arrays`` (:493-699).  Scene subtraction:
``scene_star_image`` (island_stage.py:2040-2117) and ``member_addback``
(island_solver.py:1795-1818).

The Jacobian is formed the way the reference forms it: one explicit
whitened-residual tangent per live parameter, per pixel, through the
closed-form VarPro flux solve (d f = A^-1 (d b - d A f)), then
H = J J^T and g = J r.  (The HIP kernel uses Gram-matrix identities
instead, so the two are independent formulations.)

Parity pins: tests/test_oracle_island.py checks the VarPro fluxes against
``numpy.linalg.lstsq``, the Jacobian against float64 central finite
differences of the residual, and the LM invariants (chi2 non-increasing
over accepts, bounds and clips respected, frozen parameters untouched,
padded slots inert, convergence prongs), and the scene subtraction against
an independent Gaussian-PSF field (stars at integer positions cancel to
rounding, extended light untouched, members restored, non-integer shifts to
the Lanczos interpolation error).  Trajectories of the reference
JAX program itself are parity unpinned (JAX is not importable here).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import core

LM_LAMBDA0 = 1e-3
LM_UP, LM_DOWN = 10.0, 0.1
STALL_REJECT_K = 4


@dataclass
class LMOptions:
    """Legacy-step constants (island_solver.py:146-152, island_plan.py:55-74)."""
    reg: float = 1e-9
    pos_step_max: float = 0.5
    shape_step_max: float = 0.15
    un_step_max: float = 0.5
    ftol_rel: float = 1e-4
    step_tol: float = 0.1
    shape_step_tol: float = 0.01
    max_steps: int = 30


def clip_vec(N: int, o: LMOptions) -> np.ndarray:
    return np.tile([o.pos_step_max, o.pos_step_max, o.shape_step_max,
                    o.shape_step_max, o.shape_step_max, o.un_step_max], N)


_PNAMES = ("x", "y", "lr", "e1", "e2", "u")


class Island:
    """One island: window data + member description (float64 copies).

    Fields follow the ``isl`` pytree of island_solver.py:2742-2778:
    data, w (nb,B,B); x0, y0; stype, n_frozen, free_n, mmask (N,);
    keep (N,6); lower, upper (6N,).  ``psf`` is a list over bands of dicts
    (coefs (T,T,nt), x_origin, x_scale, y_origin, y_scale).
    """

    def __init__(self, data, w, x0, y0, stype, n_frozen, free_n, mmask, keep,
                 lower, upper, psf, T):
        self.data = np.asarray(data, np.float64)
        self.w = np.asarray(w, np.float64)
        self.nb, self.B = self.data.shape[0], self.data.shape[1]
        self.x0, self.y0 = float(x0), float(y0)
        self.stype = np.asarray(stype, np.int64)
        self.N = self.stype.size
        self.n_frozen = np.asarray(n_frozen, np.float64)
        self.free_n = np.asarray(free_n, bool)
        self.mmask = np.asarray(mmask, bool)
        self.keep = np.asarray(keep, bool) & self.mmask[:, None]
        self.lower = np.asarray(lower, np.float64)
        self.upper = np.asarray(upper, np.float64)
        self.psf = psf
        self.T = int(T)

    def live(self):
        """(member, param) index pairs of kept parameters (keep & mmask)."""
        return [(k, j) for k in range(self.N) for j in range(6) if self.keep[k, j]]


def models_band(p, isl: Island, b: int, want_grad: bool):
    """Placed unit-flux member models (N,B,B) for band b and, with
    ``want_grad``, a dict {(k, j): placed tangent plane} for live params."""
    P = p.reshape(isl.N, 6)
    B, T = isl.B, isl.T
    m = np.zeros((isl.N, B, B))
    tan = {}
    live = isl.live() if want_grad else []
    for k in range(isl.N):
        if not isl.mmask[k]:
            continue
        x, y = P[k, 0], P[k, 1]
        st = int(isl.stype[k])
        if st == core.SOURCE_STAR:
            lr, e1, e2 = 0.0, 1.0, 0.0
        else:
            lr, e1, e2 = P[k, 2], P[k, 3], P[k, 4]
            if e1 == 0.0 and e2 == 0.0:
                e1 = 1e-4
        if isl.free_n[k]:
            n_val = float(core.sersic_u_to_n(P[k, 5]))
        else:
            n_val = isl.n_frozen[k]
        jk = [j for (kk, j) in live if kk == k]
        want = []
        for j in jk:
            want.append({0: "x", 1: "y", 2: "lr", 3: "e1", 4: "e2", 5: "n"}[j])
        stamp, grads = core.member_stamp(isl.psf[b], x, y, isl.x0, isl.y0, st,
                                         lr, e1, e2, n_val, want=tuple(want))
        ix, iy = core.stamp_top_left(x, y, isl.x0, isl.y0, T)
        m[k] = core.place(stamp, ix, iy, B)
        for j in jk:
            name = {0: "x", 1: "y", 2: "lr", 3: "e1", 4: "e2", 5: "n"}[j]
            g = grads[name]
            if j == 5:
                nmin, nmax = core.SERSIC_N_MIN, core.SERSIC_N_MAX
                g = g * (n_val - nmin) * (nmax - n_val) / (nmax - nmin)
            tan[(k, j)] = core.place(g, ix, iy, B)
    return m, tan


def fluxes_band(m, d, w, mmask, reg):
    """Closed-form member fluxes (island_solver.py:570-590).  Returns
    (f (N,), A64 (N,N) with the padded/ridge diagonal, A_raw)."""
    N = m.shape[0]
    wm = w[None] * m
    A = np.einsum("kij,lij->kl", wm, m)
    b = np.einsum("kij,ij->k", wm, d)
    A64 = A + np.diag(np.where(mmask, reg, 1.0))
    f = np.linalg.solve(A64, b) * mmask
    return f, A64, A


def chi2_of(p, isl: Island, o: LMOptions) -> float:
    c = 0.0
    for b in range(isl.nb):
        m, _ = models_band(p, isl, b, False)
        f, _, _ = fluxes_band(m, isl.data[b], isl.w[b], isl.mmask, o.reg)
        r = np.sqrt(isl.w[b]) * (isl.data[b] - np.einsum("k,kij->ij", f, m))
        c += float(np.sum(r * r))
    return c


def hg(p, isl: Island, o: LMOptions):
    """Un-masked, un-damped flux-marginalised Gauss-Newton (H, g) over the
    live parameters, plus the chi2 at p (island_solver.py:812-865)."""
    live = isl.live()
    L = len(live)
    H = np.zeros((L, L))
    g = np.zeros(L)
    chi2 = 0.0
    for b in range(isl.nb):
        d, w = isl.data[b], isl.w[b]
        sw = np.sqrt(w)
        m, tan = models_band(p, isl, b, True)
        f, A64, _ = fluxes_band(m, d, w, isl.mmask, o.reg)
        model = np.einsum("k,kij->ij", f, m)
        r = sw * (d - model)
        chi2 += float(np.sum(r * r))
        J = np.zeros((L, isl.B * isl.B))
        for i, (k, j) in enumerate(live):
            t = tan[(k, j)]
            # dA = e_k u^T + u e_k^T, u_l = sum w m_l t ; db = e_k sum w t d
            u = np.einsum("lij,ij->l", w[None] * m, t)
            dA = np.zeros_like(A64)
            dA[k, :] += u
            dA[:, k] += u
            db = np.zeros(isl.N)
            db[k] = np.sum(w * t * d)
            df = np.linalg.solve(A64, db - dA @ f) * isl.mmask
            dmodel = np.einsum("l,lij->ij", df, m) + f[k] * t
            J[i] = (-sw * dmodel).ravel()
        H += J @ J.T
        g += J @ r.ravel()
    return H, g, chi2, live


def kappa_scaled(H) -> float:
    """1-norm condition number of H scaled to a unit diagonal,
    |Hs|_1 |Hs^-1|_1 with Hs = D^-1/2 H D^-1/2, D = diag(H): the measure the
    float32 fit uses to send an island to the float64 kernels (no
    reference counterpart; the reference has no such policy)."""
    H = np.asarray(H, np.float64)
    if H.size == 0:
        return 1.0
    s = 1.0 / np.sqrt(np.diag(H))
    Hs = H * s[:, None] * s[None, :]
    return float(np.abs(Hs).sum(axis=0).max() * np.abs(np.linalg.inv(Hs)).sum(axis=0).max())


def one_step(state: dict, isl: Island, o: LMOptions) -> dict:
    """Legacy damped LM step with accept/reject and the three convergence
    prongs (island_solver.py:921-1002).  ``state`` keys: p, lam, chi2,
    consec, nrej, conv, stall, nst."""
    if state["conv"]:
        return dict(state)
    N = isl.N
    nP = 6 * N
    p = state["p"]
    Hl, gl, _, live = hg(p, isl, o)
    H = np.zeros((nP, nP))
    g = np.zeros(nP)
    idx = np.array([6 * k + j for (k, j) in live], dtype=np.int64)
    if idx.size:
        H[np.ix_(idx, idx)] = Hl
        g[idx] = gl
    keep = np.zeros(nP, bool)
    keep[idx] = True
    H = np.where(keep[:, None] & keep[None, :], H, 0.0) + np.diag(np.where(keep, 0.0, 1.0))
    g = np.where(keep, g, 0.0)
    Hd = H + state["lam"] * np.diag(np.diag(H)) + 1e-8 * np.eye(nP)
    # the parameter vector is float32 state, as in the reference: the
    # float64 step is cast to float32, clipped and added in float32
    # (island_solver.py:963-969, p_flat f32)
    delta = (-np.linalg.solve(Hd, g)).astype(np.float32)
    delta = np.where(keep, delta, np.float32(0.0))
    cv = clip_vec(N, o).astype(np.float32)
    delta = np.clip(delta, -cv, cv)
    p_try = np.clip(p + delta, isl.lower.astype(np.float32), isl.upper.astype(np.float32))
    chi2_try = chi2_of(p_try, isl, o)
    chi2_cur = state["chi2"]
    accept = chi2_try < chi2_cur
    p_new = p_try if accept else p
    lam = state["lam"]
    lam_new = max(lam * LM_DOWN, 1e-9) if accept else min(lam * LM_UP, 1e6)
    chi2_new = chi2_try if accept else chi2_cur
    dchi2 = chi2_cur - chi2_try
    step = (p_new - p).reshape(N, 6)
    mmf = isl.mmask.astype(np.float32)
    pos_step = np.max(mmf * np.hypot(step[:, 0], step[:, 1])) if N else np.float32(0.0)
    shape_step = np.max(mmf[:, None] * np.abs(step[:, 2:6])) if N else np.float32(0.0)
    small = accept and (dchi2 < o.ftol_rel * max(chi2_cur, 1.0))
    consec = state["consec"] + 1 if small else (0 if accept else state["consec"])
    step_small = accept and pos_step < o.step_tol and shape_step < o.shape_step_tol
    nrej = 0 if accept else state["nrej"] + 1
    stall_now = nrej >= STALL_REJECT_K
    conv = bool(consec >= 2 or step_small or stall_now)
    return dict(p=p_new, lam=lam_new, chi2=chi2_new, consec=consec, nrej=nrej,
                conv=conv, stall=stall_now, nst=state["nst"] + 1,
                accept=accept, chi2_try=chi2_try, p_try=p_try)


def propose(p, lam, isl: Island, o: LMOptions, hg_at_p=None):
    """The legacy step's trial point at damping ``lam`` (one_step,
    island_solver.py:921-969, float32 parameter state).  ``hg_at_p`` reuses
    an (H, g, chi2, live) of :func:`hg` at p."""
    N = isl.N
    nP = 6 * N
    Hl, gl, _, live = hg(p, isl, o) if hg_at_p is None else hg_at_p
    H = np.zeros((nP, nP))
    g = np.zeros(nP)
    idx = np.array([6 * k + j for (k, j) in live], dtype=np.int64)
    if idx.size:
        H[np.ix_(idx, idx)] = Hl
        g[idx] = gl
    keep = np.zeros(nP, bool)
    keep[idx] = True
    H = np.where(keep[:, None] & keep[None, :], H, 0.0) + np.diag(np.where(keep, 0.0, 1.0))
    g = np.where(keep, g, 0.0)
    Hd = H + lam * np.diag(np.diag(H)) + 1e-8 * np.eye(nP)
    delta = (-np.linalg.solve(Hd, g)).astype(np.float32)
    delta = np.where(keep, delta, np.float32(0.0))
    cv = clip_vec(N, o).astype(np.float32)
    delta = np.clip(delta, -cv, cv)
    p = np.asarray(p, np.float32)
    return np.clip(p + delta, isl.lower.astype(np.float32), isl.upper.astype(np.float32))


def diverged_probe(p, lam_final, isl: Island, o: LMOptions) -> bool:
    """Post-hoc divergence test of a stalled island (island_solver.py:
    1335-1392): the STALL_REJECT_K rejected proposals, rebuilt at
    lam_final * 10^(j - K), are all identical."""
    h = hg(p, isl, o)
    props = [propose(p, lam_final * 10.0 ** (j - STALL_REJECT_K), isl, o, h)
             for j in range(STALL_REJECT_K)]
    return all(np.array_equal(q, props[0]) for q in props[1:])


PLATEAU_LAMBDAS = tuple(float(np.float32(v)) for v in [1e-9] + [10.0 ** e for e in range(-8, 7)])


def plateau_probe(p, isl: Island, o: LMOptions) -> float:
    """Best chi2 change over the damping ladder 1e-9 ... 1e6 of a stalled
    island (island_solver.py:1394-1465): min_j chi2(proposal(lam_j)) -
    chi2(p); negative when a damped step still improves."""
    h = hg(p, isl, o)
    c0 = chi2_of(p, isl, o)
    return min(chi2_of(propose(p, lm, isl, o, h), isl, o) for lm in PLATEAU_LAMBDAS) - c0


def run_island(p0, isl: Island, o: LMOptions, max_steps=None, trace=False):
    """Fresh-chunk LM loop for one island (run_class_chunk, :2890-2977):
    lam = 1e-3, chi2 = chi2(p0), until converged or ``max_steps``.  The
    parameters are float32 state (the reference's ``p0``/``p_flat``
    dtype); chi2 and the damping are float64."""
    max_steps = o.max_steps if max_steps is None else max_steps
    st = dict(p=np.asarray(p0, np.float32).ravel().copy(), lam=LM_LAMBDA0,
              chi2=chi2_of(p0, isl, o), consec=0, nrej=0, conv=False,
              stall=False, nst=0)
    hist = []
    for _ in range(max_steps):
        if st["conv"]:
            break
        st = one_step(st, isl, o)
        if trace:
            hist.append(dict(st))
    st["history"] = hist
    return st


def final_fluxes(p, isl: Island, o: LMOptions):
    """Per-band fluxes and sqrt(diag(A^-1)) (island_solver.py:1467-1482)."""
    f_all = np.zeros((isl.nb, isl.N))
    e_all = np.zeros((isl.nb, isl.N))
    for b in range(isl.nb):
        m, _ = models_band(p, isl, b, False)
        f, A64, _ = fluxes_band(m, isl.data[b], isl.w[b], isl.mmask, o.reg)
        f_all[b] = f
        e_all[b] = np.sqrt(np.maximum(np.diag(np.linalg.inv(A64)), 1e-30))
    return f_all, e_all


def final_errors(p, isl: Island, o: LMOptions):
    """Marginal variances of the nonlinear parameters from the inverse of
    the keep-masked GN Hessian (island_solver.py:1484-1547, no priors).
    Returns (N, 7): var_x, var_y, var_lr, var_e1, var_e2, var_u, cov_e1e2."""
    N = isl.N
    nP = 6 * N
    Hl, _, _, live = hg(p, isl, o)
    H = np.eye(nP)
    idx = np.array([6 * k + j for (k, j) in live], dtype=np.int64)
    if idx.size:
        H[np.ix_(idx, idx)] = Hl
    S = np.linalg.inv(H)
    var6 = np.diag(S).reshape(N, 6)
    s6 = 6 * np.arange(N)
    cov = S[s6 + 3, s6 + 4]
    return np.concatenate([var6, cov[:, None]], axis=1)


def flux_sensitivities(p, isl: Island, o: LMOptions, b: int):
    """dF*/dtheta (N, L) of band b's VarPro fluxes over the live parameters:
    the closed-form derivative of the flux solve, df = A^-1 (db - dA f),
    with dA = e_k u^T + u e_k^T, u_l = sum w m_l t, db = e_k sum w t d."""
    live = isl.live()
    d, w = isl.data[b], isl.w[b]
    m, tan = models_band(p, isl, b, True)
    f, A64, _ = fluxes_band(m, d, w, isl.mmask, o.reg)
    J = np.zeros((isl.N, len(live)))
    for i, (k, j) in enumerate(live):
        t = tan[(k, j)]
        u = np.einsum("lij,ij->l", w[None] * m, t)
        dA = np.zeros_like(A64)
        dA[k, :] += u
        dA[:, k] += u
        db = np.zeros(isl.N)
        db[k] = np.sum(w * t * d)
        J[:, i] = np.linalg.solve(A64, db - dA @ f) * isl.mmask
    return J, live


def final_flux_marg(p, isl: Island, o: LMOptions):
    """Per-band shape-marginalisation flux variance increment (nb, N)
    (island_solver.py:1549-1612, no priors): diag(J Sigma J^T) over the
    extended members' live parameters, Sigma the inverse of the GN Hessian
    final_errors inverts; star rows and padded slots 0."""
    Hl, _, _, live = hg(p, isl, o)
    out = np.zeros((isl.nb, isl.N))
    if not live:
        return out
    S = np.linalg.inv(Hl)
    ext = (isl.stype != core.SOURCE_STAR) & isl.mmask
    cols = np.array([ext[k] for (k, _) in live], np.float64)
    for b in range(isl.nb):
        J, _ = flux_sensitivities(p, isl, o, b)
        Jm = J * cols[None]
        dv = np.einsum("ka,ab,kb->k", Jm, S, Jm)
        out[b] = np.where(ext, np.maximum(dv, 0.0), 0.0)
    return out


def source_metrics(p, isl: Island, o: LMOptions):
    """Per-member (rchisq, fracflux, fracin, fracmasked) (N, 4) and the
    per-band window flux f * sum(m) (nb, N) at p (island_solver.py:1643-
    1715): unit models m on the window, VarPro fluxes f, model = sum f m;
    per band fin = clip(sum m, 0, 1), fracflux ratio sum (model - f m) f m /
    sum (f m)^2, masked fraction clip(sum over zero-weight pixels of m, 0, 1),
    rchisq numerator sum m w (d - model)^2 over denominator |sum m|."""
    N, nb = isl.N, isl.nb
    z = np.zeros(N)
    ff_n, ff_d, fin_n, fin_d, fm_n, rc_n, rc_d = (z.copy() for _ in range(7))
    fwin = np.zeros((nb, N))
    for b in range(nb):
        d, w = isl.data[b], isl.w[b]
        m, _ = models_band(p, isl, b, False)
        f, _, _ = fluxes_band(m, d, w, isl.mmask, o.reg)
        model = np.einsum("k,kij->ij", f, m)
        own = f[:, None, None] * m
        s_m = m.sum(axis=(1, 2))
        fin = np.clip(s_m, 0.0, 1.0)
        bot = (own * own).sum(axis=(1, 2))
        top = ((model[None] - own) * own).sum(axis=(1, 2))
        ratio = np.where(bot > 0, top / np.where(bot > 0, bot, 1.0), 0.0)
        fm = np.clip(((w == 0.0)[None] * m).sum(axis=(1, 2)), 0.0, 1.0)
        rc = (m * (w * (d - model) ** 2)[None]).sum(axis=(1, 2))
        ff_n += fin * ratio
        ff_d += fin
        fin_n += np.abs(s_m)
        fin_d += 1.0
        fm_n += fin * fm
        rc_n += np.abs(rc)
        rc_d += np.abs(s_m)
        fwin[b] = f * s_m
    mm = isl.mmask
    with np.errstate(divide="ignore", invalid="ignore"):
        fracflux = np.where((ff_d > 0) & mm, ff_n / ff_d, 0.0)
        fracin = np.where((fin_d > 0) & mm, fin_n / fin_d, np.nan)
        fracmasked = np.where((ff_d > 0) & mm, fm_n / ff_d, 0.0)
        rchisq = np.where((rc_d > 0) & mm, rc_n / rc_d, np.nan)
    return np.stack([rchisq, fracflux, fracin, fracmasked], axis=1), np.where(mm[None], fwin,
                                                                             np.nan)


def scene_star_image(shape, pad: int, psfs, x, y, stype, flux):
    """Field-wide STAR scene of the scene subtraction (island_stage.py:1846-
    1851, :2040-2117) on a padded (nb, Hp, Wp) stack: every STAR k with
    flux[k, b] != 0 contributes flux[k, b] times its native unit stamp (PSF at
    (x, y), Lanczos-shifted by x - round(x), y - round(y)) at top-left
    round(pos) - T // 2 + pad, clipped to the stack.  The stack minus this
    image is what the islands gather."""
    nb, Hp, Wp = shape
    out = np.zeros(shape)
    T = psfs[0]["coefs"].shape[0]
    for k in range(len(x)):
        if int(stype[k]) != core.SOURCE_STAR:
            continue
        ix, iy = core.stamp_top_left(float(x[k]), float(y[k]), 0.0, 0.0, T)
        ix += pad
        iy += pad
        y0, x0 = max(0, iy), max(0, ix)
        y1, x1 = min(Hp, iy + T), min(Wp, ix + T)
        if y1 <= y0 or x1 <= x0:
            continue
        for b in range(nb):
            f = float(flux[k][b])
            if f == 0.0:
                continue
            st, _ = core.member_stamp(psfs[b], float(x[k]), float(y[k]), 0.0, 0.0,
                                      core.SOURCE_STAR)
            out[b, y0:y1, x0:x1] += f * st[y0 - iy:y1 - iy, x0 - ix:x1 - ix]
    return out


def member_addback(psfs_tier, x0b, y0b, B: int, xs, ys, stypes, flux):
    """Member add-back of one island (island_solver.py:1795-1818 at the seeds,
    member fluxes from island_plan.py:664-680): (nb, B, B) sum over member
    STARS of flux[k, b] times the unit stamp at the class's stamp tier placed
    in the B x B window at origin (x0b, y0b); extended members add nothing."""
    nb = len(psfs_tier)
    T = psfs_tier[0]["coefs"].shape[0]
    out = np.zeros((nb, B, B))
    for k in range(len(xs)):
        if int(stypes[k]) != core.SOURCE_STAR:
            continue
        ix, iy = core.stamp_top_left(float(xs[k]), float(ys[k]), float(x0b), float(y0b), T)
        for b in range(nb):
            f = float(flux[k][b])
            if f == 0.0:
                continue
            st, _ = core.member_stamp(psfs_tier[b], float(xs[k]), float(ys[k]), float(x0b),
                                      float(y0b), core.SOURCE_STAR)
            out[b] += f * core.place(st, ix, iy, B)
    return out
