"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point the header declares, the ctypes structs have the C
layout (compiled against the header with gcc), argument errors come back
as status codes with a message before any device work, and the host
planning logic (islands, buckets, grids, stamp tiers, PSF device layout)
agrees with brute-force restatements."""
from __future__ import annotations

import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from uniphot_amd import _lib
from uniphot_amd.grid import assign_sources_to_grids, divide_image_into_grids
from uniphot_amd.island_plan import plan_island_classes
from uniphot_amd.islands import (ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES, bucket_islands,
                                 build_islands)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "uniphot_amd.h")


def _header_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(uph_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def L():
    if not _lib.LIB_PATH.exists():
        from uniphot_amd import build
        build.build()
    return _lib.lib()


def test_library_exports_every_declared_entry_point(L):
    declared = _header_functions()
    assert declared == sorted(_lib.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name
    assert L.uph_version() >= 100


_STRUCTS = {
    "uph_lm_options": _lib.LMOptions,
    "uph_island_class": _lib.IslandClass,
    "uph_island_inputs": _lib.IslandInputs,
    "uph_island_outputs": _lib.IslandOutputs,
    "uph_island_batch": _lib.IslandBatch,
    "uph_island_split": _lib.IslandSplit,
}


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_ctypes_structs_match_the_c_layout(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"',
             "int main(void) {"]
    for cname, py in _STRUCTS.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for ln in out.splitlines():
        s, f, v = ln.split()
        got[(s, f)] = int(v)
    for cname, py in _STRUCTS.items():
        assert got[(cname, "size")] == C.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])


def test_scratch_size_queries(L):
    a = L.uph_island_scratch_per_block(4, 48, 55, 8, 0)
    b = L.uph_island_scratch_per_block(4, 48, 55, 8, 1)
    assert a % 256 == 0 and b > a                       # float64 planes are larger
    assert L.uph_island_scratch_per_block(4, 64, 55, 8, 0) > a
    gal = L.uph_island_scratch_per_block(4, 48, 55, 24, 0)
    assert gal > L.uph_island_scratch_per_block(4, 48, 55, 8, 0)  # galaxy workspace
    r1, r4 = L.uph_render_scratch_bytes(111, 1), L.uph_render_scratch_bytes(111, 4)
    assert r4 == 4 * r1 > 0
    assert L.uph_band_scratch_bytes(10, 40, 1) > L.uph_band_scratch_bytes(10, 20, 1) > 0


def test_argument_errors_return_status_and_message(L):
    st = L.uph_island_fit(None, None, None, None, 0, None, 0, None, None)
    assert st == 1 and b"null argument" in L.uph_last_error()
    opts = _lib.LMOptions(1e-9, 0.5, 0.15, 0.5, 1e-4, 0.1, 0.01, 30)
    cls = _lib.IslandClass()
    cls.n_isl, cls.N, cls.B, cls.T, cls.n_terms, cls.n_bands = 1, 3, 40, 55, 15, 2
    st = L.uph_island_fit(C.byref(cls), C.byref(_lib.IslandInputs()),
                          C.byref(_lib.IslandOutputs()), C.byref(opts), 0, None, 0, None, None)
    assert st == 1 and b"multiple of 16" in L.uph_last_error()
    with pytest.raises(_lib.UphError) as ei:
        _lib.check(L.uph_render_stamps(4, 54, 15, *([None] * 13), 0, None, None, 0, None))
    assert ei.value.code == 1 and "odd" in str(ei.value)
    with pytest.raises(_lib.UphError):
        _lib.check(L.uph_band_solve(-1, 55, 1, 32, *([None] * 2), 0, 0, 0, 0, *([None] * 3),
                                    0.0, 1.0, 0.0, 1.0, *([None] * 4), 1e-9, *([None] * 4),
                                    None, 0, None))
    with pytest.raises(_lib.UphError) as ei:
        _lib.check(L.uph_stamp_sum(1, 64, 55, 1, 3, 0, 2, *([None] * 7), -1.0, 64, 64, None,
                                   None))
    assert "bands [0, 3) of 2" in str(ei.value)
    with pytest.raises(_lib.UphError) as ei:
        _lib.check(L.uph_gather_windows(1, 32, 2, None, 64, 64, None, None, 64, 64, None, None,
                                        None))
    assert "null argument" in str(ei.value)


# ---------------------------------------------------------------------------
# host planning logic


def _union_find_labels(px, py, r):
    n = px.size
    parent = list(range(n))

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for i in range(n):
        for j in range(i + 1, n):
            if (px[i] - px[j]) ** 2 + (py[i] - py[j]) ** 2 < (r[i] + r[j]) ** 2:
                a, b = find(i), find(j)
                if a != b:
                    parent[max(a, b)] = min(a, b)
    roots = [find(i) for i in range(n)]
    # label islands in order of their smallest member index
    order = {}
    for i in range(n):
        order.setdefault(roots[i], len(order))
    return np.array([order[roots[i]] for i in range(n)])


def test_islands_match_brute_force_union_find():
    rng = np.random.default_rng(5)
    par = rng.uniform(0, 300, (60, 2))
    pts = np.concatenate([par, par[rng.integers(60, size=90)] + rng.normal(0, 5, (90, 2))])
    px, py = pts[:, 0], pts[:, 1]
    r = rng.uniform(2, 9, px.size)
    iset = build_islands(px, py, None, None, None, 27.5, radius=r)
    ref = _union_find_labels(px, py, r)
    np.testing.assert_array_equal(iset.labels, ref)
    for k, m in enumerate(iset.members):
        assert m.min() == np.where(ref == k)[0].min()
        assert iset.bbox[k, 0] == np.floor((px[m] - r[m]).min())
        assert iset.bbox[k, 3] == np.ceil((py[m] + r[m]).max())


def test_bucketing_picks_the_smallest_class_or_giant():
    rng = np.random.default_rng(8)
    par = rng.uniform(0, 2000, (400, 2))
    k = rng.integers(1, 40, 400)
    pts = np.concatenate([par[i] + rng.normal(0, 2 + k[i] / 4, (k[i], 2)) for i in range(400)])
    r = np.full(len(pts), 4.0)
    iset = build_islands(pts[:, 0], pts[:, 1], None, None, None, 27.5, radius=r)
    bk = bucket_islands(iset, ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES)
    seen = np.zeros(iset.n_islands, int)
    for (B, N), b in bk.buckets.items():
        for row, i in enumerate(b.island_ids):
            seen[i] += 1
            size, n = iset.bbox_size[i], iset.n_src[i]
            assert size <= B and n <= N
            smaller_b = [c for c in ISLAND_BBOX_CLASSES if c < B]
            smaller_n = [c for c in ISLAND_NSRC_CLASSES if c < N]
            assert not smaller_b or size > max(smaller_b)
            assert not smaller_n or n > max(smaller_n)
            assert b.member_mask[row].sum() == n
    for i in bk.giant_ids:
        seen[i] += 1
        assert iset.bbox_size[i] > max(ISLAND_BBOX_CLASSES) or iset.n_src[i] > max(
            ISLAND_NSRC_CLASSES)
    assert np.all(seen == 1)


def test_grid_tiling_owns_each_source_once():
    grids = divide_image_into_grids(2048, 2000, 8, 7, 30)
    rng = np.random.default_rng(1)
    x, y = rng.uniform(0, 2000, 3000), rng.uniform(0, 2048, 3000)
    assign_sources_to_grids(grids, x, y)
    owned = np.concatenate([g.source_indices for g in grids])
    assert np.array_equal(np.sort(owned), np.arange(3000))
    cover = np.zeros((2048, 2000), int)
    for g in grids:
        cover[g.y0_core:g.y1_core, g.x0_core:g.x1_core] += 1
        assert g.x0_pad == max(0, g.x0_core - 30) and g.y1_pad == min(2048, g.y1_core + 30)
        h = g.halo_source_indices
        assert np.all((x[h] >= g.x0_pad) & (x[h] < g.x1_pad) & (y[h] >= g.y0_pad)
                      & (y[h] < g.y1_pad))
    assert np.all(cover == 1)


def test_stamp_tier_covers_twice_the_profile_extent():
    rng = np.random.default_rng(2)
    n = 300
    px, py = rng.uniform(0, 3000, n), rng.uniform(0, 3000, n)
    st = rng.integers(0, 4, n).astype(np.int32)
    sn = np.where(st == 2, 4.0, np.where(st == 3, rng.uniform(0.5, 6, n), 1.0))
    lr = np.log(rng.uniform(1, 12, n))
    iset = build_islands(px, py, st, lr, sn, 27.5, radius=np.full(n, 20.0))
    plans, capped = plan_island_classes(iset, bucket_islands(iset, ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES), st, lr, sn, 55)
    for p in plans:
        for row in range(p.n_islands):
            m = p.midx[row][p.mmask[row]]
            gal = m[st[m] != 0]
            if p.kind == "star":
                assert gal.size == 0 and p.tier == 55
                continue
            nre = np.where(sn[gal] < 1.5, 4.0, np.where(sn[gal] < 3.0, 6.0, 8.0))
            need = 2 * (nre * np.exp(lr[gal])).max()
            ladder = (55, 111, 221)
            want = next((t for t in ladder if need <= t), 221)
            assert p.tier == want


def test_scene_tiles_cover_each_clipped_stamp_once():
    """Scene-subtraction tiles: every stamp pixel inside the image is in
    exactly one tile entry of its stamp, tiles are disjoint and in the image,
    entries of a tile are in stamp order."""
    from uniphot_amd.scene import tile_lists
    rng = np.random.default_rng(0)
    H, W, T, tile = 300, 230, 55, 64
    ix = rng.integers(-60, W + 5, 400)
    iy = rng.integers(-60, H + 5, 400)
    off, win, lst = tile_lists(ix, iy, T, H, W, tile)
    cover = np.zeros((ix.size, H, W), np.int8)
    seen = np.zeros((H, W), np.int32)
    for q in range(win.shape[0]):
        x0, y0, h, w = (int(v) for v in win[q])
        assert 0 <= x0 and x0 + w <= W and 0 <= y0 and y0 + h <= H and h > 0 and w > 0
        seen[y0:y0 + h, x0:x0 + w] += 1
        ent = lst[off[q]:off[q + 1]]
        assert np.all(np.diff(ent) > 0)
        for s in ent:
            ys0, ys1 = max(y0, iy[s]), min(y0 + h, iy[s] + T)
            xs0, xs1 = max(x0, ix[s]), min(x0 + w, ix[s] + T)
            assert ys1 > ys0 and xs1 > xs0
            cover[s, ys0:ys1, xs0:xs1] += 1
    assert seen.max() == 1
    for s in range(ix.size):
        area = max(0, min(H, iy[s] + T) - max(0, iy[s])) * max(0, min(W, ix[s] + T)
                                                                 - max(0, ix[s]))
        assert int(cover[s].sum()) == area and cover[s].max() <= 1


def test_mosaic_shelves_are_disjoint():
    from uniphot_amd.scene import shelf_pack
    rng = np.random.default_rng(1)
    sides = rng.choice([32, 64, 128, 256], 300, p=[0.5, 0.3, 0.15, 0.05])
    mx, my, Hm, Wm = shelf_pack(sides)
    img = np.zeros((Hm, Wm), np.int32)
    for x, y, s in zip(mx, my, sides):
        assert x + s <= Wm and y + s <= Hm
        img[y:y + s, x:x + s] += 1
    assert img.max() == 1
    assert img.sum() / img.size > 0.7      # shelves of sorted sides pack densely


def test_every_module_imports_without_a_gpu():
    """Every package module, the bench and the entry points import on a
    machine with no GPU (the library is loaded lazily)."""
    import importlib
    import pkgutil
    import uniphot_amd
    for m in pkgutil.iter_modules(uniphot_amd.__path__):
        if m.name.startswith("lib"):
            continue  # the HIP library itself (a C ABI loaded with ctypes)
        importlib.import_module(f"uniphot_amd.{m.name}")
    importlib.import_module("bench")
    importlib.import_module("__graft_entry__")


def test_bucketing_a_subset_of_islands_equals_bucketing_them_in_the_full_set():
    """An island set restricted to some islands (as the accuracy and cutout
    paths build) buckets those islands exactly as the full set does."""
    from uniphot_amd.pipeline import sub_island_set
    rng = np.random.default_rng(4)
    pts = rng.uniform(0, 600, (300, 2))
    iset = build_islands(pts[:, 0], pts[:, 1], None, None, None, 27.5,
                         radius=np.full(300, 6.0))
    ids = np.sort(rng.choice(iset.n_islands, iset.n_islands // 3, replace=False))
    full = bucket_islands(iset, ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES)
    part = bucket_islands(sub_island_set(iset, ids), ISLAND_BBOX_CLASSES, ISLAND_NSRC_CLASSES)
    for key, b in part.buckets.items():
        f = full.buckets[key]
        rows = np.searchsorted(f.island_ids, ids[b.island_ids])
        assert np.array_equal(f.island_ids[rows], ids[b.island_ids])
        assert np.array_equal(f.member_idx[rows], b.member_idx)


def test_catalog_source_params_rules():
    """source_model.py:141-215 rules: TYPE strings (case and spaces
    ignored, unknown -> STAR), Re fallback and clip, ab = 1 - ELL clipped
    (NaN ELL -> 0, no column -> 0.8), THETA -> phi, DEV n fixed at 4,
    SERSIC n from the column where finite else 2, clipped to [0.29, 6.3];
    the same result from a DataFrame with lower-case columns."""
    import pandas as pd
    from uniphot_amd.source_model import (catalog_source_params,
                                          ellipse_soft_from_re_ab_phi_np,
                                          ellipse_soft_to_re_ab_phi_np, type_codes)
    assert type_codes([" exp", "DEV", "ser", "Sersic", "psf", "QSO", ""]).tolist() == \
        [1, 2, 3, 3, 0, 0, 0]
    t = np.array(["STAR", "EXP", "DEV", "SERSIC", "SERSIC", "DEV"])
    re = np.array([np.nan, 0.1, 500.0, 2.0, 3.0, 4.0])
    ell = np.array([0.5, np.nan, 0.95, 0.2, 0.0, 0.3])
    th = np.array([30.0, -30.0, np.nan, 0.0, 90.0, -90.0])
    sn = np.array([np.nan, 5.0, 7.0, 9.0, np.nan, 1.0])
    st, lr, e1, e2, n = catalog_source_params(t, re, ell, th, sn)
    assert st.tolist() == [0, 1, 2, 3, 3, 2]
    np.testing.assert_allclose(np.exp(lr), [3.0, 0.3, 100.0, 2.0, 3.0, 4.0], rtol=1e-6)
    np.testing.assert_allclose(n, [1.0, 1.0, 4.0, 6.3, 2.0, 4.0], rtol=1e-7)
    phi = -np.where(np.nan_to_num(th) >= 0, 90.0 - np.nan_to_num(th), -90.0 - np.nan_to_num(th))
    ab = np.clip(1.0 - np.nan_to_num(ell), 0.1, 1.0)
    want = ellipse_soft_from_re_ab_phi_np(np.clip(np.nan_to_num(re, nan=3.0), 0.3, 100), ab, phi)
    for a, b in zip((lr, e1, e2), want):
        np.testing.assert_array_equal(a, b)
    df = pd.DataFrame(dict(type=t, re=re, ell=ell, theta=th, sersic_n=sn))
    for a, b in zip(catalog_source_params(df), (st, lr, e1, e2, n)):
        np.testing.assert_array_equal(a, b)
    # no ELL column: ab = 0.8
    _, _, f1, f2, _ = catalog_source_params(t, re, None, th, sn)
    _, ab8, _ = ellipse_soft_to_re_ab_phi_np(lr, f1, f2)
    np.testing.assert_allclose(ab8, 0.8, rtol=1e-6)


def test_ellipse_soft_round_trip_and_round_case():
    """to(from(re, ab, phi)) returns re, ab and phi mod 180 (float32); a
    round source (ab = 1) has ee1 = ee2 = 0."""
    from uniphot_amd.source_model import (ellipse_soft_from_re_ab_phi_np,
                                          ellipse_soft_to_re_ab_phi_np)
    rng = np.random.default_rng(5)
    re = rng.uniform(0.2, 50, 200)
    ab = rng.uniform(0.15, 1.0, 200)
    phi = rng.uniform(0, 180, 200)
    r2, ab2, p2 = ellipse_soft_to_re_ab_phi_np(*ellipse_soft_from_re_ab_phi_np(re, ab, phi))
    np.testing.assert_allclose(r2, re, rtol=1e-6)
    np.testing.assert_allclose(ab2, ab, rtol=1e-5)
    d = np.abs(((p2 - phi) + 90.0) % 180.0 - 90.0)
    assert d.max() < 1e-3
    _, z1, z2 = ellipse_soft_from_re_ab_phi_np(np.array([2.0]), np.array([1.0]), np.array([33.0]))
    assert z1[0] == 0.0 and z2[0] == 0.0
