"""Serving loop: tiles admitted continuously by HBM budget run on their own
streams with their work overlapping on the device; each tile's result is
the same alone and in the mix."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tiles(cuda):
    from scenes.tiles import PsfBank, make_tile
    bank = PsfBank(n_bands=8)
    rng = np.random.default_rng(11)
    shapes = [(256, 3), (512, 2), (256, 5), (512, 4)]
    reqs = [make_tile(rng, i, 0.0, s, nb, bank) for i, (s, nb) in enumerate(shapes)]
    print("tiles", [(r.side, r.n_bands, r.n_sources) for r in reqs], flush=True)
    return bank, reqs


def _same(a, b):
    for k in ("flux", "flux_err", "x", "y", "logre", "island_nst", "island_conv"):
        assert np.array_equal(getattr(a, k), getattr(b, k), equal_nan=True), k


def test_tiles_alone_equal_in_the_mix(tiles):
    from uniphot_amd.serving import TileServer, latency_summary
    bank, reqs = tiles
    alone = [TileServer(bank, max_inflight=1).serve([r], realtime=False)[0] for r in reqs]
    mix = TileServer(bank, max_inflight=4).serve(reqs, realtime=False)
    print(latency_summary(mix), flush=True)
    for a, m in zip(alone, mix):
        assert a.rid == m.rid
        _same(a.result, m.result)
        assert np.isfinite(m.result.flux).mean() > 0.9


def test_budget_serialises_admission(tiles):
    from uniphot_amd.serving import TileServer
    bank, reqs = tiles
    srv = TileServer(bank, max_inflight=4)
    budget = 1.5 * max(srv.footprint(r) for r in reqs)
    srv.hbm_budget = budget
    served = srv.serve(reqs, realtime=False)
    ref = TileServer(bank, max_inflight=1).serve(reqs, realtime=False)
    for a, b in zip(ref, served):
        _same(a.result, b.result)
    # never two tiles in flight whose footprints exceed the budget together
    ev = sorted([(s.admitted, 1, s.rid) for s in served] + [(s.finished, -1, s.rid) for s in served])
    live = set()
    for t, kind, rid in ev:
        if kind == 1:
            live.add(rid)
            assert sum(srv.footprint(reqs[i]) for i in live) <= budget or len(live) == 1
        else:
            live.discard(rid)
