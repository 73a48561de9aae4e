"""Serving loop for field tiles: continuous admission by HBM budget, each
admitted tile fitted on its own HIP stream so the work of tiles in flight
overlaps on the device, PSF coefficients resident across requests.

A request (``scenes.tiles.TileRequest``) carries its images on the device,
its band PSFs (frames of the server's resident bank) and its catalog.  The
server admits arrived requests in arrival order while the sum of their
estimated device footprints stays within ``hbm_budget`` and fewer than
``max_inflight`` are running; each runs ``pipeline.fit_field`` (two linear
passes and the band-split island fit) on a private stream.  A tile's result
does not depend on what else is in flight: every kernel's arithmetic per
work item is fixed and items never share data across tiles.
"""
from __future__ import annotations

import threading
import time
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass

import numpy as np

from .pipeline import DeviceField, FitConfig, fit_field
from .psf_polynomial import psf_device_arrays


@dataclass
class Served:
    rid: int
    arrival: float
    admitted: float
    finished: float
    n_sources: int
    n_bands: int
    side: int
    result: object            # pipeline.FitResult

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class TileServer:
    def __init__(self, bank, tiers=(55, 111, 221), hbm_budget=200e9, max_inflight=4,
                 device="cuda", cfg: FitConfig | None = None):
        import torch
        self.bank = bank
        self.hbm_budget = float(hbm_budget)
        self.max_inflight = int(max_inflight)
        self.device = device
        # tiles in flight share the device: a smaller island workspace each
        self.cfg = cfg or FitConfig(island_workspace=16 * 2 ** 30)
        # resident PSF coefficients of every bank band at every stamp tier
        # (the frame-dependent geometry is per tile)
        self._coefs = {}
        for T in tiers:
            c, _ = psf_device_arrays(bank.polys, T)
            self._coefs[T] = torch.from_numpy(c).to(device)
        self._band_of = {id(p.coefs): b for b, p in enumerate(bank.polys)}
        self._lock = threading.Lock()

    def footprint(self, req) -> float:
        """Estimated device bytes while the tile is in flight: images and
        padded stacks, grid-band linear-solve buffers, island work buffers,
        the island workspace and, with scene subtraction, the island cutout
        mosaic (at most one 64 x 64 window of data and weights per source and
        band)."""
        px = float(req.n_bands) * req.side * req.side
        stacks = px * 4.0 * 2 * 1.6
        linear = 4.0 * req.n_sources * req.n_bands * 111.0 ** 2
        island = 2e6 * req.n_sources * max(req.n_bands, 1) / 20.0 + (2 << 30)
        mosaic = 8.0 * req.n_bands * req.n_sources * 64.0 ** 2 if self.cfg.scene_subtract else 0.0
        # island workspace: at most 1024 workgroups of up to ~16 MB each
        work = float(min(self.cfg.island_workspace, 34e6 * max(1, req.n_sources * req.n_bands)))
        return px * 9.0 + stacks + linear + island + mosaic + work

    def _psf_dev(self, req):
        import torch
        bands = [self._band_of[id(p.coefs)] for p in req.psfs]
        idx = torch.as_tensor(bands, device=self.device)
        out = {}
        for T, c in self._coefs.items():
            geo = np.array([[p.x_origin, p.x_scale, p.y_origin, p.y_scale] for p in req.psfs],
                           np.float32)
            out[T] = (c.index_select(0, idx).contiguous(), torch.from_numpy(geo).to(self.device))
        return out

    def _run(self, req, t0):
        import torch
        stream = torch.cuda.Stream(device=self.device)
        t_adm = time.perf_counter() - t0
        with torch.cuda.stream(stream):
            fld = DeviceField.from_device(req.data, req.invvar, req.good, req.psfs,
                                          psf_dev=self._psf_dev(req))
            g = max(1, req.side // 256)
            cfg = FitConfig(**{**self.cfg.__dict__, "n_rows": g, "n_cols": g})
            res = fit_field(fld, req.x, req.y, req.stype, req.logre, req.ee1, req.ee2,
                            req.sersic_n, cfg, sync=stream.synchronize)
            stream.synchronize()
        return Served(req.rid, req.arrival, t_adm, time.perf_counter() - t0, req.n_sources,
                      req.n_bands, req.side, res)

    def serve(self, requests, realtime=True):
        """Serve ``requests`` (arrival times relative to the call) and return
        one ``Served`` per request, in request order.  With ``realtime`` a
        request is admitted no earlier than its arrival time."""
        pending = sorted(requests, key=lambda r: r.arrival)
        done = {}
        inflight = {}
        used = 0.0
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=self.max_inflight) as pool:
            while pending or inflight:
                now = time.perf_counter() - t0
                while (pending and (not realtime or pending[0].arrival <= now)
                       and len(inflight) < self.max_inflight):
                    need = self.footprint(pending[0])
                    if inflight and used + need > self.hbm_budget:
                        break      # wait for memory to free up
                    req = pending.pop(0)
                    used += need
                    inflight[pool.submit(self._run, req, t0)] = need
                if not inflight:
                    time.sleep(max(0.0, min(0.05, pending[0].arrival - now)))
                    continue
                timeout = None
                if pending and realtime:
                    timeout = max(1e-3, pending[0].arrival - (time.perf_counter() - t0))
                fin, _ = wait(list(inflight), timeout=timeout, return_when=FIRST_COMPLETED)
                for f in fin:
                    used -= inflight.pop(f)
                    s = f.result()
                    done[s.rid] = s
        return [done[r.rid] for r in requests]


def latency_summary(served) -> dict:
    lat = np.array([s.latency for s in served])
    span = max(s.finished for s in served) - min(s.arrival for s in served)
    n_src = sum(s.n_sources for s in served)
    return dict(requests=len(served), sources=n_src, makespan_s=span,
                sources_per_s=n_src / max(span, 1e-9), p50_latency_s=float(np.percentile(lat, 50)),
                p99_latency_s=float(np.percentile(lat, 99)), max_latency_s=float(lat.max()))
