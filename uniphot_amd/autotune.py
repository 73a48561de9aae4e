"""Launch-shape autotuner of the island fit.

The band-split island sequence runs every evaluation and pre-render launch
as a persistent work-queue grid whose size (workgroups, each holding its
own slice of the island workspace) trades residency against the tail of the
largest (island, band) items.  The best size depends on the kernels, the
device and the workload's mix of island classes, so it is measured: each
candidate grid runs the workload's own launch sequence, the fastest wins,
and the choice is cached on disk keyed by a hash of the kernel sources, the
device and a coarse workload signature.  The reference's counterpart is its
measured chunk-width ladder (island_solver.py:2257-2685, cache key
:1928-1973), which sizes JAX chunks by memory rather than by time.

Caches: ``uniphot_amd/data/autotune.json`` (shipped, read only) and a user
cache (``$UPH_AUTOTUNE_CACHE``, else ``~/.cache/uniphot_amd/autotune.json``)
where new results are written; the user cache wins.
"""
from __future__ import annotations

import hashlib
import sys
import json
import math
import os
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
SHIPPED = HERE / "data" / "autotune.json"
# persistent workgroups per launch: 1-8 per CU of a 256-CU device
BLOCK_CANDIDATES = (256, 512, 768, 1024, 1536, 2048)

_hash = None
_log = []


# the sources the island launch sequences are built from
ISLAND_SOURCES = ("island.hip", "prerender.hip", "prerender64.hip", "uph_device.hpp",
                  "uph_galaxy.hpp", "uph_island.hpp", "uph_prerender.hpp")


def kernel_hash() -> str:
    """Hash of the HIP sources the tuned launches run (16 hex digits)."""
    global _hash
    if _hash is None:
        h = hashlib.sha256()
        for name in ISLAND_SOURCES:
            h.update(name.encode())
            h.update((HERE / "csrc" / name).read_bytes())
        _hash = h.hexdigest()[:16]
    return _hash


def device_key(device=None) -> str:
    import torch
    dev = torch.device(device or "cuda")
    pr = torch.cuda.get_device_properties(dev)
    arch = getattr(pr, "gcnArchName", "").split(":")[0]
    return f"{arch or pr.name}/{pr.multi_processor_count}cu"


def user_cache_path() -> Path:
    env = os.environ.get("UPH_AUTOTUNE_CACHE")
    return Path(env) if env else Path.home() / ".cache" / "uniphot_amd" / "autotune.json"


def _read(path: Path) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def lookup(key: str):
    """Cached entry {"choice": ..., "ms": {...}} for ``key``, or None."""
    hit = _read(user_cache_path()).get(key)
    return hit if hit is not None else _read(SHIPPED).get(key)


def store(key: str, entry: dict) -> None:
    path = user_cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        data = _read(path)
        data[key] = entry
        tmp = path.with_suffix(".tmp")
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: the choice still holds for this process


def workload_key(precision: str, n_bands: int, tiers, mixed: bool, n_islands: int,
                 device=None) -> str:
    """Cache key: kernel sources, device, precision and a coarse workload
    signature (bands, largest stamp tier, galaxies present, islands to the
    nearest power of two)."""
    b = 1 << max(0, int(round(math.log2(max(n_islands, 1)))))
    return (f"island_blocks|{kernel_hash()}|{device_key(device)}|{precision}|nb{n_bands}|"
            f"T{max(tiers)}|{'mixed' if mixed else 'star'}|n~{b}")


def choose(key: str, run, candidates=BLOCK_CANDIDATES, reps: int = 1) -> int:
    """The cached choice for ``key``, else time ``run(c)`` (seconds of the
    workload at candidate c, synchronised) for each candidate after one
    warm-up run, cache the fastest and return it."""
    hit = lookup(key)
    if hit is not None:
        _log.append(dict(key=key, choice=hit["choice"], cached=True))
        return int(hit["choice"])
    run(candidates[len(candidates) // 2])
    ms = {}
    for c in candidates:
        ms[int(c)] = round(1e3 * min(run(c) for _ in range(max(1, reps))), 3)
        # one line per measured candidate: a long tuning run shows its progress
        print(f"[autotune] {key}: {int(c)} workgroups {ms[int(c)]} ms", file=sys.stderr,
              flush=True)
    best = min(ms, key=ms.get)
    entry = dict(choice=int(best), ms=ms, tuned_at=time.strftime("%Y-%m-%d %H:%M:%S"))
    store(key, entry)
    _log.append(dict(key=key, choice=int(best), cached=False, ms=ms))
    return int(best)


def log() -> list:
    """Choices made in this process (for the benchmark record)."""
    return list(_log)
