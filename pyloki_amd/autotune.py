"""Launch shapes chosen by measurement on the device, cached per device and workload.

The tunable shape today is the tile of the round-based fused FFA kernels (the last
level, or the last two levels, fused with scoring): a block owns 256 output rows, ``rows``
along the outer grid axis (e.g. accelerations) by ``256 / rows`` along the inner one
(frequencies). Taller tiles reuse each source row across more outer indices from L2;
wider ones keep more frequency neighbours together. Which wins depends on the grid and on
the device, so ``tune`` times every candidate on the real workload (HIP events around a
full step, the median of a few repetitions) and keeps the fastest. The tile never
changes a result, only speed.

Cache: a JSON file (``$PYLOKI_AMD_TUNE_CACHE``, default
``~/.cache/pyloki_amd/autotune.json``) keyed by device name and a caller-given workload
signature; ``apply`` sets a cached choice without measuring.
"""

from __future__ import annotations

import json
import os
from pathlib import Path

import numpy as np
import torch

from pyloki_amd import _lib

TILE_ROWS = (1, 2, 4, 8, 16, 32, 64)


def cache_path() -> Path:
    return Path(os.environ.get("PYLOKI_AMD_TUNE_CACHE",
                               "~/.cache/pyloki_amd/autotune.json")).expanduser()


def _device() -> str:
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{p.name}:{getattr(p, 'gcnArchName', '')}"


def _load() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _store(db: dict) -> None:
    path = cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(db, indent=1, sort_keys=True))
        tmp.replace(path)
    except OSError:
        pass  # a read-only home: the choice still applies to this process


def set_tile_rows(rows: int) -> None:
    _lib.check(_lib.load().pyl_ffa_set_tile_rows(int(rows)))


def get_tile_rows() -> int:
    return int(_lib.load().pyl_ffa_get_tile_rows())


def _time(step, reps: int, warmup: int) -> float:
    for _ in range(warmup):
        step()
    ms = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        step()
        b.record()
        b.synchronize()
        ms.append(a.elapsed_time(b))
    return float(np.median(ms))


def tune_tile_rows(signature: str, step, candidates=TILE_ROWS, reps: int = 3,
                   warmup: int = 1, use_cache: bool = True) -> dict:
    """Time ``step()`` (one full unit of the workload on the current stream) with every
    candidate tile, set the fastest and cache it under ``signature``. Returns
    {"rows": best, "ms": {rows: median ms}, "cached": bool}."""
    db = _load()
    key = f"ffa_tile_rows/{signature}"
    dev = _device()
    hit = db.get(dev, {}).get(key)
    if use_cache and hit is not None:
        set_tile_rows(hit["rows"])
        return {**hit, "cached": True}
    timings = {}
    for rows in candidates:
        set_tile_rows(rows)
        timings[int(rows)] = _time(step, reps, warmup)
    best = min(timings, key=timings.get)
    set_tile_rows(best)
    entry = {"rows": int(best), "ms": {str(k): v for k, v in timings.items()}}
    db.setdefault(dev, {})[key] = entry
    _store(db)
    return {**entry, "cached": False}


def apply(signature: str) -> int | None:
    """Set the cached tile for ``signature`` on this device, if any; returns it."""
    hit = _load().get(_device(), {}).get(f"ffa_tile_rows/{signature}")
    if hit is None:
        return None
    set_tile_rows(hit["rows"])
    return int(hit["rows"])
