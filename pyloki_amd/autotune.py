"""Launch shapes chosen by measurement on the device, cached per device and workload.

Tunable shapes (``KNOBS``), none of which changes a result, only speed:

- ``ffa_tile_rows``: the tile of the round-based fused FFA kernels (the last level, or
  the last two levels, fused with scoring): a block owns 256 output rows, ``rows`` along
  the outer grid axis (e.g. accelerations) by ``256 / rows`` along the inner one
  (frequencies). Taller tiles reuse each source row across more outer indices from L2;
  wider ones keep more frequency neighbours together.
- ``bf_segments_per_block``: how many segments a brute-fold block stages in LDS (0: as
  many as fit a 64 KB budget; 1, 2 or 4 up to the 160 KB a gfx950 workgroup holds). More
  segments amortise each block's bin lists over more samples; fewer give more blocks.

Which wins depends on the grid and on the device, so ``tune`` times every candidate on
the real workload (HIP events around a full step, the median of a few repetitions) and
keeps the fastest; ``tune_all`` tunes the knobs one after the other.

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


def set_bf_segments(segs: int) -> None:
    _lib.check(_lib.load().pyl_brutefold_set_segments_per_block(int(segs)))


# knob -> (setter, candidates, default)
KNOBS = {
    "ffa_tile_rows": (set_tile_rows, TILE_ROWS, 16),
    "bf_segments_per_block": (set_bf_segments, (0, 1, 2, 4), 0),
}


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


def tune(knob: str, signature: str, step, candidates=None, reps: int = 3, warmup: int = 1,
         use_cache: bool = True) -> dict:
    """Time ``step()`` (one full unit of the workload on the current stream) with every
    candidate value of ``knob``, set the fastest and cache it under ``signature``. Returns
    {"value": best, "ms": {value: median ms}, "cached": bool}."""
    setter, default_cands, _ = KNOBS[knob]
    candidates = default_cands if candidates is None else candidates
    db = _load()
    key = f"{knob}/{signature}"
    dev = _device()
    hit = db.get(dev, {}).get(key)
    if use_cache and hit is not None:
        setter(hit["value"])
        return {**hit, "cached": True}
    timings = {}
    for v in candidates:
        setter(v)
        timings[int(v)] = _time(step, reps, warmup)
    best = min(timings, key=timings.get)
    setter(best)
    entry = {"value": int(best), "ms": {str(k): v for k, v in timings.items()}}
    db.setdefault(dev, {})[key] = entry
    _store(db)
    return {**entry, "cached": False}


def tune_all(signature: str, step, knobs=("ffa_tile_rows", "bf_segments_per_block"),
             **kw) -> dict:
    """``tune`` every knob in turn (each with the earlier knobs at their best)."""
    return {k: tune(k, signature, step, **kw) for k in knobs}


def tune_tile_rows(signature: str, step, candidates=TILE_ROWS, reps: int = 3,
                   warmup: int = 1, use_cache: bool = True) -> dict:
    """``tune`` of the fused FFA tile: {"rows": best, "ms": ..., "cached": ...}."""
    r = tune("ffa_tile_rows", signature, step, candidates, reps, warmup, use_cache)
    return {"rows": r["value"], "ms": r["ms"], "cached": r["cached"]}


def apply(signature: str, knob: str = "ffa_tile_rows") -> int | None:
    """Set the cached value of ``knob`` for ``signature`` on this device, if any."""
    hit = _load().get(_device(), {}).get(f"{knob}/{signature}")
    if hit is None:
        return None
    value = hit["value"] if "value" in hit else hit["rows"]
    KNOBS[knob][0](value)
    return int(value)
