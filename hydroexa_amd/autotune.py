"""Autotuner for the fused advance's launch shape.

Candidates are (rows marched per wavefront, minimum waves per SIMD) for a
single level and additionally the schedule (one wavefront per strip, or
persistent wavefronts on a work queue) for an ensemble.  Each candidate is
timed with HIP events over a few real steps of the caller's workload; the
fastest is applied through hx_set_tuning / hx_ensemble_set_tuning.  Results
are cached per device name and workload signature in a JSON file
(``HYDROEXA_TUNE_CACHE``, default ~/.cache/hydroexa_amd/tuning.json).
"""
from __future__ import annotations

import json
import os
from typing import Dict, Iterable, Optional, Tuple

import torch

from .solver import Level, SolverVariant, set_tuning

ROWS = (8, 16, 32, 64)
OCCS = (1, 3)


def _cache_path() -> str:
    return os.environ.get("HYDROEXA_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "hydroexa_amd", "tuning.json"))


def _load_cache() -> Dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(cache: Dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, path)
    except OSError:
        pass


def _time(fn, reps: int) -> float:
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn()                                   # warm-up (code load, caches)
    torch.cuda.synchronize()
    e0.record(st)
    for _ in range(reps):
        fn()
    e1.record(st)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_level(level: Level, model: str, cfl: float = 0.45, steps: int = 6,
               rows: Iterable[int] = ROWS, occs: Iterable[int] = OCCS,
               use_cache: bool = True) -> Tuple[Tuple[int, int], Dict]:
    """Pick (rows, occupancy) for this level's shape and dtype.  The level's
    state is advanced by the trial steps; pass a scratch copy to keep it."""
    key = f"{torch.cuda.get_device_name()}|level|{model}|{level.dtype}|{level.ny}x{level.nx}"
    cache = _load_cache() if use_cache else {}
    if key in cache:
        best = tuple(cache[key]["best"])
        set_tuning(model, level.dtype, *best)
        return best, cache[key]["times"]
    sv = SolverVariant(model)
    times = {}
    for r in rows:
        for o in occs:
            set_tuning(model, level.dtype, r, o)
            times[f"{r},{o}"] = _time(lambda: sv.run(level, cfl, steps), 2) / steps
    best = tuple(int(v) for v in min(times, key=times.get).split(","))
    set_tuning(model, level.dtype, *best)
    if use_cache:
        cache[key] = {"best": list(best), "times": times}
        _store_cache(cache)
    return best, times


def tune_ensemble(ens, cfl: float = 0.45, steps: int = 6, rows: Iterable[int] = ROWS,
                  occs: Iterable[int] = OCCS, schedules: Iterable[bool] = (False, True),
                  use_cache: bool = True,
                  stream: Optional[torch.cuda.Stream] = None) -> Tuple[Tuple[int, int, bool], Dict]:
    """Pick (rows, occupancy, persistent) for a batched ensemble; trial steps
    advance the ensemble's state."""
    sig = ",".join(f"{lv.ny}x{lv.nx}" for lv in ens.levels)
    key = (f"{torch.cuda.get_device_name()}|ensemble|{ens.model}|{ens.dtype}|"
           f"{len(ens.levels)}|{hash(sig) & 0xffffffff:08x}")
    cache = _load_cache() if use_cache else {}
    if key in cache:
        best = tuple(cache[key]["best"])
        ens.set_tuning(best[0], best[1], bool(best[2]))
        return best, cache[key]["times"]
    st = stream or torch.cuda.Stream()
    times = {}
    for r in rows:
        for o in occs:
            for p in schedules:
                ens.set_tuning(r, o, p)
                with torch.cuda.stream(st):
                    times[f"{r},{o},{int(p)}"] = _time(lambda: ens.run(cfl, steps, graph=True, stream=st), 2) / steps
    b = min(times, key=times.get).split(",")
    best = (int(b[0]), int(b[1]), bool(int(b[2])))
    ens.set_tuning(*best)
    if use_cache:
        cache[key] = {"best": [best[0], best[1], int(best[2])], "times": times}
        _store_cache(cache)
    return best, times
