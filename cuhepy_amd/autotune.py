"""Measurement-driven tuning of launch shapes (lanes per bignum, window width,
persistent waves per CU).

Every candidate is run on the device and timed with a device synchronisation
around it; the fastest is kept. Results are cached in memory and, when
``CHX_TUNE_CACHE`` names a file, in that JSON file, keyed by the device name
and a signature of the workload shape.
"""

from __future__ import annotations

import json
import os
import time
from typing import Callable, Iterable, Optional

_MEM: dict[str, dict] = {}


def _cache_path() -> Optional[str]:
    return os.environ.get("CHX_TUNE_CACHE") or None


def _load() -> dict:
    p = _cache_path()
    if p and os.path.exists(p):
        try:
            with open(p) as f:
                return json.load(f)
        except (OSError, ValueError):
            return {}
    return {}


def _store(key: str, value: dict) -> None:
    _MEM[key] = value
    p = _cache_path()
    if p:
        data = _load()
        data[key] = value
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1)
        os.replace(tmp, p)


def lookup(key: str) -> Optional[dict]:
    if key in _MEM:
        return _MEM[key]
    v = _load().get(key)
    if v is not None:
        _MEM[key] = v
    return v


def time_it(fn: Callable[[], None], sync: Callable[[], None], reps: int = 2) -> float:
    """Median wall time of `reps` runs of fn (after one warm run)."""
    fn()
    sync()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        sync()
        times.append(time.perf_counter() - t0)
    times.sort()
    return times[len(times) // 2]


def search(candidates: Iterable[dict], run: Callable[[dict], Callable[[], None]], sync: Callable[[], None],
           reps: int = 2) -> tuple[dict, list[tuple[dict, float]]]:
    """Time each candidate (``run(cand)`` returns a no-argument runner)."""
    results = []
    for cand in candidates:
        try:
            t = time_it(run(cand), sync, reps)
        except (ValueError, RuntimeError):
            continue  # a shape that is not built or not valid for this workload
        results.append((cand, t))
    if not results:
        raise RuntimeError("no tuning candidate ran")
    best = min(results, key=lambda r: r[1])[0]
    return best, results


def tune_ragged(ctx, workload, device_name: str, variants=(0, 1, 2, 3), windows=(4, 5, 6), waves=(4, 8),
                reps: int = 2, force: bool = False) -> tuple[dict, list]:
    """Coordinate search for the ragged persistent launch: variant (lanes per
    bignum per width, one kernel per width or one mixed kernel), then window
    width, then persistent waves per CU."""
    from .engine import RaggedPlan

    widths = sorted({(s.limbs, s.count // 1024) for s in workload.subs})
    key = f"ragged|{device_name}|{len(workload.subs)}|{workload.total_items()}|{hash(tuple(widths)) & 0xFFFFFFFF}"
    if not force:
        hit = lookup(key)
        if hit is not None:
            return hit, []
    plans = {}

    def run(c):
        k = (c["variant"], c["window"], c["waves"])
        if k not in plans:
            plans[k] = RaggedPlan(ctx, workload.desc, c["variant"], c["window"], c["waves"])
        p = plans[k]
        return lambda: p.run(graph=True)

    log = []
    best = {"variant": variants[0], "window": 5, "waves": 8}
    for dim, values in (("variant", variants), ("window", windows), ("waves", waves)):
        cands = [dict(best, **{dim: v}) for v in values]
        best, res = search(cands, run, ctx.sync, reps)
        log.extend(res)
    for p in plans.values():
        p.close()
    _store(key, best)
    return best, log


def tune_lazy(ctx, op: int, limbs: int, run: Callable[[], None], tpis: Iterable[int], windows: Iterable[int],
              reps: int = 2) -> tuple[dict, list]:
    """Pick lanes per bignum and window for one exponentiation op by timing
    `run` under each (tpi, window) set through the context tuning knob."""
    def runner(c):
        ctx.set_tuning(op, limbs, c["tpi"], c["window"])
        return run

    cands = [{"tpi": t, "window": w} for t in tpis for w in windows]
    best, res = search(cands, runner, ctx.sync, reps)
    ctx.set_tuning(op, limbs, best["tpi"], best["window"])
    return best, res
