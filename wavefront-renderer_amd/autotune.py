"""Autotuner for the wavefront hot path: picks the wavefront size (ray slots in flight), the
traversal block shape (resident 128-lane blocks per CU, refill threshold of the persistent
kernels), the sample lanes per pixel and the radix-sort digit width by timing the real workload
on the device.

Search: coordinate descent from the library defaults. Each knob in turn is set to every value of
its candidate list with the others held at their current best, the fastest value is kept, and the
sweep repeats until a round changes nothing (at most `rounds` rounds). Every setting only changes
the order and grouping of work, never which paths are traced, so a tuned render traces the same
rays; sample lanes and the wavefront size change the order of the float additions of a pixel
(images equal up to reassociation), the block shape and digit width do not change the image at
all.

Two objectives are provided: one plain render (`tune_render`, wf_render with the regenerating
schedule) and one batch of jobs on an engine (`tune_engine`, the batched form's engine
capacity, lanes and context tuning). Results can be cached in a JSON file keyed by device,
workload and library tuning space.
"""
from __future__ import annotations

import json
import time
from pathlib import Path
from typing import Callable, Dict, Iterable, List, Optional

# Candidate values of each knob. Wavefront sizes are ray slots per framebuffer pixel.
RENDER_SPACE: Dict[str, List] = {
    "wavefront_per_pixel": [1, 2, 4, 8],
    "sample_lanes": [8, 16, 32, 64],
    "trace_blocks_per_cu": [6, 8, 10, 12, 16],
    "radix_bits": [6, 8, 11],
    "refill_lanes": [32, 48, 64],
}
ENGINE_SPACE: Dict[str, List] = {
    "capacity": [2 << 20, 4 << 20, 8 << 20, 16 << 20, 32 << 20],
    "sample_lanes": [8, 16, 32, 64],
    "trace_blocks_per_cu": [4, 6, 8, 10, 12],
    "radix_bits": [6, 8, 11],
    "refill_lanes": [32, 48, 64],
}
CONTEXT_KNOBS = ("trace_blocks_per_cu", "radix_bits", "refill_lanes")


def coordinate_descent(measure: Callable[[dict], float], space: Dict[str, List], start: dict,
                       rounds: int = 2, log: Optional[Callable[[str], None]] = None) -> dict:
    """Minimises measure(setting) over the product of the candidate lists, one knob at a time.
    Returns {"best": setting, "seconds": its measurement, "trials": [(setting, seconds)]}."""
    best = dict(start)
    for knob, values in space.items():   # a start value outside the space moves to the nearest
        if best.get(knob) not in values:
            cur = best.get(knob)
            numeric = isinstance(cur, (int, float)) and all(isinstance(v, (int, float))
                                                            for v in values)
            best[knob] = min(values, key=lambda v: abs(v - cur)) if numeric else values[0]
    seen: Dict[tuple, float] = {}
    trials = []

    def cost(setting):
        key = tuple(sorted(setting.items()))
        if key not in seen:
            seen[key] = float(measure(dict(setting)))
            trials.append((dict(setting), seen[key]))
            if log:
                log(f"[autotune] {setting} -> {seen[key] * 1e3:.1f} ms")
        return seen[key]

    best_t = cost(best)
    for _ in range(max(1, rounds)):
        changed = False
        for knob, values in space.items():
            for v in values:
                if v == best.get(knob):
                    continue
                cand = dict(best, **{knob: v})
                t = cost(cand)
                if t < best_t:
                    best, best_t, changed = cand, t, True
        if not changed:
            break
    return {"best": best, "seconds": best_t, "trials": trials}


def _ctx_part(setting: dict) -> dict:
    return {k: setting[k] for k in CONTEXT_KNOBS if k in setting}


def _cached(cache: Optional[Path], key: str):
    if cache and Path(cache).exists():
        try:
            return json.loads(Path(cache).read_text()).get(key)
        except (OSError, ValueError):
            return None
    return None


def _store(cache: Optional[Path], key: str, value: dict):
    if not cache:
        return
    p = Path(cache)
    try:
        data = json.loads(p.read_text()) if p.exists() else {}
    except (OSError, ValueError):
        data = {}
    data[key] = value
    p.parent.mkdir(parents=True, exist_ok=True)
    p.write_text(json.dumps(data, indent=1, sort_keys=True))


def tune_render(ctx, scene, params_for: Callable[[int, int], object], pixels: int,
                space: Optional[Dict[str, List]] = None, rounds: int = 2, repeats: int = 2,
                cache: Optional[Path] = None, key: str = "", log=None) -> dict:
    """Tunes one plain render. params_for(sample_lanes, wavefront) returns the wf_render_params
    of the workload with those two fields set. The objective is the best of `repeats` frame
    times after one warm frame (which sizes the cached engine and its graph)."""
    from . import render
    space = dict(space or RENDER_SPACE)
    cache_key = f"render|{key}|{json.dumps(space, sort_keys=True)}"
    hit = _cached(cache, cache_key)
    if hit:
        ctx.set_tuning(**_ctx_part(hit["best"]))
        return hit
    base = ctx.tuning()
    start = {"wavefront_per_pixel": 8, "sample_lanes": 64,
             **{k: base[k] for k in CONTEXT_KNOBS}}

    def measure(s):
        ctx.set_tuning(**_ctx_part(s))
        p = params_for(s["sample_lanes"], s["wavefront_per_pixel"] * pixels)
        render(scene, p, keep_on_device=True)
        ctx.synchronize()
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            render(scene, p, keep_on_device=True)
            ctx.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    try:
        res = coordinate_descent(measure, space, start, rounds, log)
    finally:
        ctx.set_tuning(**base)
    ctx.set_tuning(**_ctx_part(res["best"]))
    out = {"best": res["best"], "seconds": res["seconds"], "trials": len(res["trials"])}
    _store(cache, cache_key, out)
    return out


def tune_engine(ctx, scene, make_jobs: Callable[[], Iterable], pixel_budget: int, max_jobs: int,
                engine_kw: dict, space: Optional[Dict[str, List]] = None, rounds: int = 2,
                cache: Optional[Path] = None, key: str = "", log=None) -> dict:
    """Tunes a batched engine. make_jobs() returns the calibration jobs (wf_job_desc); the
    objective is the wall time of admitting them all and running the engine until every one has
    finished, on an engine created with the candidate capacity and sample lanes (its second run,
    after a warm run that captures the iteration graph)."""
    from . import Engine
    space = dict(space or ENGINE_SPACE)
    cache_key = f"engine|{key}|{json.dumps(space, sort_keys=True)}"
    hit = _cached(cache, cache_key)
    if hit:
        ctx.set_tuning(**_ctx_part(hit["best"]))
        return hit
    base = ctx.tuning()
    start = {"capacity": 8 << 20, "sample_lanes": 64, **{k: base[k] for k in CONTEXT_KNOBS}}

    def run_all(eng):
        ids = [eng.admit(j) for j in make_jobs()]
        left = set(ids)
        while left:
            for jid in eng.run():
                left.discard(jid)
                eng.result(jid, keep_on_device=True)

    def measure(s):
        ctx.set_tuning(**_ctx_part(s))
        eng = Engine(ctx, scene, capacity=s["capacity"], pixel_budget=pixel_budget,
                     max_jobs=max_jobs, sample_lanes=s["sample_lanes"], use_graph=True,
                     **engine_kw)
        try:
            run_all(eng)
            ctx.synchronize()
            t0 = time.perf_counter()
            run_all(eng)
            ctx.synchronize()
            return time.perf_counter() - t0
        finally:
            eng.close()

    try:
        res = coordinate_descent(measure, space, start, rounds, log)
    finally:
        ctx.set_tuning(**base)
    ctx.set_tuning(**_ctx_part(res["best"]))
    out = {"best": res["best"], "seconds": res["seconds"], "trials": len(res["trials"])}
    _store(cache, cache_key, out)
    return out
