"""Autotuner of the batched view launches (Renderer.render_views and
render_views_backward).

The tuning (pale_batch_tuning) sets only the launch shape -- pixel slots per
batch (the queue of one persistent launch and one adjoint wavefront),
camera-gather entries per slot and round, rounds captured in the HIP graph,
and the size of the persistent layer-walk grid -- never the results (the
batched images equal the single-view ones bit for bit whatever the tuning).
The tuner measures device time (the renderer's own HIP events) of the given
workload, or of a sample of it, and walks the parameters one at a time from
the defaults (coordinate descent), keeping the fastest value of each.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional, Sequence

CANDIDATES: Dict[str, Sequence] = {
    "slots_per_batch": (1 << 20, 2 << 20, 4 << 20, 8 << 20, 16 << 20, 32 << 20, 64 << 20),
    "entries_per_slot": (16, 24, 32),
    "graph_rounds": (0, 1, 2),
    "grid_fraction": (1.0, 0.75),
}
DEFAULT = {"slots_per_batch": 16 << 20, "entries_per_slot": 24, "graph_rounds": 1, "grid_fraction": 1.0}


def measure(renderer, tuning: dict, cameras, spp, seeds, dloss=None, repeats: int = 1) -> float:
    """Device ms of render_views (+ render_views_backward when dloss is
    given) under `tuning`; the best of `repeats` runs after one warm-up."""
    renderer.set_batch_tuning(**tuning)
    best = float("inf")
    for k in range(repeats + 1):
        renderer.render_views(cameras, spp, seeds)
        ms = renderer.stats()["forward_ms"]
        if dloss is not None:
            renderer.render_views_backward(cameras, spp, dloss)
            ms += renderer.stats()["backward_ms"]
        if k > 0:
            best = min(best, ms)
    return best


def tune(renderer, cameras, spp, seeds, dloss=None, candidates: Optional[Dict[str, Sequence]] = None,
         repeats: int = 1, total_slots: Optional[int] = None) -> dict:
    """Coordinate descent over `candidates` (default CANDIDATES), starting
    from DEFAULT.  Batch sizes far above the workload's own slot count are
    skipped (they all give one batch).  Sets the best tuning on the renderer
    and returns {"tuning", "ms", "table", "seconds"}."""
    cand = dict(CANDIDATES if candidates is None else candidates)
    if total_slots is None:
        total_slots = sum(c.width * c.height * s for c, s in zip(cameras, spp))
    sizes = [s for s in cand["slots_per_batch"] if s < 2 * total_slots]
    larger = [s for s in cand["slots_per_batch"] if s >= 2 * total_slots]
    cand["slots_per_batch"] = sizes + larger[:1]
    t0 = time.time()
    best = dict(DEFAULT)
    if best["slots_per_batch"] not in cand["slots_per_batch"]:
        best["slots_per_batch"] = cand["slots_per_batch"][-1]
    table: List[dict] = []
    seen: Dict[tuple, float] = {}

    def run(t):
        key = tuple(sorted(t.items()))
        if key not in seen:
            seen[key] = measure(renderer, t, cameras, spp, seeds, dloss, repeats)
            table.append(dict(t, ms=round(seen[key], 3)))
        return seen[key]

    best_ms = run(best)
    for name, values in cand.items():
        for v in values:
            t = dict(best, **{name: v})
            ms = run(t)
            if ms < best_ms:
                best, best_ms = t, ms
    renderer.set_batch_tuning(**best)
    return {"tuning": best, "ms": round(best_ms, 3), "table": table, "seconds": round(time.time() - t0, 2)}
