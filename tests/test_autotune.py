"""Autotuner (differentiablepointrendering_amd/autotune.py) against a stand-in
renderer whose device time is a known function of the tuning: coordinate
descent must end on the fastest value of each parameter, skip batch sizes far
above the workload, and leave the chosen tuning set on the renderer."""
from types import SimpleNamespace

from differentiablepointrendering_amd import autotune


class StandIn:
    def __init__(self, cost):
        self.cost = cost
        self.tuning = None
        self.calls = 0

    def set_batch_tuning(self, slots_per_batch=0, entries_per_slot=0, graph_rounds=0, grid_fraction=0.0):
        self.tuning = dict(slots_per_batch=slots_per_batch, entries_per_slot=entries_per_slot,
                           graph_rounds=graph_rounds, grid_fraction=grid_fraction)

    def render_views(self, cameras, spp, seeds):
        self.calls += 1

    def render_views_backward(self, cameras, spp, dloss):
        pass

    def stats(self):
        ms = self.cost(self.tuning)
        return {"forward_ms": ms, "backward_ms": 0.5 * ms}


def _cams(n, w=640, h=480):
    return [SimpleNamespace(width=w, height=h) for _ in range(n)]


def test_coordinate_descent_finds_the_separable_minimum():
    best = {"slots_per_batch": 8 << 20, "entries_per_slot": 32, "graph_rounds": 0, "grid_fraction": 0.75}

    def cost(t):
        return (10.0 + abs(t["slots_per_batch"] - best["slots_per_batch"]) / (1 << 20)
                + abs(t["entries_per_slot"] - best["entries_per_slot"]) + 3.0 * abs(t["graph_rounds"])
                + 20.0 * abs(t["grid_fraction"] - best["grid_fraction"]))

    r = StandIn(cost)
    cams = _cams(8)
    out = autotune.tune(r, cams, [16] * 8, list(range(8)), dloss=[None] * 8)
    assert out["tuning"] == best
    assert r.tuning == best
    assert out["ms"] == round(1.5 * 10.0, 3)
    # every measured point is in the table, each once
    keys = [tuple(sorted((k, v) for k, v in row.items() if k != "ms")) for row in out["table"]]
    assert len(keys) == len(set(keys))


def test_batch_sizes_far_above_the_workload_are_skipped():
    seen = []

    def cost(t):
        seen.append(t["slots_per_batch"])
        return 1.0

    r = StandIn(cost)
    # 2 views of 64x64 at 4 spp: 32,768 slots; only the smallest candidate at
    # or above twice that is kept
    autotune.tune(r, _cams(2, 64, 64), [4, 4], [0, 1])
    assert max(seen) == autotune.CANDIDATES["slots_per_batch"][0]
