"""Per-size-class tuning of the batched pipeline.

For a batch of systems of one size class (systems of about the same particle count), the autotuner times the
full batched solve (the captured-graph pipeline by default) for candidate launch configurations and keeps the
fastest. A configuration has five dimensions:

* group_size: targets per wavefront (64 or 32);
* waves_per_block: the traversal's block size (2, 4 or 8 wavefronts; every size keeps 6 waves per SIMD);
* waves_per_cu: traversal waves launched per CU (blocks per CU = waves_per_cu / waves_per_block);
* group_ranges: contiguous group ranges with their own work counters (8: one per XCD, 1: one global counter);
* lifo: LIFO entries per wavefront held in LDS (0: full depth, 22 x 64 entries, which never overflows; a
  shallower LIFO uses less LDS per block, and the groups that overflow it are traversed again at full depth).

The search is one pass of coordinate descent from the default configuration: each dimension in turn takes
the value that is fastest with the others held at the best so far. Results are cached per size class (log2 of
the median system size) and can be saved as JSON. An explicit candidate list is searched exhaustively instead;
its entries are dicts or tuples (group_size, blocks_per_cu, group_ranges[, lifo[, waves_per_block]]).
"""

from __future__ import annotations

import json
import math
import time

import numpy as np

from . import Batch, Context, default_params

DEFAULT = {"group_size": 64, "waves_per_block": 4, "waves_per_cu": 24, "group_ranges": 8, "lifo": 0}
DIMENSIONS = {
    "group_size": (64, 32),
    "waves_per_block": (4, 2, 8),
    "waves_per_cu": (24, 16, 32),
    "group_ranges": (8, 1),
    "lifo": (0, 8 * 64),
}


def size_class(systems) -> int:
    return int(round(math.log2(max(1.0, float(np.median([s.n for s in systems]))))))


def as_config(cfg) -> dict:
    """a configuration dict from a dict or a tuple (group_size, blocks_per_cu, group_ranges[, lifo[, waves]])"""
    if isinstance(cfg, dict):
        return {**DEFAULT, **cfg}
    gs, blocks, ranges = cfg[:3]
    lifo = cfg[3] if len(cfg) > 3 else 0
    waves = cfg[4] if len(cfg) > 4 else 4
    return {"group_size": gs, "waves_per_block": waves, "waves_per_cu": blocks * waves if blocks else 24,
            "group_ranges": ranges, "lifo": lifo}


def _key(cfg: dict):
    return tuple(cfg[k] for k in DIMENSIONS)


def time_solve(batch: Batch, prm, reps: int = 2, graph: bool = True) -> float:
    batch.solve(prm, graph=graph)  # warm-up (allocations, graph capture, caches)
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.solve(prm, graph=graph)
    return (time.perf_counter() - t0) / reps


class Autotuner:
    def __init__(self, ctx: Context, candidates=None, reps: int = 2, graph: bool = True):
        self.ctx = ctx
        self.candidates = None if candidates is None else [as_config(c) for c in candidates]
        self.reps = reps
        self.graph = graph
        self.table = {}

    def apply(self, cfg):
        """set the traversal launch shape of the context; returns the params for the group size"""
        c = as_config(cfg)
        waves = int(c["waves_per_block"])
        blocks = max(1, int(c["waves_per_cu"]) // waves)
        self.ctx.call("ryo_ctx_set_traversal_shape", int(c["lifo"]), blocks, int(c["group_ranges"]))
        self.ctx.call("ryo_ctx_set_traversal_waves", waves)
        return default_params(group_size=int(c["group_size"]))

    def reset(self):
        self.ctx.call("ryo_ctx_set_traversal_shape", 0, 0, 0)
        self.ctx.call("ryo_ctx_set_traversal_waves", 0)

    def tune(self, systems, batch: Batch | None = None):
        """time the candidates on this batch; returns (best config, [(config, seconds), ...])"""
        own = batch is None
        if own:
            batch = Batch(self.ctx, sum(s.n for s in systems), len(systems))
            batch.load(systems)
        seen = {}

        def measure(cfg):
            k = _key(cfg)
            if k not in seen:
                prm = self.apply(cfg)
                seen[k] = (dict(cfg), time_solve(batch, prm, self.reps, self.graph))
            return seen[k][1]

        try:
            if self.candidates is not None:
                for cfg in self.candidates:
                    measure(cfg)
            else:
                best = dict(DEFAULT)
                for dim, values in DIMENSIONS.items():
                    scores = {v: measure({**best, dim: v}) for v in values}
                    best[dim] = min(scores, key=scores.get)
        finally:
            self.reset()
            if own:
                batch.close()
        timings = list(seen.values())
        best = min(timings, key=lambda t: t[1])[0]
        self.table[size_class(systems)] = best
        return best, timings

    def config_for(self, systems) -> dict:
        """the tuned configuration of the systems' size class (nearest tuned class), or the defaults"""
        if not self.table:
            return dict(DEFAULT)
        c = size_class(systems)
        k = min(self.table, key=lambda t: abs(t - c))
        return self.table[k]

    def save(self, path):
        with open(path, "w") as f:
            json.dump({str(k): v for k, v in self.table.items()}, f, indent=1)

    def load(self, path):
        with open(path) as f:
            self.table = {int(k): as_config(v) for k, v in json.load(f).items()}
