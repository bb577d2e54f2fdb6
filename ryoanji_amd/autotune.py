"""Per-size-class tuning of the batched pipeline.

For a batch of systems of one size class (systems of about the same particle count), the autotuner times the
full batched solve (the captured-graph pipeline by default) for each candidate launch configuration --
target-group size (32 or 64 targets per wavefront), traversal blocks per CU, the number of XCD group ranges and
the LIFO depth per wavefront held in LDS (full depth, 22 x 64 entries, never overflows; a shallower LIFO uses
less LDS per block, and the groups that overflow it are traversed again at full depth) -- and keeps the
fastest. Results are cached per size class (log2 of the median system size) and can be saved as JSON.
A configuration is (group size, blocks per CU, group ranges[, LIFO entries per wave]); 0 keeps a default.
"""

from __future__ import annotations

import json
import math
import time

import numpy as np

from . import Batch, Context, default_params

CANDIDATES = [(gs, blocks, ranges, stack) for gs in (64, 32) for blocks in (4, 6, 8) for ranges in (8, 1)
              for stack in (0, 8 * 64)]


def size_class(systems) -> int:
    return int(round(math.log2(max(1.0, float(np.median([s.n for s in systems]))))))


def time_solve(batch: Batch, prm, reps: int = 2, graph: bool = True) -> float:
    batch.solve(prm, graph=graph)  # warm-up (allocations, graph capture, caches)
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.solve(prm, graph=graph)
    return (time.perf_counter() - t0) / reps


class Autotuner:
    def __init__(self, ctx: Context, candidates=None, reps: int = 2, graph: bool = True):
        self.ctx = ctx
        self.candidates = [tuple(c) for c in (candidates or CANDIDATES)]
        self.reps = reps
        self.graph = graph
        self.table = {}

    def apply(self, cfg):
        """set the traversal launch shape of the context; returns the params for the group size"""
        gs, blocks, ranges = cfg[:3]
        stack = cfg[3] if len(cfg) > 3 else 0
        self.ctx.call("ryo_ctx_set_traversal_shape", int(stack), int(blocks), int(ranges))
        return default_params(group_size=int(gs))

    def reset(self):
        self.ctx.call("ryo_ctx_set_traversal_shape", 0, 0, 0)

    def tune(self, systems, batch: Batch | None = None):
        """time every candidate on this batch; returns (best config, [(config, seconds), ...])"""
        own = batch is None
        if own:
            batch = Batch(self.ctx, sum(s.n for s in systems), len(systems))
            batch.load(systems)
        timings = []
        try:
            for cfg in self.candidates:
                prm = self.apply(cfg)
                timings.append((cfg, time_solve(batch, prm, self.reps, self.graph)))
        finally:
            self.reset()
            if own:
                batch.close()
        best = min(timings, key=lambda t: t[1])[0]
        self.table[size_class(systems)] = best
        return best, timings

    def config_for(self, systems):
        """the tuned configuration of the systems' size class (nearest tuned class), or the defaults"""
        if not self.table:
            return (64, 0, 0, 0)
        c = size_class(systems)
        k = min(self.table, key=lambda t: abs(t - c))
        return self.table[k]

    def save(self, path):
        with open(path, "w") as f:
            json.dump({str(k): list(v) for k, v in self.table.items()}, f, indent=1)

    def load(self, path):
        with open(path) as f:
            self.table = {int(k): tuple(v) for k, v in json.load(f).items()}
