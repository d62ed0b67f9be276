"""Autotuner for the tracer's launch and grid parameters.

Tunable parameters (all leave the traced results unchanged except near-ties
that depend on the grid):
  lam     grid resolution factor (cells ~ lam * triangles; the reference fixed
          k = 6 in UniformGridAccStructBuilder.cpp:91 and exposed it in its GUI,
          WdgTransformEdit.cpp:49-69)
  sched   per traversal stage, 's' one ray per lane, 'w' one ray per lane with
          the wave's triangle tests pooled, 'p' persistent refill, or '2', '4',
          '8' lanes per ray
  refill  idle lanes that trigger a refill in the persistent schedule
  block   threads per block of the traversal kernels

The tuner measures candidates with a caller-supplied timing function and keeps
the results (optionally in a JSON file, the "tuned parameters" file) keyed by a
workload description, so later runs reuse them.
"""
from __future__ import annotations

import json
import os
from typing import Callable, Iterable


class Autotuner:
    def __init__(self, cache_path: str | None = None):
        self.cache_path = cache_path
        self.cache: dict = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)
        self.log: list = []

    def lookup(self, key: str):
        return self.cache.get(key)

    def tune(self, key: str, candidates: Iterable[dict], measure: Callable[[dict], float],
             reuse: bool = True) -> dict:
        """Return the candidate with the lowest measured cost (seconds)."""
        if reuse and key in self.cache:
            return self.cache[key]["best"]
        best, best_cost, tried = None, float("inf"), []
        for cand in candidates:
            cost = float(measure(cand))
            tried.append({"params": cand, "cost": cost})
            self.log.append((key, cand, cost))
            if cost < best_cost:
                best, best_cost = cand, cost
        if best is None:
            raise ValueError("autotune: no candidates")
        self.cache[key] = {"best": best, "cost": best_cost, "tried": tried}
        if self.cache_path:
            tmp = self.cache_path + ".tmp"
            with open(tmp, "w") as f:
                json.dump(self.cache, f, indent=1)
            os.replace(tmp, self.cache_path)
        return best


def per_stage_schedule(stage_ms: dict) -> str:
    """Pick, stage by stage, the fastest traversal schedule: stage_ms maps a
    schedule character to its list of per-stage milliseconds."""
    chars = list(stage_ms)
    n = len(stage_ms[chars[0]])
    return "".join(min(chars, key=lambda ch: stage_ms[ch][i]) for i in range(n))
