"""Autotuner for the transport loop's launch configuration.

Chooses, for a problem and a source of primaries, the track-slot ordering
(TrackOrder), the workgroup size of the step kernels and the number of
track slots by timing short calibration transports on the device, and
caches the winner in a JSON file keyed by the problem's description. The
search is exhaustive over a small grid (the loop is limited by divergence
and memory latency, where the best point is found empirically, not
modelled).
"""

from __future__ import annotations

import hashlib
import json
import os
import time
from typing import Callable, Iterable, Sequence

import numpy as np

#: Candidate grid
TRACK_ORDERS = ("none", "init_charge", "reindex_step_limit_action", "reindex_both_action")
BLOCK_SIZES = (64, 128, 256)
SLOT_COUNTS = (1 << 19, 1 << 20, 1 << 21)


def _key(problem: dict, primaries: np.ndarray) -> str:
    desc = {k: v for k, v in problem.items() if k not in ("physics", "geometry")}
    h = hashlib.sha256(json.dumps(desc, sort_keys=True, default=str).encode())
    h.update(str(len(primaries)).encode())
    h.update(np.ascontiguousarray(primaries["energy"]).tobytes())
    return h.hexdigest()[:16]


def measure(amd, problem: dict, primaries: np.ndarray, track_order: str, block_size: int,
            slots: int, use_graph: bool = True, repeat: int = 1) -> float:
    """Track-steps per second of full transports of ``primaries``."""
    prob = dict(problem)
    prob["track_init"] = dict(problem.get("track_init", {}), track_order=track_order)
    params = amd.CoreParams(prob)
    stepper = amd.Stepper(params, slots)
    stepper.state.set_block_size(block_size)
    if use_graph:
        stepper.capture_step()
    stepper.run(primaries)  # warm-up
    stepper.state.synchronize()
    t0 = time.perf_counter()
    steps = sum(stepper.run(primaries)["track_steps"] for _ in range(repeat))
    stepper.state.synchronize()
    return steps / (time.perf_counter() - t0)


def tune(amd, problem: dict, primaries: np.ndarray,
         track_orders: Iterable[str] = TRACK_ORDERS,
         block_sizes: Iterable[int] = BLOCK_SIZES,
         slot_counts: Iterable[int] = SLOT_COUNTS,
         cache: str | None = None, log: Callable[[str], None] | None = None) -> dict:
    """Best (track_order, block_size, slots) for the problem; reads and
    updates the JSON cache when given."""
    key = _key(problem, primaries)
    table = {}
    if cache and os.path.exists(cache):
        with open(cache) as f:
            table = json.load(f)
        if key in table:
            return table[key]
    results = []
    # Coordinate search: order at the default launch shape, then block
    # size, then slots (the axes interact weakly)
    best = {"track_order": "none", "block_size": 256, "slots": 1 << 20}

    def trial(cfg):
        rate = measure(amd, problem, primaries, cfg["track_order"], cfg["block_size"],
                       cfg["slots"])
        results.append(dict(cfg, track_steps_per_s=rate))
        if log:
            log(f"autotune {cfg}: {rate:.4g} track-steps/s")
        return rate

    for axis, values in (("track_order", track_orders), ("block_size", block_sizes),
                         ("slots", slot_counts)):
        scores = {}
        for v in values:
            scores[v] = trial(dict(best, **{axis: v}))
        best[axis] = max(scores, key=scores.get)
    out = dict(best, trials=results)
    if cache:
        table[key] = out
        with open(cache, "w") as f:
            json.dump(table, f, indent=1)
    return out
