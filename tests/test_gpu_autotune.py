"""Autotuner: the chosen launch configuration is the fastest of its
trials, and a second call is served from the cache."""

import json

import pytest

import celeritas_amd as amd_pkg
from celeritas_amd import autotune
from celeritas_amd import workloads as wl

pytestmark = pytest.mark.gpu


def test_autotune_picks_fastest_and_caches(amd, tmp_path):
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics, init_capacity=1 << 20, max_events=1 << 10)
    prim = wl.testem3_primaries(16, emin=1e3, emax=5e3, seed=4)
    cache = str(tmp_path / "tune.json")
    best = autotune.tune(amd, prob, prim, track_orders=("none", "reindex_both_action"),
                         block_sizes=(128, 256), slot_counts=(1 << 14,), cache=cache,
                         log=print)
    trials = best["trials"]
    assert len(trials) == 5
    fastest = max(t["track_steps_per_s"] for t in trials
                  if t["track_order"] == best["track_order"]
                  and t["block_size"] == best["block_size"])
    assert fastest > 0
    order_trials = [t for t in trials[:2]]
    assert best["track_order"] == max(order_trials, key=lambda t: t["track_steps_per_s"])[
        "track_order"]
    again = autotune.tune(amd, prob, prim, cache=cache)
    assert again == json.load(open(cache))[next(iter(json.load(open(cache))))]
    assert again["track_order"] == best["track_order"]
