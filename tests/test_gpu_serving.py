"""Continuous-batching serving loop on the device.

- Every event completes, admission follows arrival, and completion follows
  admission; the per-event deposited energies add up to the calorimeter
  tally of the same run.
- Events admitted later overlap with those in flight (the loop does not
  drain the device between events).
- Each event's result is the same alone and in the mix within statistical
  error: the per-event deposited fraction of a set of events transported
  one at a time and the same events transported together agree (mean
  difference within 4 standard errors; the RNG streams are per track
  slot, so individual showers differ and only the distribution is
  comparable).
"""

import numpy as np
import pytest

import celeritas_amd as amd_pkg
from celeritas_amd import serving
from celeritas_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def params(amd):
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    return amd.CoreParams(wl.testem3_problem(physics=physics, init_capacity=1 << 22,
                                             max_events=1 << 12,
                                             track_order="reindex_both_action"))


def _small_events(n, seed):
    return wl.served_events(n, seed=seed, nmin=2, nmax=20, emin=100.0, emax=2000.0)


def test_serving_completes_and_tallies(amd, params):
    server = serving.Server(params, 1 << 14)
    events = _small_events(24, seed=3)
    arrivals = serving.poisson_arrivals(len(events), rate_hz=400.0, seed=1)
    r = server.run(events, arrivals, max_steps=200_000)
    print(f"{len(events)} events in {r['wall_time']:.3f} s, {r['steps']} steps, "
          f"p50 {1e3 * r['p50_latency']:.1f} ms, p99 {1e3 * r['p99_latency']:.1f} ms",
          flush=True)
    assert np.all(r["admit"] >= arrivals - 1e-9)
    assert np.all(r["done"] >= r["admit"])
    assert np.all(r["latency"] > 0)
    calo = server.state.get("calo_edep", 100)
    # Nothing is deposited in the vacuum world: the per-event tally and the
    # calorimeter tally count the same deposits
    assert r["edep"].sum() == pytest.approx(calo.sum(), rel=1e-12)
    # Deposited energy stays below the event's energy (kinetic plus the rest
    # energy that positron primaries release when they annihilate) and is
    # substantial
    e0 = np.array([e["energy"].sum() for e in events])
    rest = np.array([2 * 0.51099895 * np.sum(e["particle_id"] == wl.POSITRON)
                     for e in events])
    assert np.all(r["edep"] < (e0 + rest) * (1 + 1e-9))
    assert np.mean(r["edep"] / e0) > 0.5


def test_events_overlap(amd, params):
    server = serving.Server(params, 1 << 14)
    events = _small_events(16, seed=5)
    r = server.run(events, np.zeros(len(events)), max_steps=200_000)
    # All arrive at once: several are admitted before the first completes
    first_done = r["done"].min()
    assert np.sum(r["admit"] <= first_done) >= 2


def test_device_admission_in_order_within_limit(amd, params):
    """The device admits staged events in arrival order while the queued
    initializers stay within the limit, and an event larger than the limit
    once nothing is queued or alive."""
    server = serving.Server(params, 1 << 14, steps_per_sync=1)
    events = _small_events(12, seed=41)
    big = wl.served_events(1, seed=42, nmin=60, nmax=60, emin=100.0, emax=200.0)[0]
    events.insert(5, big)
    limit = 40
    assert len(big) > limit and sum(len(e) for e in events) > 2 * limit
    r = server.run(events, np.zeros(len(events)), max_queued=limit, max_steps=200_000)
    print("admit", np.round(r["admit"], 4), "done", np.round(r["done"], 4), flush=True)
    assert np.all(r["done"] > 0)
    # In order, and not all at once: the limit holds later events back
    assert np.all(np.diff(r["admit"]) >= 0)
    assert r["admit"][-1] > r["admit"][0]
    # The large event waits for the device to empty: everything admitted
    # before it has completed by the time it is admitted
    assert np.all(r["done"][:5] <= r["admit"][5])
    calo = server.state.get("calo_edep", 100)
    assert r["edep"].sum() == pytest.approx(calo.sum(), rel=1e-12)


def test_same_alone_and_in_mix(amd, params):
    events = _small_events(40, seed=11)
    e0 = np.array([e["energy"].sum() for e in events])
    server = serving.Server(params, 1 << 14)
    mix = server.run(events, np.zeros(len(events)), max_steps=400_000)
    alone = np.zeros(len(events))
    for i, ev in enumerate(events):
        r = server.run([ev], [0.0], max_steps=200_000)
        alone[i] = r["edep"][0]
    fa, fm = alone / e0, mix["edep"] / e0
    d = fm - fa
    se = d.std(ddof=1) / np.sqrt(len(d))
    print(f"mean deposited fraction alone {fa.mean():.4f}, in mix {fm.mean():.4f}, "
          f"difference {d.mean():+.4f} +- {se:.4f}")
    assert abs(d.mean()) < 4 * se + 1e-12


# --------------------------------------------------------------------------- #
# Mixed precision (the served configuration)
# --------------------------------------------------------------------------- #
@pytest.fixture(scope="module")
def mixed_params(amd):
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    return amd.CoreParams(wl.testem3_problem(physics=physics, init_capacity=1 << 22,
                                             max_events=1 << 12, precision="mixed",
                                             track_order="reindex_both_action"))


def test_mixed_served_events_vs_fp64_oracle(amd, mixed_params):
    """Per-event deposition of the mixed served run against the FP64 oracle
    transporting each event alone (energy deposited in every volume, as the
    served per-event tally counts it): the mean difference of the deposited
    fractions over the events is within 3 standard errors (independent RNG
    streams, so only the distribution is comparable)."""
    from tests import oracle_bind as ob

    # Per-event differences are heavy-tailed (events of two primaries can
    # leak much of their energy), so the standard error needs many events
    events = _small_events(160, seed=23)
    e0 = np.array([e["energy"].sum() for e in events])
    server = serving.Server(mixed_params, 1 << 14, steps_per_sync=4)
    r = server.run(events, serving.poisson_arrivals(len(events), rate_hz=300.0, seed=9),
                   max_steps=400_000)
    assert server.state.counters()["num_sampling_failures"] == 0
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    oprob = ob.Problem(wl.testem3_problem(physics=physics, seed=777, init_capacity=1 << 16,
                                          max_events=16))
    oracle = np.zeros(len(events))
    for i, ev in enumerate(events):
        st = ob.State(oprob, 1 << 12, stream_id=i + 1)  # independent RNG streams per event
        prim = ev.copy()
        prim["event_id"] = 0
        ob.run(st, prim)
        oracle[i] = st.get_f64("total_edep")[0]
    fm, fo = r["edep"] / e0, oracle / e0
    d = fm - fo
    se = d.std(ddof=1) / np.sqrt(len(d))
    print(f"mixed served deposited fraction {fm.mean():.4f}, FP64 oracle {fo.mean():.4f}: "
          f"difference {d.mean():+.4f} +- {se:.4f} (per-event FP64-oracle error)")
    assert abs(d.mean()) < 3 * se + 1e-12


def test_mixed_same_alone_and_in_mix(amd, mixed_params):
    events = _small_events(40, seed=31)
    e0 = np.array([e["energy"].sum() for e in events])
    server = serving.Server(mixed_params, 1 << 14)
    mix = server.run(events, np.zeros(len(events)), max_steps=400_000)
    alone = np.zeros(len(events))
    for i, ev in enumerate(events):
        alone[i] = server.run([ev], [0.0], max_steps=200_000)["edep"][0]
    d = mix["edep"] / e0 - alone / e0
    se = d.std(ddof=1) / np.sqrt(len(d))
    print(f"mixed: deposited fraction difference in mix - alone {d.mean():+.4f} +- {se:.4f}")
    assert abs(d.mean()) < 4 * se + 1e-12


def test_device_completion_list(amd, params):
    """Every event appears exactly once in the device's completion list, and
    several steps per counter readback give the same completions."""
    events = _small_events(20, seed=41)
    for steps in (1, 8):
        server = serving.Server(params, 1 << 14, steps_per_sync=steps)
        r = server.run(events, np.zeros(len(events)), max_steps=200_000)
        c = server.state.counters()
        assert c["num_completed"] == len(events)
        assert np.all(r["done"] >= r["admit"])
        assert r["track_steps"] > 0 and r["steps"] > 0
