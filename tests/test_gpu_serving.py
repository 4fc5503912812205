"""Continuous-batching serving loop on the device.

- Every event completes, admission follows arrival, and completion follows
  admission; the per-event deposited energies add up to the calorimeter
  tally of the same run (the event tally also counts the vacuum world).
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
    assert r["edep"].sum() == pytest.approx(calo.sum(), rel=1e-3)
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
