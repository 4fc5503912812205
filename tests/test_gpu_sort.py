"""Track ordering on the device.

- init_charge: the slot layout after initialization equals the reference's
  PartitionDataTest strings (TrackSort.test.cc:523-581) and the oracle's.
- Reindexing: after each sort action, track_slots is a permutation of the
  slots, the keys are non-decreasing in thread order, equal keys keep slot
  order (this build's sort is stable; the reference's is not, so only the
  key ranges are comparable), and action_offsets[k] equals the number of
  slots whose key is below k (count_tracks_per_action,
  TrackSortUtils.cc:80-157), which sorted launches read as their range.
- reindex_shuffle: the host permutation of the reference
  (TrackSlotUtils.cc:21-31), a permutation other than the identity.
- Physics is independent of the order: per-detector energy deposition
  with sorted launches agrees with the unsorted loop within statistical
  error (chi^2 of batch means, p > 1e-3).
"""

import numpy as np
import pytest
from scipy import stats

import celeritas_amd as amd_pkg
from celeritas_amd import workloads as wl
from tests import oracle_bind as ob
from tests import sort_cases as sc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def physics():
    return amd_pkg.generate_physics(wl.testem3_physics_spec())


def _problem(physics, order, seed=20220511, capacity=1 << 16, max_events=1 << 12):
    return wl.testem3_problem(physics=physics, seed=seed, init_capacity=capacity,
                              max_events=max_events, track_order=order)


@pytest.mark.parametrize("slots,count,expected",
                         sc.PARTITION_INIT_HOST + sc.PARTITION_INIT_DEVICE)
def test_init_charge_layout(amd, physics, slots, count, expected):
    prob = _problem(physics, "init_charge", capacity=4096, max_events=128)
    params = amd.CoreParams(prob)
    st = amd.CoreState(params, slots)
    st.insert_primaries(sc.cycling_primaries(count))
    st.execute("extend-from-primaries")
    st.execute("initialize-tracks")
    got = sc.charge_string(st.get("status"), st.get("particle_id"))
    print(f"{slots} slots, {count} primaries: {got}", flush=True)
    assert got == expected
    # Same layout from the oracle
    ost = ob.State(ob.Problem(prob), slots)
    ost.insert_primaries(sc.cycling_primaries(count))
    ost.action("extend-from-primaries")
    ost.action("initialize-tracks")
    assert sc.charge_string(ost.get_u32("status"), ost.get_u32("particle")) == got


SORTS = {
    # order: [(sort action label, key field)]
    "reindex_status": [("sort-tracks-status", "status")],
    "reindex_particle_type": [("sort-tracks-start", "particle_id")],
    # The along-step sort runs before the pre-step, keyed by the action the
    # pre-step will choose: the particle's (by charge)
    "reindex_along_step_action": [("sort-tracks-along-step", "along_of_particle")],
    "reindex_step_limit_action": [("sort-tracks-post-step", "post_step_action")],
    "reindex_both_action": [("sort-tracks-post-step", "post_step_action"),
                            ("sort-tracks-along-step", "along_of_particle")],
}


def _along_of_particle(params, physics):
    labels = params.action_labels()
    neutral = labels.index("along-step-neutral")
    charged = labels.index("along-step-general-linear")
    return np.array([neutral if p["charge"] == 0 else charged
                     for p in physics["particles"]], dtype=np.int64)


def _host_keys(st, field, num_keys, along=None):
    status = st.get("status")
    if field == "status":
        keys = np.zeros(len(status), dtype=np.int64)
    elif field == "along_of_particle":
        keys = np.minimum(along[st.get("particle_id").astype(np.int64)], num_keys - 1)
        keys[status == 3] = num_keys - 1  # errored: the pre-step chooses no action
    else:
        keys = np.minimum(st.get(field).astype(np.int64), num_keys - 1)
    keys[status == 0] = num_keys - 1
    return keys


def _check_sorted(st, field, num_keys, along=None):
    n = st.size
    keys = _host_keys(st, field, num_keys, along)
    slots = st.get("track_slots").astype(np.int64)
    np.testing.assert_array_equal(np.sort(slots), np.arange(n))
    k = keys[slots]
    assert np.all(np.diff(k) >= 0), "keys not sorted"
    # Stable: within a key, slots ascend
    same = np.diff(k) == 0
    assert np.all(np.diff(slots)[same] > 0), "equal keys out of slot order"
    offsets = st.get("action_offsets", count=num_keys + 1).astype(np.int64)
    expected = np.searchsorted(np.sort(keys), np.arange(num_keys + 1), side="left")
    expected[num_keys] = n
    np.testing.assert_array_equal(offsets, expected)


@pytest.mark.parametrize("order", sorted(SORTS))
def test_sort_invariants(amd, physics, order):
    params = amd.CoreParams(_problem(physics, order))
    labels = params.action_labels()
    num_actions = len(labels)
    stepper = amd.Stepper(params, 2048)
    prim = wl.testem3_primaries(8, emin=1e3, emax=1e4, seed=11)
    stepper(prim)
    for step in range(12):
        for label, field in SORTS[order]:
            assert label in labels
            num_keys = {"status": 2, "particle_id": 4}.get(field, num_actions + 1)
            stepper.state.execute(label)
            print(f"{order} step {step}: {label}", flush=True)
            _check_sorted(stepper.state, field, num_keys,
                          _along_of_particle(params, physics))
        r = stepper()
        if not r["active"] and not r["queued"]:
            break


def test_shuffle_permutation(amd, physics):
    params = amd.CoreParams(_problem(physics, "reindex_shuffle"))
    st = amd.CoreState(params, 1000)
    slots = st.get("track_slots").astype(np.int64)
    np.testing.assert_array_equal(np.sort(slots), np.arange(1000))
    assert np.any(slots != np.arange(1000))


def _edep_batches(amd, physics, order, seed, nbatch, nprim, energy):
    params = amd.CoreParams(_problem(physics, order, seed=seed, capacity=1 << 18))
    stepper = amd.Stepper(params, 1 << 14)
    out = []
    for b in range(nbatch):
        prim = wl.testem3_primaries(nprim, emin=energy, emax=energy * 1.0001,
                                    seed=100 + b, gamma_fraction=0.5)
        before = stepper.state.get("calo_edep", 100).copy()
        stepper.run(prim, max_steps=100_000)
        out.append(stepper.state.get("calo_edep", 100) - before)
        print(f"{order} batch {b}: {out[-1].sum():.1f} MeV", flush=True)
    t = np.array(out)
    return t.mean(axis=0), t.std(axis=0, ddof=1) / np.sqrt(nbatch)


@pytest.mark.parametrize("order", ["init_charge", "reindex_both_action"])
def test_order_does_not_change_physics(amd, physics, order):
    nbatch, nprim, energy = 10, 6, 2e3
    um, us = _edep_batches(amd, physics, "none", 1001, nbatch, nprim, energy)
    sm, ss = _edep_batches(amd, physics, order, 3003, nbatch, nprim, energy)
    sel = (us > 0) & (ss > 0) & (um + sm > 1e-3 * (um + sm).max())
    chi2 = float((((um - sm) ** 2) / (us**2 + ss**2))[sel].sum())
    ndf = int(sel.sum())
    p = stats.chi2.sf(chi2, ndf)
    print(f"{order}: chi2 = {chi2:.1f} for {ndf} detectors, p = {p:.3g}")
    assert ndf > 20
    assert p > 1e-3
