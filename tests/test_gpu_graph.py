"""Graph-captured steps against eager steps.

The step sequence captured as a hipGraph is replayed for every step
without new primaries; after each replay (with the program's own copies
and synchronizations in between) the state equals that of an eager
stepper run from the same seed and primaries.
"""

import numpy as np
import pytest

import celeritas_amd as amd_pkg
from celeritas_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", ["none", "reindex_both_action"])
def test_graph_replays_match_eager(amd, order):
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics, init_capacity=1 << 18, max_events=1 << 10,
                              track_order=order)
    eager = amd.Stepper(amd.CoreParams(prob), 4096)
    graph = amd.Stepper(amd.CoreParams(prob), 4096)
    prim = wl.testem3_primaries(6, emin=1e3, emax=2e3, seed=9)
    r_e, r_g = eager(prim), graph(prim)
    assert dict(r_e) == dict(r_g)
    graph.capture_step()
    for step in range(40):
        r_e, r_g = eager(), graph()
        print(f"step {step}: {dict(r_g)}", flush=True)
        assert dict(r_e) == dict(r_g)
        for field in ("status", "particle_id", "post_step_action"):
            np.testing.assert_array_equal(eager.state.get(field), graph.state.get(field))
        for field in ("energy", "pos_x", "pos_y", "pos_z"):
            np.testing.assert_array_equal(eager.state.get(field), graph.state.get(field))
        if not r_e:
            break
    # New primaries mid-run take the eager path, then replays resume
    more = wl.testem3_primaries(2, emin=1e3, emax=2e3, seed=10, first_event=6)
    assert dict(eager(more)) == dict(graph(more))
    for _ in range(10):
        assert dict(eager()) == dict(graph())
    # The tally sums with atomics, so only its rounding may differ
    np.testing.assert_allclose(eager.state.get("calo_edep", 100),
                               graph.state.get("calo_edep", 100), rtol=1e-12)


def test_concurrent_actions_match_serial(amd):
    """Track-disjoint actions forked onto side streams, eagerly and inside a
    captured graph, leave the state a serial launch sequence leaves."""
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics, init_capacity=1 << 20, max_events=1 << 10,
                              track_order="reindex_both_action")
    serial = amd.Stepper(amd.CoreParams(dict(prob, concurrent_actions=False)), 1 << 15)
    eager = amd.Stepper(amd.CoreParams(prob), 1 << 15)
    graph = amd.Stepper(amd.CoreParams(prob), 1 << 15)
    steppers = (serial, eager, graph)
    prim = wl.testem3_primaries(24, emin=1e3, emax=1e4, seed=21)
    first = [dict(s(prim)) for s in steppers]
    assert first[0] == first[1] == first[2]
    graph.capture_step()
    for step in range(400):
        rs = [dict(s()) for s in steppers]
        if step % 20 == 0:
            print(f"step {step}: {rs[0]}", flush=True)
        assert rs[0] == rs[1] == rs[2]
        if step % 10 == 0 or not (rs[0]["alive"] or rs[0]["queued"]):
            for field in ("status", "particle_id", "post_step_action", "energy",
                          "pos_x", "pos_y", "pos_z"):
                ref = serial.state.get(field)
                np.testing.assert_array_equal(ref, eager.state.get(field))
                np.testing.assert_array_equal(ref, graph.state.get(field))
        if not (rs[0]["alive"] or rs[0]["queued"]):
            break
    assert not (rs[0]["alive"] or rs[0]["queued"])
    ref = serial.state.get("calo_edep", 100)
    np.testing.assert_allclose(eager.state.get("calo_edep", 100), ref, rtol=1e-12)
    np.testing.assert_allclose(graph.state.get("calo_edep", 100), ref, rtol=1e-12)
