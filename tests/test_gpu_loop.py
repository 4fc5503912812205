"""Whole step loop on the GPU against the reference anchors and the oracle."""

import numpy as np
import pytest

from tests import oracle_bind as ob
from tests import problems

pytestmark = pytest.mark.gpu


def run_stepper(stepper, primaries, max_iters=1_000_000):
    """StepperTestBase::run (test/celeritas/global/StepperTestBase.cc:64-100)."""
    r = stepper(primaries)
    active, queued = [r["active"]], [r["queued"]]
    while r and len(active) < max_iters:
        r = stepper()
        active.append(r["active"])
        queued.append(r["queued"])
    return np.array(active), np.array(queued)


def test_simple_compton_anchor(amd):
    params = amd.CoreParams(problems.simple_compton())
    stepper = amd.Stepper(params, 64)
    assert stepper.action_sequence() == [
        "extend-from-primaries", "initialize-tracks", "pre-step", "along-step-neutral",
        "physics-discrete-select", "scat-klein-nishina", "geo-boundary", "tracking-cut",
        "extend-from-secondaries"]
    active, queued = run_stepper(stepper, problems.simple_compton_primaries(32))
    assert len(active) == 919
    assert active.sum() / 32 == 53.8125
    assert (int(np.argmax(queued)), int(queued.max())) == (1, 6)
    assert ob.emptying_step(active) == 3

    # Step-by-step identical to the oracle
    ost = ob.State(ob.Problem(problems.simple_compton()), 64)
    oact, oq = ob.run(ost, problems.simple_compton_primaries(32))
    np.testing.assert_array_equal(active, oact)
    np.testing.assert_array_equal(queued, oq)


def test_simple_compton_final_state(amd):
    """After a fixed number of steps the slot states match the oracle."""
    prob = problems.simple_compton()
    params = amd.CoreParams(prob)
    stepper = amd.Stepper(params, 64)
    ost = ob.State(ob.Problem(prob), 64)
    prim = problems.simple_compton_primaries(32)
    stepper(prim)
    ost.insert_primaries(prim)
    import ctypes
    f = ob.load().orc_step_once if hasattr(ob.load(), "orc_step_once") else None
    ob.step_once(ost)
    for _ in range(40):
        stepper()
        ob.step_once(ost)
    st = stepper.state
    np.testing.assert_array_equal(st.get("status"), ost.get_u32("status"))
    np.testing.assert_array_equal(st.get("num_steps"), ost.get_u32("num_steps"))
    live = st.get("status") != 0
    for name, oname in (("energy", "energy"), ("time", "time")):
        np.testing.assert_allclose(st.get(name)[live], ost.get_f64(oname)[live], rtol=1e-12)
    pos = np.stack([st.get("pos_x"), st.get("pos_y"), st.get("pos_z")], axis=1)
    opos = ost.get_f64("pos").reshape(-1, 3)
    np.testing.assert_allclose(pos[live], opos[live], rtol=1e-12, atol=1e-12)
