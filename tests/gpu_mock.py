"""Test doubles acting on a device state through the C ABI.

mock_interact mirrors MockInteractExecutor
(test/celeritas/track/MockInteractExecutor.hh): alive tracks are killed when
their flag is false and allocate the given number of secondaries (gamma
label 0, 5 MeV, along +x). pre_step_status mirrors the status part of
PreStepExecutor (initializing -> alive, secondaries cleared) for problems
without physics processes.
"""

import numpy as np

from celeritas_amd import SECONDARY_DTYPE

ALIVE, KILLED, INITIALIZING = 2, 4, 1


def pre_step_status(state):
    status = state.get("status")
    status[status == INITIALIZING] = ALIVE
    state.set("status", status)
    state.set("sec_count", np.zeros(state.size, dtype=np.uint32))
    state.set_counters({"secondary_stack_size": 0})


def mock_interact(state, num_secondaries, alive):
    n = state.size
    status = state.get("status")
    nsec = np.asarray(num_secondaries, dtype=np.uint32)
    alive = np.asarray(alive, dtype=bool)
    is_alive = status == ALIVE
    status[is_alive & ~alive] = KILLED
    counts = np.where(is_alive, nsec, 0).astype(np.uint32)
    starts = np.zeros(n, dtype=np.uint32)
    starts[1:] = np.cumsum(counts)[:-1]
    starts[counts == 0] = 0
    total = int(counts.sum())
    if total:
        sec = np.zeros(total, dtype=SECONDARY_DTYPE)
        sec["particle_id"] = 0
        sec["energy"] = 5.0
        sec["dir"] = [1.0, 0.0, 0.0]
        cap_bytes = sec.nbytes
        buf = state.get("secondaries", count=total)
        buf[:] = sec
        state.set("secondaries", buf)
    state.set("status", status)
    state.set("sec_start", starts)
    state.set("sec_count", counts)
    state.set_counters({"secondary_stack_size": total})
