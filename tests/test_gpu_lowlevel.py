"""The per-action launch ABI (celer_hip_*_launch, celer_hip_sort_tracks).

A caller that owns the step sequence drives the kernels of each action
directly on a state, with its own thread maps, as the reference's step
actions launch their executors (ActionLauncher.device.hh:102-135). Two
states of the same params are stepped in lockstep: one through the
library's step actions (celer_state_execute_action), the other through
the low-level entry points, choosing for each action the kernel and the
thread map (sorted range or all slots) the action would. The two must
stay bit-identical in every slot: SimpleCompton (no reindexing) and
TestEm3 with action-sorted launches (reindex_both_action), in FP64 and
in the mixed-precision build.
"""

import ctypes

import numpy as np
import pytest

import celeritas_amd as amd_pkg
from celeritas_amd import workloads as wl
from tests import problems

pytestmark = pytest.mark.gpu

MODEL_OF_LABEL = {
    "scat-klein-nishina": 0, "photoel-livermore": 1, "conv-bethe-heitler": 2,
    "ioni-moller-bhabha": 3, "brems-sb": 4, "brems-rel": 5, "annihil-2-gamma": 6,
    "brems-combined": 7,
}
ALONG, POST = 7, 11  # StepActionOrder
SORT_KEY = {"sort-tracks-post-step": 3, "sort-tracks-status": 0, "sort-tracks-particle": 1}


class ThreadSlots(ctypes.Structure):
    _fields_ = [("range", ctypes.c_void_p), ("slots", ctypes.c_void_p)]


def _lib():
    lib = amd_pkg.lib()
    vp, u32 = ctypes.c_void_p, ctypes.c_uint32
    sigs = {
        "celer_params_device_ref": (vp, [vp]),
        "celer_state_device_ref": (vp, [vp]),
        "celer_state_stream": (vp, [vp]),
        "celer_state_track_slots": (vp, [vp]),
        "celer_state_action_offsets": (vp, [vp]),
        "celer_params_action_order": (ctypes.c_int, [vp, u32]),
        "celer_hip_pre_step_launch": (ctypes.c_int, [vp, vp, vp, vp]),
        "celer_hip_along_step_launch": (ctypes.c_int, [vp, vp, u32, ctypes.c_int, vp, vp]),
        "celer_hip_discrete_select_launch": (ctypes.c_int, [vp, vp, vp, vp]),
        "celer_hip_interaction_launch": (ctypes.c_int, [vp, vp, u32, u32, vp, vp]),
        "celer_hip_boundary_launch": (ctypes.c_int, [vp, vp, vp, vp]),
        "celer_hip_tracking_cut_launch": (ctypes.c_int, [vp, vp, vp, vp]),
        "celer_hip_calo_tally_launch": (ctypes.c_int, [vp, vp, u32, vp, vp]),
        "celer_hip_sort_tracks": (ctypes.c_int, [vp, u32, u32, vp]),
        "celer_hip_sort_tracks_along_before_pre": (ctypes.c_int, [vp, vp, u32, vp]),
        "celer_hip_initialize_tracks": (ctypes.c_int, [vp, vp, vp]),
        "celer_hip_extend_from_secondaries": (ctypes.c_int, [vp, vp, vp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype, fn.argtypes = res, args
    return lib


def _low_level_step(lib, params, state, labels, sorted_launch, num_detectors):
    """One step of the action sequence through the per-action entry points."""
    p = lib.celer_params_device_ref(params.handle)
    s = lib.celer_state_device_ref(state._h)
    st = lib.celer_state_stream(state._h)
    slots = lib.celer_state_track_slots(state._h) if sorted_launch is not None else None
    offsets = lib.celer_state_action_offsets(state._h)
    nact = params.num_actions()
    for aid, label in labels:
        order = lib.celer_params_action_order(params.handle, aid)
        ts = ThreadSlots(None, slots)
        if sorted_launch and order in (ALONG, POST):
            ts.range = offsets + 4 * aid
        tsp = ctypes.byref(ts)
        if label in ("extend-from-primaries",):
            state.execute(aid)  # host primaries: the library's action copies them
            rc = 0
        elif label == "initialize-tracks":
            rc = lib.celer_hip_initialize_tracks(p, s, st)
        elif label == "extend-from-secondaries":
            rc = lib.celer_hip_extend_from_secondaries(p, s, st)
        elif label == "pre-step":
            rc = lib.celer_hip_pre_step_launch(p, s, tsp, st)
        elif label.startswith("along-step"):
            rc = lib.celer_hip_along_step_launch(p, s, aid, label != "along-step-neutral",
                                                 tsp, st)
        elif label == "physics-discrete-select":
            rc = lib.celer_hip_discrete_select_launch(p, s, tsp, st)
        elif label in MODEL_OF_LABEL:
            rc = lib.celer_hip_interaction_launch(p, s, aid, MODEL_OF_LABEL[label], tsp, st)
        elif label == "geo-boundary":
            rc = lib.celer_hip_boundary_launch(p, s, tsp, st)
        elif label == "tracking-cut":
            rc = lib.celer_hip_tracking_cut_launch(p, s, tsp, st)
        elif label == "simple-calo":
            rc = lib.celer_hip_calo_tally_launch(p, s, num_detectors,
                                                 ctypes.byref(ThreadSlots(None, None)), st)
        elif label in SORT_KEY:
            rc = lib.celer_hip_sort_tracks(s, SORT_KEY[label], nact + 1, st)
        elif label == "sort-tracks-along-step":
            rc = lib.celer_hip_sort_tracks_along_before_pre(p, s, nact + 1, st)
        else:
            raise AssertionError(f"no low-level launch for action {label}")
        assert rc == 0, (label, rc)


FIELDS = ("status", "particle_id", "energy", "pos_x", "pos_y", "pos_z", "dir_x", "volume",
          "num_steps", "interaction_mfp", "energy_deposition", "post_step_action")


def _lockstep(amd, prob, prim, slots, steps, sorted_launch, num_detectors=0):
    lib = _lib()
    params = amd.CoreParams(prob)
    ref_stepper = amd.Stepper(params, slots)
    labels = [(params.find_action(l), l) for l in ref_stepper.action_sequence()]
    a = amd.CoreState(params, slots)
    b = amd.CoreState(params, slots)
    for k in range(steps):
        if k == 0:
            a.insert_primaries(prim)
            b.insert_primaries(prim)
        for aid, _ in labels:
            a.execute(aid)
        _low_level_step(lib, params, b, labels, sorted_launch, num_detectors)
        a.synchronize()
        b.synchronize()
        for f in FIELDS:
            np.testing.assert_array_equal(a.get(f), b.get(f), err_msg=f"step {k} field {f}")
        ca, cb = a.counters(), b.counters()
        assert ca == cb, f"step {k}: {ca} != {cb}"
    if num_detectors:
        # Deposits are summed in LDS with atomics, whose order varies
        np.testing.assert_allclose(a.get("calo_edep", num_detectors),
                                   b.get("calo_edep", num_detectors), rtol=1e-13, atol=0)
    return a


def test_simple_compton_lockstep(amd):
    a = _lockstep(amd, problems.simple_compton(), problems.simple_compton_primaries(32),
                  64, 60, sorted_launch=None)
    assert (a.get("status") == 2).any() or a.counters()["num_alive"] == 0


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_testem3_sorted_lockstep(amd, precision):
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics, init_capacity=1 << 16, max_events=64,
                              track_order="reindex_both_action", precision=precision)
    prim = wl.testem3_primaries(8, emin=1e3, emax=2e3, seed=5)
    a = _lockstep(amd, prob, prim, 4096, 40, sorted_launch=True, num_detectors=100)
    assert a.counters()["num_active"] > 8
