"""Kernel attribute registry (the reference's KernelRegistry): every step
kernel launched by a TestEm3 step is recorded with the runtime's register,
scratch and LDS figures and its occupancy at the launch shape."""

import pytest

import celeritas_amd as amd_pkg
from celeritas_amd import workloads as wl

pytestmark = pytest.mark.gpu


def test_step_kernels_recorded(amd):
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics, init_capacity=1 << 16, max_events=64,
                              track_order="reindex_both_action")
    stepper = amd.Stepper(amd.CoreParams(prob), 4096)
    stepper(wl.testem3_primaries(4, emin=1e3, emax=2e3, seed=3))
    for _ in range(3):
        stepper()
    # The registry is process-wide: keep this stepper's launch shapes
    # (workgroups of 256, the tables staged in LDS where a kernel stages)
    entries = [a for a in amd_pkg.kernel_attributes() if a["block_size"] == 256]
    print(sorted({a["name"] for a in entries}))
    # Physics tables (pre-step, discrete select, interactions), geometry
    # tables (neutral along-step, boundary) or both (charged along-step) in
    # dynamic LDS
    staged = {"pre-step": 32, "discrete-select": 32, "interactions": 32,
              "along-step-neutral": 32, "boundary": 32, "along-step-linear": 52}
    # (the discrete selection runs inside the along-step kernels)
    for name in ("pre-step", "along-step-linear", "along-step-neutral", "interactions",
                 "boundary", "tracking-cut", "init-tracks", "process-secondaries"):
        shapes = [a for a in entries if a["name"] == name]
        if name in staged:
            shapes = [a for a in shapes if 0 < a["dynamic_lds"] <= staged[name] * 1024]
        else:
            shapes = [a for a in shapes if a["dynamic_lds"] == 0]
        assert shapes, name
        for a in shapes:
            assert a["num_regs"] > 0, name
            assert a["max_blocks_per_cu"] >= 1, name
            # A CU holds at most 32 waves (eight per SIMD)
            assert 0 < a["waves_per_simd"] <= 8, name
            assert 0 < a["occupancy"] <= 1, name
            assert a["local_mem"] == 0, name  # no register spills on the hot path


def test_kernel_timing(amd):
    """Launches carry dispatch start/stop events only while timing is on."""
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics, init_capacity=1 << 16, max_events=64,
                              track_order="reindex_both_action")
    stepper = amd.Stepper(amd.CoreParams(prob), 4096)
    prim = wl.testem3_primaries(4, emin=1e3, emax=2e3, seed=5)
    amd_pkg.kernel_times(reset=True)
    amd_pkg.kernel_timing(True)
    try:
        stepper(prim)
        for _ in range(5):
            stepper()
        times = {}
        for t in amd_pkg.kernel_times(reset=True):
            s, n = times.get(t["name"], (0.0, 0))
            times[t["name"]] = (s + t["seconds"], n + t["launches"])
    finally:
        amd_pkg.kernel_timing(False)
    print(times)
    for name in ("pre-step", "along-step-linear", "along-step-neutral", "interactions",
                 "boundary", "init-tracks", "process-secondaries"):
        secs, launches = times[name]
        assert launches == 6, name
        assert 0 < secs / launches < 0.1, name
    # Timing off: launches are no longer counted
    stepper()
    assert amd_pkg.kernel_times(reset=True) == []
