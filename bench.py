#!/usr/bin/env python3
"""Throughput of the per-step transport loop on the TestEm3 calorimeter.

Workload (the FP64 GPU-loop configuration): 10,000 primaries, e-/gamma
50/50, energies log-uniform in 1-100 GeV, directions uniform in a 5 degree
cone about +x, started at (-22, 0, 0) cm in front of 50 layers of
2.3 mm Pb + 5.7 mm liquid argon (40 x 40 cm), 2^20 track slots, per-layer
energy deposition tallied. One bench step transports the whole set of
primaries to completion through the HIP stepping loop (one primary per
event). Inputs are synthetic; physics tables come from the built-in
generator.

Prints one JSON line with the metric (track-steps per second over the
timed steps; events/s beside it), a roofline entry for the dominant kernel
(algorithmic bytes per launch over its average device time, from events on
the state stream) and the CPU baseline (the oracle, the C++ restatement of
the same loop, timed single-threaded on a bounded sample of the same
primaries).
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

#: Algorithmic bytes per processed track and per slot scanned, from the SoA
#: layout in FP64 (fields each kernel must read and write once per track).
BYTES_PER_UNIT = {
    # status, particle, energy, volume, mfp read; deposition, secondary
    # span, element, pre-volume, step, actions, per-process and total XS,
    # range written
    "pre-step": (93, 1, "pre_step"),
    # geometry state (position, direction, volume, surface, next-step
    # cache) read and written, step, actions, energy, time, range, MFP,
    # macro XS, deposition, step count
    "along-step-general-linear": (218, 2, "along_charged"),
    "along-step-neutral": (194, 2, "along_neutral"),
    # geometry state read and written, material, status
    "geo-boundary": (120, 2, "boundary"),
}


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed full transports")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--primaries", type=int, default=10000)
    ap.add_argument("--slots", type=int, default=1 << 20)
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--track-order", default="reindex_both_action",
                    help="track slot ordering (none, init_charge, reindex_status, "
                         "reindex_particle_type, reindex_along_step_action, "
                         "reindex_step_limit_action, reindex_both_action)")
    ap.add_argument("--cpu-primaries", type=int, default=48,
                    help="primaries in the CPU baseline sample (0: skip)")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel on a buffer of this size "
                         "(reference for memory-counter profiles) and exit")
    return ap.parse_args()


def calibrate_copy(mib):
    """Plain copy of known size: FETCH_SIZE and WRITE_SIZE of this kernel
    convert the counters of the transport kernels into bytes."""
    import ctypes

    import celeritas_amd as amd

    f = amd.lib().celer_calibration_copy
    f.argtypes = [ctypes.c_size_t, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)]
    f.restype = ctypes.c_int
    ms = ctypes.c_float(0)
    nbytes = mib << 20
    rc = f(nbytes, 10, ctypes.byref(ms))
    if rc != 0:
        raise SystemExit(f"calibration copy failed: {rc}")
    print(json.dumps({"calibration_copy_bytes_read": nbytes,
                      "calibration_copy_bytes_written": nbytes,
                      "kernel_ms": ms.value,
                      "gb_per_s": 2 * nbytes / (ms.value * 1e-3) / 1e9}), flush=True)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for {args.gpus}); events are "
              "independent and the loop would shard by event, but no multi-GPU "
              "path is built", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return

    import celeritas_amd as amd
    from celeritas_amd import workloads as wl

    if not amd.device_available():
        raise SystemExit("bench.py: no HIP device available")

    log("generating physics tables")
    physics = amd.generate_physics(wl.testem3_physics_spec())
    problem = wl.testem3_problem(physics=physics, init_capacity=1 << 26,
                                 max_events=max(args.primaries, 1),
                                 track_order=args.track_order)
    primaries = wl.testem3_primaries(args.primaries, seed=args.seed)
    params = amd.CoreParams(problem)

    stepper = amd.Stepper(params, args.slots)
    log(f"state: {args.slots} slots")
    for w in range(args.warmup):
        r = stepper.run(primaries)
        log(f"warmup {w}: {r}")

    stepper.state.synchronize()
    t0 = time.perf_counter()
    totals = []
    for k in range(args.steps):
        totals.append(stepper.run(primaries))
        log(f"step {k}: {totals[-1]} at {time.perf_counter() - t0:.2f}s")
    stepper.state.synchronize()
    elapsed = time.perf_counter() - t0

    track_steps = sum(t["track_steps"] for t in totals)
    value = track_steps / elapsed
    events_per_s = args.steps * args.primaries / elapsed
    edep = stepper.state.get("calo_edep", max(problem["volume_to_detector"]) + 1)

    line = {
        "metric": "steps/s and events/s, FP32-mixed vs FP64 on 1 GPU; p50/p99 event latency",
        "value": value,
        "unit": "track-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "events_per_s": events_per_s,
        "track_steps_per_event": track_steps / (args.steps * args.primaries),
        "iterations_per_step": totals[-1]["iterations"],
        "config": {
            "workload": "TestEm3 50 x (2.3 mm Pb + 5.7 mm lAr), 10000 primaries e-/gamma "
                        "50/50, log-uniform 1-100 GeV, 5 deg cone, 2^20 slots, FP64",
            "primaries": args.primaries,
            "slots": args.slots,
            "track_order": args.track_order,
            "physics": "Compton, Livermore PE, Bethe-Heitler, e+ annihilation, "
                       "Moller-Bhabha ionisation, SB/relativistic brems, Urban MSC, "
                       "Urban energy-loss fluctuations",
            "edep_total_mev": float(edep.sum()),
        },
    }

    if not args.no_roofline:
        line["roofline"] = roofline(amd, params, primaries, args.slots)
    if args.cpu_primaries > 0:
        line["cpu_baseline"] = cpu_baseline(amd, wl, problem, args)
    print(json.dumps(line), flush=True)


def roofline(amd, params, primaries, slots):
    """Dominant kernel's algorithmic bandwidth from an instrumented run."""
    st = amd.Stepper(params, slots, action_times=True)
    st.diagnostics(enable=True)
    st.run(primaries)
    times = st.action_times()
    diag = st.diagnostics()
    ranked = sorted(times.items(), key=lambda kv: -kv[1][0])
    log("action device times [s]: " + ", ".join(f"{k}={v[0]:.4f}" for k, v in ranked))
    log(f"work counters: {diag}")
    label, (secs, launches) = next((kv for kv in ranked if kv[0] in BYTES_PER_UNIT),
                                   ranked[0])
    per_unit, per_slot, counter = BYTES_PER_UNIT[label]
    units = diag[counter]
    algorithmic = per_unit * units + per_slot * slots * launches
    per_launch = algorithmic / launches
    avg_launch = secs / launches
    achieved = per_launch / avg_launch / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(label)
    return {
        "bound": "hbm",
        "kernel": label,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "algorithmic_bytes_per_launch": per_launch,
        "avg_launch_us": 1e6 * avg_launch,
        "launches": launches,
        "share_of_step_time": secs / sum(v[0] for v in times.values()),
    }


def cpu_baseline(amd, wl, problem, args):
    """The oracle (C++ restatement of the loop) on a sample of the primaries."""
    from tests import oracle_bind as ob

    sample = wl.testem3_primaries(args.cpu_primaries, seed=args.seed + 1)
    for i in range(len(sample)):
        sample["event_id"][i] = i
    oprob = ob.Problem(problem)
    st = ob.State(oprob, 4096)
    t0 = time.perf_counter()
    active, _ = ob.run(st, sample, max_iters=10_000_000)
    dt = time.perf_counter() - t0
    steps = int(active.sum())
    return {
        "value": steps / dt,
        "unit": "track-steps/s",
        "cores": 1,
        "kind": "port",
        "events_per_s": len(sample) / dt,
        "sample": f"{len(sample)} primaries of the same distribution (seed {args.seed + 1}), "
                  f"4096 slots, single thread; {steps} track-steps in {dt:.1f} s",
    }


if __name__ == "__main__":
    main()
