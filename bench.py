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
    # geometry state (position, direction, volume, surface, sense, boundary
    # flags, next-step cache: 72 B read, 68 B written), RNG state (24 + 24),
    # MSC range state (24 + 24), step, actions, energy, time, range, MFP,
    # macro XS, deposition, step count (184 B read, 161 B written in all)
    "along-step-general-linear": (345, 2, "along_charged"),
    # the same plus the looping-step counter
    "along-step-uniform-msc": (353, 2, "along_charged"),
    # geometry state, step, actions, time, MFP, macro XS, energy, steps
    "along-step-neutral": (216, 2, "along_neutral"),
    # geometry state read and written, status, action
    "geo-boundary": (142, 2, "boundary"),
}


def parse_args():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("testem3", "cylinder"), default="testem3",
                    help="testem3: the TestEm3 calorimeter (default); cylinder: the "
                         "barrel detector in a 3.8 T field")
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
    ap.add_argument("--serve-events", type=int, default=64,
                    help="events of the served TestEm3 workload (0: skip serving)")
    ap.add_argument("--serve-baseline-events", type=int, default=6,
                    help="events of the same stream run one at a time through the plain "
                         "FP64 loop, for the serving speedup")
    ap.add_argument("--mixed", action="store_true",
                    help="also measure the same workload in the mixed-precision mode "
                         "(off by default: validated at small scale only)")
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
    if args.workload == "cylinder":
        physics = amd.generate_physics(wl.cylinder_physics_spec())
        make_problem = wl.cylinder_problem
        problem = wl.cylinder_problem(physics=physics, init_capacity=1 << 26,
                                      max_events=max(args.primaries, 1),
                                      track_order=args.track_order)
        primaries = wl.cylinder_primaries(args.primaries, seed=args.seed)
        ndet = max(wl.cylinder_geometry()[2]) + 1
    else:
        physics = amd.generate_physics(wl.testem3_physics_spec())
        make_problem = wl.testem3_problem
        problem = wl.testem3_problem(physics=physics, init_capacity=1 << 26,
                                     max_events=max(args.primaries, 1),
                                     track_order=args.track_order)
        primaries = wl.testem3_primaries(args.primaries, seed=args.seed)
        ndet = max(problem["volume_to_detector"]) + 1
    params = amd.CoreParams(problem)

    stepper = amd.Stepper(params, args.slots)
    log(f"state: {args.slots} slots")
    for w in range(args.warmup):
        r = stepper.run(primaries, max_steps=20_000)
        log(f"warmup {w}: {r}")

    stepper.state.synchronize()
    t0 = time.perf_counter()
    totals = []
    for k in range(args.steps):
        totals.append(stepper.run(primaries, max_steps=20_000))
        log(f"step {k}: {totals[-1]} at {time.perf_counter() - t0:.2f}s")
    stepper.state.synchronize()
    elapsed = time.perf_counter() - t0

    track_steps = sum(t["track_steps"] for t in totals)
    value = track_steps / elapsed
    events_per_s = args.steps * args.primaries / elapsed
    edep = stepper.state.get("calo_edep", ndet)
    if args.workload == "cylinder":
        workload = ("Cylindrical detector (Si tracker 10 x 0.3 mm, PbWO4 ECAL 23 cm, "
                    "HCAL 20 x (5 cm brass + 4 mm scintillator)) in a uniform 3.8 T "
                    f"field, {args.primaries} primaries e-/e+/gamma 40/20/40, "
                    "log-uniform 0.5-50 GeV, isotropic, 2^20 slots, FP64, "
                    "Dormand-Prince propagation")
        physics_label = ("Compton, Livermore PE, Bethe-Heitler, e+ annihilation, "
                         "Moller-Bhabha ionisation, Seltzer-Berger and relativistic "
                         "brems, Urban MSC, Urban energy-loss fluctuations")
    else:
        workload = ("TestEm3 50 x (2.3 mm Pb + 5.7 mm lAr), 10000 primaries e-/gamma "
                    "50/50, log-uniform 1-100 GeV, 5 deg cone, 2^20 slots, FP64")
        physics_label = ("Compton, Livermore PE, Bethe-Heitler, e+ annihilation, "
                         "Moller-Bhabha ionisation, SB/relativistic brems, Urban MSC, "
                         "Urban energy-loss fluctuations")

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
            "workload": workload,
            "primaries": args.primaries,
            "slots": args.slots,
            "track_order": args.track_order,
            "physics": physics_label,
            "edep_total_mev": float(edep.sum()),
        },
    }

    if args.mixed:
        line["mixed"] = mixed_run(amd, make_problem, physics, primaries, args, value,
                                  events_per_s)
    if args.serve_events > 0 and args.workload == "testem3":
        line["serve"] = serve_run(amd, wl, physics, args)
    if not args.no_roofline:
        line["roofline"] = roofline(amd, params, primaries, args.slots, args.track_order)
    if args.cpu_primaries > 0:
        line["cpu_baseline"] = cpu_baseline(amd, wl, problem, args)
    print(json.dumps(line), flush=True)


def mixed_run(amd, make_problem, physics, primaries, args, fp64_value, fp64_events):
    """Same workload with the physics kernels in the mixed-precision mode."""
    problem = make_problem(physics=physics, init_capacity=1 << 26,
                           max_events=max(args.primaries, 1), track_order=args.track_order,
                           precision="mixed")
    params = amd.CoreParams(problem)
    stepper = amd.Stepper(params, args.slots)
    for w in range(args.warmup):
        log(f"mixed warmup {w}: {stepper.run(primaries, max_steps=20_000)}")
    stepper.state.synchronize()
    t0 = time.perf_counter()
    steps = 0
    for k in range(args.steps):
        r = stepper.run(primaries, max_steps=20_000)
        steps += r["track_steps"]
        log(f"mixed step {k}: {r} at {time.perf_counter() - t0:.2f}s")
    stepper.state.synchronize()
    dt = time.perf_counter() - t0
    value = steps / dt
    events = args.steps * args.primaries / dt
    log(f"mixed precision: {value:.4g} track-steps/s, {events:.4g} events/s")
    return {"value": value, "unit": "track-steps/s", "events_per_s": events,
            "ms_per_step": 1e3 * dt / args.steps, "speedup_events_vs_fp64": events / fp64_events,
            "dtype": "f32 transcendentals and draws in sampling, interpolation and MSC; "
                     "f64 positions, safety, field, energies and tallies"}


def serve_run(amd, wl, physics, args):
    """Streamed serving on TestEm3 (the served configuration): mixed
    precision, continuous batching on 2^21 slots with graph replay and
    device-side completion, saturated and at Poisson arrival rates swept
    below saturation; compared with the plain FP64 loop (each event
    transported to completion before the next, unsorted, no graph, as
    LocalTransporter flushes do) over the same whole stream, and with the
    same served loop in FP64 (the precision's own share)."""
    from celeritas_amd import serving

    events = wl.served_events(args.serve_events, seed=args.seed)
    n = len(events)
    energy_gev = float(sum(e["energy"].sum() for e in events)) / 1e3

    def served(precision):
        params = amd.CoreParams(wl.testem3_problem(
            physics=physics, init_capacity=1 << 26, max_events=1 << 16,
            track_order="reindex_both_action", precision=precision))
        server = serving.Server(params, 1 << 21)
        server.run(events[:2], [0.0, 0.0], max_steps=200_000)  # warm-up, graph capture
        return server

    server = served("mixed")
    sat = server.run(events, serving.poisson_arrivals(n, float("inf")), max_steps=400_000)
    log(f"served mixed (saturated): {sat['events_per_s']:.3g} events/s, "
        f"{sat['track_steps'] / sat['wall_time']:.3g} track-steps/s, "
        f"p50 {sat['p50_latency']:.3g} s, p99 {sat['p99_latency']:.3g} s")
    sweep = []
    for frac in (0.25, 0.5, 0.75, 0.9):
        rate = frac * sat["events_per_s"]
        r = server.run(events, serving.poisson_arrivals(n, rate, seed=args.seed + 7),
                       max_steps=400_000)
        sweep.append({"rate_hz": rate, "fraction_of_saturated": frac,
                      "events_per_s": r["events_per_s"],
                      "p50_latency_s": r["p50_latency"], "p99_latency_s": r["p99_latency"]})
        log(f"served mixed at {rate:.3g} Hz: p50 {r['p50_latency']:.3g} s, "
            f"p99 {r['p99_latency']:.3g} s")
    fails = server.state.counters()["num_sampling_failures"]
    del server

    fp64_server = served("fp64")
    sat64 = fp64_server.run(events, serving.poisson_arrivals(n, float("inf")),
                            max_steps=400_000)
    log(f"served FP64 (saturated): {sat64['events_per_s']:.3g} events/s")
    del fp64_server

    # Plain FP64 loop over the same whole stream, one event at a time
    plain = amd.CoreParams(wl.testem3_problem(physics=physics, init_capacity=1 << 26,
                                              max_events=1 << 16))
    stepper = amd.Stepper(plain, 1 << 20)
    stepper.run(events[0][:4])  # warm-up
    stepper.state.synchronize()
    t0 = time.perf_counter()
    for ev in events:
        stepper.run(ev, max_steps=200_000)
    stepper.state.synchronize()
    dt = time.perf_counter() - t0
    plain_rate = n / dt
    log(f"plain FP64 loop: {plain_rate:.3g} events/s over the {n} events")
    speedup = sat["events_per_s"] / plain_rate
    return {
        "workload": f"{n} TestEm3 events, primaries per event log-uniform 10-2000, "
                    "energies log-uniform 0.1-50 GeV, e-/e+/gamma 40/20/40, mixed "
                    "precision, reindex_both_action, graph replay, device-side "
                    "completion, 2^21 slots",
        "events_per_s": sat["events_per_s"],
        "track_steps_per_s": sat["track_steps"] / sat["wall_time"],
        "primary_gev_per_s": energy_gev / sat["wall_time"],
        "saturated_p50_latency_s": sat["p50_latency"],
        "saturated_p99_latency_s": sat["p99_latency"],
        "rate_sweep": sweep,
        "sampling_failures": int(fails),
        "served_fp64_events_per_s": sat64["events_per_s"],
        "plain_fp64_events_per_s": plain_rate,
        "plain_fp64_sample": f"the same {n} events, one at a time to completion, FP64, "
                             "unsorted, no graph, 2^20 slots",
        "speedup_events_vs_plain_fp64": speedup,
        "speedup_from_precision": sat["events_per_s"] / sat64["events_per_s"],
        "target_3x_met": bool(speedup >= 3.0),
    }


def sorted_launch(label, track_order):
    """Whether the action launches over its sorted thread range."""
    along = label.startswith("along-step")
    post = label == "geo-boundary"
    if track_order == "reindex_both_action":
        return along or post
    if track_order == "reindex_along_step_action":
        return along
    if track_order == "reindex_step_limit_action":
        return post
    return False


def roofline(amd, params, primaries, slots, track_order):
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
    if sorted_launch(label, track_order):
        # Launched over its own thread range: one 4-byte slot index per
        # track instead of scanning every slot's status and action
        algorithmic = (per_unit + 4) * units
    else:
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

    if args.workload == "cylinder":
        sample = wl.cylinder_primaries(args.cpu_primaries, seed=args.seed + 1)
    else:
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
