#!/usr/bin/env python3
"""Throughput of the per-step transport loop on the TestEm3 calorimeter.

Workload (the FP64 GPU-loop configuration): 10,000 primaries, e-/gamma
50/50, energies log-uniform in 1-100 GeV, directions uniform in a 5 degree
cone about +x, started at (-22, 0, 0) cm in front of 50 layers of
2.3 mm Pb + 5.7 mm liquid argon (40 x 40 cm), 2^20 track slots, per-layer
energy deposition tallied. One bench step transports the whole set of
primaries to completion through the HIP stepping loop (one primary per
event), with the step sequence replayed from a captured hipGraph, several
steps per counter readback and the workgroup size chosen by the autotuner
on a sample. Inputs are synthetic; physics tables come from the built-in
generator.

Prints one JSON line with the metric (track-steps per second over the
timed steps; events/s beside it) and, beside it: the same workload in the
mixed-precision mode with its per-detector deposition against the FP64
loop (chi^2), the barrel detector in a 3.8 T field, the served stream
(mixed precision, Poisson arrivals at swept rates, p50/p99 latency,
against the plain FP64 loop over the same stream), a roofline entry for
the dominant kernel and the whole step, and the CPU baseline (the oracle,
the C++ restatement of the same loop, on OpenMP threads over events).
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
#: Bytes of one discrete selection made inside an along-step kernel: per-
#: process cross sections (4 x 8), total cross section 8, RNG state 24 + 24,
#: element 1, post-step action 1, interaction MFP 8 (the particle, energy
#: and material are already in registers)
DISCRETE_SELECT_BYTES = 98

#: Action whose byte model applies to a kernel (kernel registry names)
KERNEL_ACTION = {
    "along-step-linear": "along-step-general-linear",
    "along-step-field": "along-step-uniform-msc",
    "pre-step": "pre-step",
    "along-step-neutral": "along-step-neutral",
    "boundary": "geo-boundary",
}

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
    ap.add_argument("--steps-per-sync", type=int, default=4,
                    help="steps launched (graph replays) per counter readback")
    ap.add_argument("--serial-actions", action="store_true",
                    help="launch every action on the state's stream (no concurrent "
                         "groups of track-disjoint actions)")
    ap.add_argument("--no-graph", action="store_true",
                    help="enqueue every step's kernels instead of replaying a graph")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default workgroup size (256)")
    ap.add_argument("--cpu-primaries", type=int, default=160,
                    help="primaries of the same workload in the CPU baseline (0: skip)")
    ap.add_argument("--cpu-config1", type=int, default=1000,
                    help="10 GeV e- primaries of the CPU-runnable configuration "
                         "(4,096 slots per thread; 0: skip)")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-precision measurement")
    ap.add_argument("--cylinder-primaries", type=int, default=10000,
                    help="primaries of the barrel-detector line (0: skip)")
    ap.add_argument("--cylinder-seeds", type=int, default=10,
                    help="transports (seeds) of the barrel detector in each precision, "
                         "for the per-volume deposition comparison")
    ap.add_argument("--serve-events", type=int, default=64,
                    help="events of the served TestEm3 workload (0: skip serving)")
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


def host_threads():
    """CPU threads for the host baseline: the job's share (OMP_NUM_THREADS
    where set), else the machine's cores."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return os.cpu_count() or 1


def autotune_block(amd, problem, primaries, slots, args):
    """Workgroup size of the step kernels from short graph-replayed
    transports of a sample of the workload (celeritas_amd.autotune)."""
    from celeritas_amd import autotune

    if args.no_autotune:
        return 256, []
    sample = primaries[: min(len(primaries), 500)]
    trials = []
    for bs in autotune.BLOCK_SIZES:
        rate = autotune.measure(amd, problem, sample, args.track_order, bs, slots,
                                use_graph=not args.no_graph)
        trials.append({"block_size": bs, "track_steps_per_s": rate})
        log(f"autotune block {bs}: {rate:.4g} track-steps/s")
    best = max(trials, key=lambda t: t["track_steps_per_s"])["block_size"]
    return best, trials


def make_stepper(amd, problem, slots, block_size, args):
    stepper = amd.Stepper(amd.CoreParams(problem), slots)
    stepper.state.set_block_size(block_size)
    if not args.no_graph:
        stepper.capture_step()
    return stepper


def timed_transports(stepper, primaries, args, ndet, label):
    """Warm-up and timed full transports; per-transport detector tallies."""
    spk = max(args.steps_per_sync, 1)
    for w in range(args.warmup):
        r = stepper.run(primaries, max_steps=20_000, steps_per_sync=spk)
        log(f"{label} warmup {w}: {r}")
    stepper.state.synchronize()
    tallies = []
    totals = []
    prev = stepper.state.get("calo_edep", ndet).copy()
    t0 = time.perf_counter()
    for k in range(args.steps):
        totals.append(stepper.run(primaries, max_steps=20_000, steps_per_sync=spk))
        # Per-transport detector tallies (one 800-byte readback per transport)
        cur = stepper.state.get("calo_edep", ndet).copy()
        tallies.append(cur - prev)
        prev = cur
        log(f"{label} step {k}: {totals[-1]} at {time.perf_counter() - t0:.2f}s")
    stepper.state.synchronize()
    elapsed = time.perf_counter() - t0
    return totals, elapsed, np.array(tallies)


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
        primaries = wl.cylinder_primaries(args.primaries, seed=args.seed)
        ndet = max(wl.cylinder_geometry()[2]) + 1
    else:
        physics = amd.generate_physics(wl.testem3_physics_spec())
        make_problem = wl.testem3_problem
        primaries = wl.testem3_primaries(args.primaries, seed=args.seed)
        ndet = 100

    def problem_of(precision="fp64"):
        prob = make_problem(physics=physics, init_capacity=1 << 26,
                            max_events=max(args.primaries, 1),
                            track_order=args.track_order, precision=precision)
        prob["concurrent_actions"] = not args.serial_actions
        return prob

    problem = problem_of()
    block_size, tune_trials = autotune_block(amd, problem, primaries, args.slots, args)
    stepper = make_stepper(amd, problem, args.slots, block_size, args)
    log(f"state: {args.slots} slots, workgroup {block_size}, "
        f"{'graph replay' if not args.no_graph else 'eager'}, "
        f"{args.steps_per_sync} steps per readback")
    totals, elapsed, tallies = timed_transports(stepper, primaries, args, ndet, "fp64")

    track_steps = sum(t["track_steps"] for t in totals)
    value = track_steps / elapsed
    events_per_s = args.steps * args.primaries / elapsed
    if args.workload == "cylinder":
        workload = ("Cylindrical detector (Si tracker 10 x 0.3 mm, PbWO4 ECAL 23 cm, "
                    "HCAL 20 x (5 cm brass + 4 mm scintillator)) in a uniform 3.8 T "
                    f"field, {args.primaries} primaries e-/e+/gamma 40/20/40, "
                    "log-uniform 0.5-50 GeV, isotropic, 2^20 slots, FP64, "
                    "Dormand-Prince propagation")
    else:
        workload = ("TestEm3 50 x (2.3 mm Pb + 5.7 mm lAr), 10000 primaries e-/gamma "
                    "50/50, log-uniform 1-100 GeV, 5 deg cone, 2^20 slots, FP64")

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
            "graph_replay": not args.no_graph,
            "concurrent_actions": not args.serial_actions,
            "steps_per_readback": args.steps_per_sync,
            "workgroup_size": block_size,
            "autotune_trials": tune_trials,
            "physics": "Compton, Livermore PE, Bethe-Heitler, e+ annihilation, "
                       "Moller-Bhabha ionisation, SB/relativistic brems, Urban MSC, "
                       "Urban energy-loss fluctuations",
            "edep_mev_per_transport": float(tallies.sum()) / args.steps,
        },
    }

    if not args.no_mixed:
        line["mixed"] = mixed_run(amd, problem_of("mixed"), primaries, args, block_size,
                                  ndet, value, events_per_s, tallies)
    if args.cylinder_primaries > 0 and args.workload == "testem3":
        line["cylinder"] = cylinder_run(amd, wl, args)
    if args.serve_events > 0 and args.workload == "testem3":
        line["serve"] = serve_run(amd, wl, physics, args)
    if not args.no_roofline:
        line["roofline"] = roofline(amd, problem, primaries, args.slots, args.track_order,
                                    block_size, value, max(args.steps_per_sync, 1))
    if args.cpu_primaries > 0 or args.cpu_config1 > 0:
        line["cpu_baseline"] = cpu_baseline(amd, wl, problem, physics, args)
    print(json.dumps(line), flush=True)


def mixed_run(amd, problem, primaries, args, block_size, ndet, fp64_value, fp64_events,
              fp64_tally):
    """Same workload with the physics kernels in the mixed-precision mode,
    and its per-detector deposition against the FP64 loop's."""
    stepper = make_stepper(amd, problem, args.slots, block_size, args)
    totals, dt, tallies = timed_transports(stepper, primaries, args, ndet, "mixed")
    steps = sum(t["track_steps"] for t in totals)
    value = steps / dt
    events = args.steps * args.primaries / dt
    log(f"mixed precision: {value:.4g} track-steps/s, {events:.4g} events/s")
    fails = stepper.state.counters()["num_sampling_failures"]
    out = {"value": value, "unit": "track-steps/s", "events_per_s": events,
           "ms_per_step": 1e3 * dt / args.steps,
           "speedup_events_vs_fp64": events / fp64_events,
           "sampling_failures": int(fails),
           "dtype": "f32 transcendentals and draws in sampling, interpolation and MSC; "
                    "f64 positions, safety, field, energies and tallies"}
    # Per-detector deposition against the FP64 loop (itself checked against
    # the FP64 oracle by the parity tests)
    cmp = edep_vs_fp64(tallies, fp64_tally)
    if cmp:
        out["edep_vs_fp64"] = cmp
    return out


def edep_vs_fp64(tallies, fp64_tally):
    """Per-volume deposition of mixed-precision transports against FP64
    ones: each transport of the same primaries is an independent sample
    (the RNG streams continue), so the spread over transports gives each
    volume's standard error. With a few transports per side each volume's
    difference is a Welch t statistic of few degrees of freedom (its square
    averages about 2, not 1), so each t is mapped to the normal quantile of
    the same probability before the squares are summed: chi^2 over the
    volumes with signal, ndf = their number."""
    from scipy import stats

    if len(tallies) < 2 or len(fp64_tally) < 2:
        return None
    nm, nf = len(tallies), len(fp64_tally)
    mm, ms = tallies.mean(0), tallies.std(0, ddof=1) / np.sqrt(nm)
    fm, fs = fp64_tally.mean(0), fp64_tally.std(0, ddof=1) / np.sqrt(nf)
    sel = (ms > 0) & (fs > 0) & (mm + fm > 1e-5 * (mm + fm).max())
    var = ms[sel] ** 2 + fs[sel] ** 2
    t = (mm[sel] - fm[sel]) / np.sqrt(var)
    dof = var ** 2 / (ms[sel] ** 4 / (nm - 1) + fs[sel] ** 4 / (nf - 1))
    z = stats.norm.isf(stats.t.sf(t, dof))
    chi2 = float((z ** 2).sum())
    ndf = int(sel.sum())
    tot_err = float(np.sqrt((ms ** 2).sum() + (fs ** 2).sum()) / fm.sum())
    rel = (mm[sel] - fm[sel]) / fm[sel]
    out = {
        "total_rel": float(mm.sum() / fm.sum() - 1), "total_rel_err": tot_err,
        "chi2": chi2, "ndf": ndf, "p_value": float(stats.chi2.sf(chi2, ndf)),
        "max_volume_rel": float(np.max(np.abs(rel))),
        "per_volume_rel": [round(float(x), 5) for x in rel],
        "per_volume_rel_err": [round(float(x), 5) for x in
                               np.sqrt(ms[sel] ** 2 + fs[sel] ** 2) / fm[sel]],
        "samples": f"{len(tallies)} mixed and {len(fp64_tally)} FP64 transports of the "
                   "same primaries (independent RNG streams)",
        "statistic": "per-volume Welch t mapped to normal quantiles, squared and summed"}
    log(f"mixed vs FP64 edep: total {out['total_rel']:+.2e} +- {tot_err:.1e}, "
        f"chi2 {chi2:.1f}/{ndf}")
    return out


def cylinder_run(amd, wl, args):
    """The barrel detector in a 3.8 T field (FP64, Dormand-Prince, graph
    replay, autotuned workgroup size)."""
    physics = amd.generate_physics(wl.cylinder_physics_spec())
    n = args.cylinder_primaries
    primaries = wl.cylinder_primaries(n, seed=args.seed)
    problem = wl.cylinder_problem(physics=physics, init_capacity=1 << 26,
                                  max_events=max(n, 1), track_order=args.track_order)
    sub = argparse.Namespace(**vars(args))
    # Each timed transport is one seed of the per-volume deposition sample
    sub.steps, sub.warmup = max(2, args.cylinder_seeds), 1
    block, trials = autotune_block(amd, problem, primaries, args.slots, sub)
    stepper = make_stepper(amd, problem, args.slots, block, sub)
    ndet = max(wl.cylinder_geometry()[2]) + 1
    totals, dt, fp64_tally = timed_transports(stepper, primaries, sub, ndet, "cylinder")
    del stepper
    steps = sum(t["track_steps"] for t in totals)
    log(f"cylinder: {steps / dt:.4g} track-steps/s, {sub.steps * n / dt:.4g} events/s")
    out = {"workload": f"Cylindrical detector (Si tracker, PbWO4 ECAL, brass/scintillator "
                       f"HCAL) in a uniform 3.8 T field, {n} primaries e-/e+/gamma "
                       "40/20/40, log-uniform 0.5-50 GeV, isotropic, 2^20 slots, FP64, "
                       "Dormand-Prince, graph replay, autotuned workgroup",
           "value": steps / dt, "unit": "track-steps/s",
           "events_per_s": sub.steps * n / dt, "timed_transports": sub.steps,
           "workgroup_size": block, "autotune_trials": trials}
    if args.no_mixed:
        return out
    # The same detector, field and source in mixed precision, one transport
    # per seed, with each volume's deposition against the FP64 transports
    mixed = dict(problem, precision="mixed")
    stepper = make_stepper(amd, mixed, args.slots, block, sub)
    totals, dt, tallies = timed_transports(stepper, primaries, sub, ndet, "cylinder mixed")
    fails = stepper.state.counters()["num_sampling_failures"]
    del stepper
    steps = sum(t["track_steps"] for t in totals)
    log(f"cylinder mixed: {steps / dt:.4g} track-steps/s, {sub.steps * n / dt:.4g} events/s")
    out["mixed"] = {"value": steps / dt, "unit": "track-steps/s",
                    "events_per_s": sub.steps * n / dt,
                    "speedup_events_vs_fp64": (sub.steps * n / dt) / out["events_per_s"],
                    "sampling_failures": int(fails),
                    "edep_vs_fp64": edep_vs_fp64(tallies, fp64_tally)}
    return out


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


#: Per-slot fields of the SoA state carried from one step to the next, in
#: bytes, and whether a step writes them back (global/CoreData.hh,
#: geo/GeoData.hh, phys/PhysicsData.hh, random/XorwowEngine.hh). Fields a
#: step computes and consumes itself (cross sections, step length and
#: actions, deposition, secondary span, element) are not persistent: in a
#: step fused into one kernel they would stay in registers.
PERSISTENT_FIELDS = [
    # (field, bytes, written back by a step)
    ("sim.track_id", 4, False), ("sim.parent_id", 4, False), ("sim.event_id", 4, False),
    ("sim.num_steps", 4, True), ("sim.time", 8, True), ("sim.status", 1, True),
    ("particle.particle_id", 1, False), ("particle.energy", 8, True),
    ("geo.pos", 24, True), ("geo.dir", 24, True), ("geo.volume", 4, True),
    ("geo.surface", 4, True), ("geo.sense", 1, True), ("geo.boundary", 1, True),
    ("geo.failed", 1, False), ("geo.next_step", 8, True), ("geo.next_surface", 4, True),
    ("geo.next_sense", 1, True), ("phys.interaction_mfp", 8, True),
    ("phys.msc_range", 24, True), ("rng.xorwow", 24, True),
]
#: Per new track: the secondary written to the stack and read back (40 B
#: each way) and the initializer written and read (80 B each way)
BYTES_PER_NEW_TRACK = 2 * 40 + 2 * 80


def step_bytes(new_tracks_per_step):
    """Algorithmic HBM bytes of one track-step: every persistent field read
    once and, if the step changes it, written once, plus the bytes of the
    tracks it creates (new tracks per track-step)."""
    persistent = sum(b * (2 if w else 1) for _, b, w in PERSISTENT_FIELDS)
    return persistent + BYTES_PER_NEW_TRACK * new_tracks_per_step, persistent


def roofline(amd, problem, primaries, slots, track_order, block_size, headline,
             steps_per_sync):
    """Dominant kernel's algorithmic bandwidth from an instrumented run, and
    the whole step's algorithmic bandwidth at the headline rate.

    The instrumented run is the same transport with eager launches, the
    actions one after another on the state's stream (as the rocprofv3
    kernel trace in profiles/ runs it), and every kernel launch carrying
    HIP start and stop events that the runtime sets at the dispatch's own
    start and end (hipExtLaunchKernel), so a kernel's time is its
    dispatch's duration."""
    st = amd.Stepper(amd.CoreParams(dict(problem, concurrent_actions=False)), slots)
    st.state.set_block_size(block_size)
    st.diagnostics(enable=True)
    amd.kernel_times(reset=True)
    amd.kernel_timing(True)
    try:
        r = st.run(primaries, steps_per_sync=steps_per_sync)
        kt = amd.kernel_times(reset=True)
    finally:
        amd.kernel_timing(False)
    times = {}
    for t in kt:
        secs, n = times.get(t["name"], (0.0, 0))
        times[t["name"]] = (secs + t["seconds"], n + int(t["launches"]))
    diag = st.diagnostics()
    new_tracks = int(st.state.get("track_counters", len(primaries)).sum())
    ranked = sorted(times.items(), key=lambda kv: -kv[1][0])
    log("kernel device times [s]: " + ", ".join(f"{k}={v[0]:.4f}" for k, v in ranked))
    log(f"work counters: {diag}")
    kname, (secs, launches) = next((kv for kv in ranked if kv[0] in KERNEL_ACTION),
                                   ranked[0])
    label = KERNEL_ACTION[kname]
    per_unit, per_slot, counter = BYTES_PER_UNIT[label]
    units = diag[counter]
    if sorted_launch(label, track_order):
        # Launched over its own thread range: one 4-byte slot index per
        # track instead of scanning every slot's status and action
        algorithmic = (per_unit + 4) * units
    else:
        algorithmic = per_unit * units + per_slot * slots * launches
    if kname == "along-step-linear" or kname == "along-step-field":
        # The discrete selections the kernel makes for the tracks whose step
        # ends in one: per-process and total cross sections, MFP and RNG
        # state read, element, action and MFP written
        algorithmic += DISCRETE_SELECT_BYTES * diag.get("discrete_in_charged", 0)
    per_launch = algorithmic / launches
    avg_launch = secs / launches
    achieved = per_launch / avg_launch / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get(label)
    # Whole step: tracks created beyond the primaries per track-step
    per_step_new = (new_tracks - len(primaries)) / max(r["track_steps"], 1)
    b_step, persistent = step_bytes(per_step_new)
    whole = headline * b_step / 1e9
    # Registers and occupancy of the kernel as the runtime reports them
    attrs = next((a for a in amd.kernel_attributes()
                  if a["name"] == kname and a["block_size"] == block_size), None)
    total = sum(v[0] for v in times.values())
    return {
        "bound": "hbm",
        "kernel": kname,
        "action": label,
        "kernel_attributes": attrs,
        "achieved": achieved,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": traffic,
        "traffic_source": "profiles/pmc_traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE "
                          "per dispatch of this kernel, scaled by the plain-copy "
                          "calibration (not measured in this run)",
        "algorithmic_bytes_per_launch": per_launch,
        "avg_launch_us": 1e6 * avg_launch,
        "launches": launches,
        "timing": "HIP start/stop events set at each dispatch's start and end "
                  "(hipExtLaunchKernel), eager launches, actions in sequence",
        "share_of_step_time": secs / total,
        "kernel_avg_us": {k: round(1e6 * v[0] / max(v[1], 1), 2) for k, v in ranked[:12]},
        "work_counters": diag,
        "step": {
            "b_step": b_step,
            "persistent_state_bytes": persistent,
            "new_tracks_per_track_step": per_step_new,
            "achieved": whole,
            "frac": whole / HBM_PEAK_GBS,
            "model": "persistent SoA fields read once and written once if changed, "
                     "plus 240 B per created track (secondary and initializer written "
                     "and read), at the headline track-steps/s",
        },
    }


def cpu_baseline(amd, wl, problem, physics, args):
    """The oracle (the C++ restatement of the loop, FP64) on the host's
    cores: events dealt round-robin to OpenMP threads, each with its own
    state (celer-sim's event-parallel mode, app/celer-sim/celer-sim.cc:
    115-137). A sample of the bench workload gives the value; the
    CPU-runnable configuration (1,000 x 10 GeV e-, 4,096 slots per thread)
    runs in full beside it."""
    from tests import oracle_bind as ob

    threads = host_threads()
    oprob = ob.Problem(problem)
    out = {"unit": "track-steps/s", "cores": threads, "kind": "port",
           "threads_note": "OpenMP threads over events; the job's host CPU share "
                           "(OMP_NUM_THREADS), os.cpu_count() = "
                           f"{os.cpu_count()}"}
    if args.cpu_primaries > 0:
        if args.workload == "cylinder":
            sample = wl.cylinder_primaries(args.cpu_primaries, seed=args.seed + 1)
        else:
            sample = wl.testem3_primaries(args.cpu_primaries, seed=args.seed + 1)
        sample["event_id"] = np.arange(len(sample))
        r = ob.run_events_parallel(oprob, sample, 4096, threads)
        out["value"] = r["track_steps"] / r["seconds"]
        out["events_per_s"] = len(sample) / r["seconds"]
        out["sample"] = (f"{len(sample)} primaries of the bench workload's distribution "
                         f"(seed {args.seed + 1}), one event each, {threads} threads x "
                         f"4096 slots; {r['track_steps']} track-steps in "
                         f"{r['seconds']:.1f} s")
        log(f"CPU baseline: {out['value']:.4g} track-steps/s on {threads} threads")
    if args.cpu_config1 > 0 and args.workload == "testem3":
        n = args.cpu_config1
        prim = wl.testem3_primaries(n, seed=args.seed + 2, emin=1e4, emax=1e4,
                                    cone_deg=0.0, gamma_fraction=0.0)
        prim["event_id"] = np.arange(n)
        cprob = ob.Problem(wl.testem3_problem(physics=physics, init_capacity=1 << 20,
                                              max_events=n))
        r = ob.run_events_parallel(cprob, prim, 4096, threads)
        out["config1"] = {
            "workload": f"TestEm3, {n} x 10 GeV e- along +x from (-22, 0, 0) cm, one "
                        f"event each, FP64, {threads} threads x 4096 slots",
            "track_steps_per_s": r["track_steps"] / r["seconds"],
            "events_per_s": n / r["seconds"], "seconds": r["seconds"],
            "track_steps": r["track_steps"]}
        log(f"CPU config 1: {n / r['seconds']:.4g} events/s")
        if "value" not in out:
            out["value"] = out["config1"]["track_steps_per_s"]
    return out


if __name__ == "__main__":
    main()
