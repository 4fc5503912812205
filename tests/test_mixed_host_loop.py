"""The mixed-precision step bodies on the CPU (no GPU).

The kernels' per-track bodies (global/StepExecutors.hh,
global/AlongStepExecutor.hh) compile for the host; tests/native/host_loop.cc
drives them slot by slot over a TestEm3 workload with the kernels'
precision split, timing every call under a watchdog.

This is how the full-scale mixed-precision stall was found: run over all
10,000 primaries of the GPU workload with the earlier arithmetic, one call
never returned. Its replay (the fixture tests/golden/mixed_stall_snapshot.json)
showed a 41.9 keV photon sent to Bethe-Heitler pair production: the 24-bit
draw (k + 1/2) 2^-24 rounds to exactly 1.0 for the top word, the discrete
selection's running sum then ends at -sigma_conv = 0 and falls through to
the last process, pair production below threshold gives a negative lepton
kinetic energy, and the Tsai-Urban loop (u <= umax < 0) can never accept.
Here: the current draw never reaches 1, the selection skips processes
without a cross section, every sampling loop has an iteration budget (the
stalled input now ends as a failed track), and a workload sample runs to
completion with no failed samples.
"""

import json
import os
import shutil
import subprocess

import pytest

import celeritas_amd as amd
from celeritas_amd import workloads as wl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "celeritas_amd", "csrc")
NATIVE = os.path.join(ROOT, "tests", "native")
FLAGS = ["-std=c++17", "-O2", "-pthread", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
         f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}", f"-I{NATIVE}"]


def product_sources():
    out = []
    for dirpath, _, files in os.walk(CSRC):
        if os.path.basename(dirpath) == "capi":
            continue
        out += [os.path.join(dirpath, f) for f in files if f.endswith(".cc")]
    return sorted(out)


@pytest.fixture(scope="module")
def tmp(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    return tmp_path_factory.mktemp("hostloop")


@pytest.fixture(scope="module")
def host_loop(tmp):
    exe = tmp / "host_loop"
    cmd = ["g++", *FLAGS, os.path.join(NATIVE, "host_loop.cc"),
           os.path.join(NATIVE, "host_loop_mixed.cc"),
           os.path.join(NATIVE, "hip_host_stubs.cc"), *product_sources(),
           "-Wl,--unresolved-symbols=ignore-all", "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=900)
    return exe


def write_testem3_problem(tmp, precision="mixed", max_events=10000):
    physics = amd.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics, init_capacity=1 << 20, max_events=max_events,
                              track_order="reindex_both_action", precision=precision)
    path = tmp / f"testem3_{precision}.json"
    path.write_text(json.dumps(prob))
    return path


def test_mixed_building_blocks(tmp):
    exe = tmp / "mixed_checks"
    subprocess.run(["g++", *FLAGS, os.path.join(NATIVE, "mixed_checks.cc"), "-o", str(exe)],
                   check=True, capture_output=True, timeout=600)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 6 and all(l.startswith("ok ") for l in lines), r.stdout
    assert r.returncode == 0


def test_stalled_input_ends_as_failed_track(host_loop, tmp):
    fixture = json.loads(open(os.path.join(ROOT, "tests", "golden",
                                           "mixed_stall_snapshot.json")).read())
    prob = write_testem3_problem(tmp, max_events=fixture["problem"]["max_events"])
    line = json.dumps(fixture["snapshot"])
    r = subprocess.run([str(host_loop), str(prob), "--replay", line, "1", "mixed"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "model class 2" in r.stderr  # Bethe-Heitler
    after = json.loads(r.stdout.strip().splitlines()[-1])
    assert after["status"] == 3  # errored, sent to the tracking cut
    assert after["post"] != fixture["snapshot"]["post"]


def test_mixed_workload_sample_completes(host_loop, tmp):
    prob = write_testem3_problem(tmp)
    prim = tmp / "primaries.bin"
    wl.testem3_primaries(200, seed=4242).tofile(prim)
    r = subprocess.run([str(host_loop), str(prob), str(prim), "16384", "4", "mixed", "20"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    summary = json.loads(r.stdout.strip().splitlines()[-1])
    assert summary["summary"] and summary["mode"] == "mixed"
    assert summary["sampling_failures"] == 0
    assert not summary["hit_iteration_limit"]
    # About 1.4e5 track-steps per primary of 1-100 GeV
    assert summary["track_steps"] > 200 * 5e4
