"""Mixed-precision samplers on the CPU (no GPU): no rejection loop fails to
end and no sampler returns a non-finite or out-of-range result.

The physics headers are compiled for the host in the mixed mode (FP32
transcendentals and 23-bit draws, as on the device) together with the
table builder, and driven over random inputs of the TestEm3 physics by
``tests/native/mixed_sampler_fuzz.cc``; a loop that runs out of its
iteration budget counts as a failure. Longer runs of the same program
(13 million sampler calls over four seeds) found one defect of the FP32
arithmetic, now fixed: a cosine rounded just outside [-1, 1] gave a NaN
direction (in-flight annihilation at GeV energies, about once in 3 million
calls); this test is the short form that runs with the CPU suite. The
samplers are only driven inside their domains (a process is sampled only
where its cross section is positive); the whole-loop check that also
covers the selection is tests/test_mixed_host_loop.py.
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


@pytest.fixture(scope="module")
def fuzz_exe(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    out = tmp_path_factory.mktemp("fuzz") / "mixed_sampler_fuzz"
    srcs = [os.path.join(ROOT, "tests", "native", "mixed_sampler_fuzz.cc")] + [
        os.path.join(CSRC, f) for f in ("phys/PhysicsParams.cc", "phys/PhysicsActions.cc",
                                        "global/ActionRegistry.cc", "corecel/Json.cc",
                                        "corecel/DeviceRuntime.cc")]
    cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}", *srcs,
           "-Wl,--unresolved-symbols=ignore-all", "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    return out


def test_mixed_samplers_terminate_and_stay_finite(fuzz_exe, tmp_path):
    physics = amd.generate_physics(wl.testem3_physics_spec())
    prob = tmp_path / "testem3.json"
    prob.write_text(json.dumps(wl.testem3_problem(physics=physics)))
    r = subprocess.run([str(fuzz_exe), str(prob), "60000", "6"], capture_output=True,
                       text=True, timeout=600)
    print(r.stdout[-2000:])
    last = r.stdout.strip().splitlines()[-1]
    assert last.startswith("calls"), r.stdout[-2000:] + r.stderr[-2000:]
    fields = dict(zip(last.split()[::2], map(int, last.split()[1::2])))
    assert fields["calls"] > 150_000
    assert fields["loops"] == 0
    assert fields["bad"] == 0
    assert r.returncode == 0
