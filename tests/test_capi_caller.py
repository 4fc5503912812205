"""A plain C program compiled against include/celeritas_amd.h and linked
with the library (tests/native/capi_caller.c): the header is valid C99,
the host-only entry points work from C, errors come back as codes with a
message, and (on a GPU) the SimpleCompton loop runs to the reference's
anchor (test/celeritas/global/Stepper.test.cc:194-225: 919 iterations,
53.8125 steps per primary for 32 primaries) through the C ABI alone."""

import json
import os
import shutil
import subprocess

import pytest

import celeritas_amd as amd
from tests import g4emlow
from tests import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def caller(tmp_path_factory):
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    tmp = tmp_path_factory.mktemp("capi")
    exe = tmp / "capi_caller"
    libdir = os.path.dirname(amd.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
                    f"-I{amd.INCLUDE_DIR}", os.path.join(ROOT, "tests", "native", "capi_caller.c"),
                    f"-L{libdir}", "-lceleritas_amd", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                   check=True, capture_output=True, timeout=300)
    return exe, tmp


def test_c_caller_host_entry_points(caller):
    exe, _ = caller
    r = subprocess.run([str(exe), g4emlow.DATA], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
    assert out["version"] == "0.1.0"
    assert out["sizes"] == f"{amd.PRIMARY_DTYPE.itemsize} 16 16"
    assert out["sb_json"].split(" ", 1)[1].startswith('{"values":')
    assert "1 <= Z <= 100" in out["error_message"]
    if not amd.device_available():
        assert out["no_device"].startswith("1 ") and "no HIP device" in out["no_device"]


@pytest.mark.gpu
def test_c_caller_runs_simple_compton(caller):
    exe, tmp = caller
    path = tmp / "simple_compton.json"
    path.write_text(json.dumps(problems.simple_compton()))
    r = subprocess.run([str(exe), g4emlow.DATA, str(path), "32"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    last = r.stdout.strip().splitlines()[-1].split()
    assert last[0] == "iterations" and int(last[1]) == 919
    assert int(last[3]) / 32 == 53.8125
