"""The library's ORANGE tracker on the CPU against the oracle.

tests/native/geo_trace.cc runs the tracker functions the device's trace
verifier runs (geo/OrangeLiteTracker.hh: initialize, find_next_step,
midpoint find_safety, move_to_boundary, cross_boundary) on params built by
the library's host code, over random rays in the simple-cms, TestEm3 and
barrel geometries. Volumes must agree exactly and distances and safeties
to 1e-12.

Regression: a crossing moves the position onto the face by a fused
multiply-add, which can leave it an ulp outside the exact bounding box of
the volume being entered; the neighbour search pruned that volume by its
box, found none and failed the track (simple-cms rays entering the barrel
through its z = -700 cm end cap). The boxes are now bumped outward.
"""

import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import celeritas_amd as amd
from celeritas_amd import workloads as wl
from tests import oracle_bind as ob
from tests import problems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "celeritas_amd", "csrc")
NATIVE = os.path.join(ROOT, "tests", "native")
FLAGS = ["-std=c++17", "-O2", "-pthread", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
         "-I/opt/rocm/include", f"-I{CSRC}", f"-I{os.path.join(ROOT, 'include')}"]
MAX_SEG = 256


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    tmp = tmp_path_factory.mktemp("geotrace")
    out = tmp / "geo_trace"
    srcs = []
    for dirpath, _, files in os.walk(CSRC):
        if os.path.basename(dirpath) != "capi":
            srcs += [os.path.join(dirpath, f) for f in files if f.endswith(".cc")]
    cmd = ["g++", *FLAGS, os.path.join(NATIVE, "geo_trace.cc"),
           os.path.join(NATIVE, "hip_host_stubs.cc"), *sorted(srcs),
           "-Wl,--unresolved-symbols=ignore-all", "-o", str(out)]
    subprocess.run(cmd, check=True, capture_output=True, timeout=900)
    return out, tmp


def host_trace(exe, prob, rays):
    path, tmp = exe
    (tmp / "p.json").write_text(json.dumps(prob))
    np.ascontiguousarray(rays, np.float64).tofile(tmp / "r.bin")
    out = subprocess.run([str(path), str(tmp / "p.json"), str(tmp / "r.bin"), str(MAX_SEG)],
                         capture_output=True, check=True, timeout=300).stdout
    rec = 4 + 20 * MAX_SEG
    res = []
    for i in range(len(rays)):
        b = out[i * rec:(i + 1) * rec]
        k = int(np.frombuffer(b[:4], np.uint32)[0])
        res.append((np.frombuffer(b[4:4 + 4 * MAX_SEG], np.uint32)[:k],
                    np.frombuffer(b[4 + 4 * MAX_SEG:4 + 12 * MAX_SEG])[:k],
                    np.frombuffer(b[4 + 12 * MAX_SEG:])[:k]))
    return res


def random_rays(geometry, n, seed, shrink=0.9):
    rng = np.random.default_rng(seed)
    lo, hi = (np.asarray(b, dtype=np.float64) for b in geometry["universes"][0]["bbox"])
    c, h = (lo + hi) / 2, shrink * (hi - lo) / 2
    pos = c + h * (2 * rng.random((n, 3)) - 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.hstack([pos, d])


def compare(exe, prob, rays):
    res = host_trace(exe, prob, rays)
    oprob = ob.Problem(prob)
    for i, r in enumerate(rays):
        vols, dists, safe = oprob.trace(r[:3], r[3:], max_segments=MAX_SEG, safety=True)
        hv, hd, hs = res[i]
        assert np.array_equal(hv, vols), f"ray {i} {r.tolist()}: {hv} != {vols}"
        np.testing.assert_allclose(hd, dists, rtol=1e-12, atol=1e-13, err_msg=f"ray {i}")
        np.testing.assert_allclose(hs, safe, rtol=1e-12, atol=1e-13, err_msg=f"ray {i}")


def test_simple_cms(exe):
    prob = problems.geometry_only("simple-cms.org.json")
    rays = random_rays(prob["geometry"], 2000, seed=11)
    # The rays that failed before the fix (ray 1110 and 1331 of this set)
    # enter the barrel through its end cap
    res = host_trace(exe, prob, rays[[1110, 1331]])
    assert [list(r[0]) for r in res] == [[7, 5, 6, 7], [7, 6, 7]]
    compare(exe, prob, rays)


def test_testem3(exe):
    physics = amd.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics)
    compare(exe, prob, random_rays(prob["geometry"], 1000, seed=12))


def test_cylinder(exe):
    physics = amd.generate_physics(wl.cylinder_physics_spec())
    prob = wl.cylinder_problem(physics=physics)
    compare(exe, prob, random_rays(prob["geometry"], 1000, seed=13))
