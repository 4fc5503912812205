"""Point location on the device (track initialization's volume search)
against the oracle's exhaustive search.

The device walks a bounding interval hierarchy to choose candidate
volumes; the oracle tests every volume's logic in order. For volumes that
do not overlap both give the same answer for every point, including the
failures: points exactly on a face (rejected) and points outside every
volume. Cases:

- TestEm3 (100 slabs along x in a world box): uniform points over a box
  slightly larger than the world, points placed exactly on every x/y/z
  plane of the geometry, and NaN positions.
- The barrel detector (cylinders and z planes with a background volume):
  the same, with points on the cylinders' radii in the z = 0 plane.
- The two-box problem used by the track-initialization tests.
"""

import numpy as np
import pytest

import celeritas_amd as amd_pkg
from celeritas_amd import workloads as wl
from tests import oracle_bind as ob
from tests import problems

pytestmark = pytest.mark.gpu


def _surfaces(geometry):
    s = geometry["universes"][0]["surfaces"]
    data = np.asarray(s["data"], dtype=np.float64)
    offsets = np.concatenate([[0], np.cumsum(s["sizes"])])
    return [(t, data[offsets[i]:offsets[i + 1]]) for i, t in enumerate(s["types"])]


def _points(geometry, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = (np.asarray(b, dtype=np.float64) for b in geometry["universes"][0]["bbox"])
    span = hi - lo
    pos = lo - 0.05 * span + 1.1 * span * rng.random((n, 3))
    on = []
    for t, d in _surfaces(geometry):
        if t in ("px", "py", "pz"):
            ax = "xyz".index(t[1])
            p = lo + span * rng.random((16, 3))
            p[:, ax] = d[0]
            on.append(p)
        elif t in ("cxc", "cyc", "czc"):
            ax = "xyz".index(t[1])
            phi = 2 * np.pi * rng.random(16)
            r = np.sqrt(d[0])
            p = lo + span * rng.random((16, 3))
            u, v = (ax + 1) % 3, (ax + 2) % 3
            p[:, u] = r * np.cos(phi)
            p[:, v] = r * np.sin(phi)
            on.append(p)
    pos = np.concatenate([pos] + on + [np.full((2, 3), np.nan)])
    d = rng.normal(size=pos.shape)
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.hstack([pos, d])


def _compare(amd, prob, n, seed):
    params = amd.CoreParams(prob)
    pts = _points(prob["geometry"], n, seed)
    dev = amd_pkg.verify_geo_locate(params, pts)
    cpu = ob.Problem(prob).locate(pts)
    bad = np.flatnonzero(dev != cpu)
    assert bad.size == 0, (f"{bad.size} of {len(pts)} differ, first at {pts[bad[0], :3]}: "
                           f"device {dev[bad[0]]}, oracle {cpu[bad[0]]}")
    return cpu


def test_testem3_locate_matches_oracle(amd):
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics)
    vols = _compare(amd, prob, 200_000, seed=1)
    # Most volumes are hit, and the on-plane points fail
    assert len(np.unique(vols)) > 90
    assert np.sum(vols == 0xFFFFFFFF) > 100


def test_cylinder_locate_matches_oracle(amd):
    physics = amd_pkg.generate_physics(wl.cylinder_physics_spec())
    prob = wl.cylinder_problem(physics=physics)
    vols = _compare(amd, prob, 200_000, seed=2)
    assert len(np.unique(vols)) > 20


def test_two_boxes_locate_matches_oracle(amd):
    _compare(amd, problems.two_boxes_track_init(), 20_000, seed=3)
