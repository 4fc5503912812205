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


def _trace_compare(amd, prob, rays, min_segments):
    """Device ray traces against the oracle's: volumes exact, distances and
    midpoint safeties to 1e-12 relative."""
    params = amd.CoreParams(prob)
    dev = amd_pkg.verify_geo_trace(params, rays, max_segments=256)
    oprob = ob.Problem(prob)
    nseg = 0
    for i, r in enumerate(rays):
        vols, dists, safe = oprob.trace(r[:3], r[3:], max_segments=256, safety=True)
        dv, dd, ds = dev[i]
        assert np.array_equal(dv, vols), f"ray {i} {r}: volumes {dv} != {vols}"
        np.testing.assert_allclose(dd, dists, rtol=1e-12, atol=1e-13, err_msg=f"ray {i}")
        np.testing.assert_allclose(ds, safe, rtol=1e-12, atol=1e-13, err_msg=f"ray {i}")
        nseg += len(vols)
    assert nseg >= min_segments
    return dev


def _random_rays(geometry, n, seed, shrink=0.9):
    rng = np.random.default_rng(seed)
    lo, hi = (np.asarray(b, dtype=np.float64) for b in geometry["universes"][0]["bbox"])
    c, h = (lo + hi) / 2, shrink * (hi - lo) / 2
    pos = c + h * (2 * rng.random((n, 3)) - 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.hstack([pos, d])


def test_simple_cms_trace_golden(amd):
    """SimpleCmsGeantTest.trace (orange/OrangeGeant.test.cc:170-200) on the
    device; the last safety of the second ray is the .org.json world
    cell's (650), as the oracle test explains."""
    prob = problems.geometry_only("simple-cms.org.json")
    names = prob["geometry"]["universes"][0]["cell_names"]
    params = amd.CoreParams(prob)
    (v1, d1, s1), (v2, d2, s2) = amd_pkg.verify_geo_trace(
        params, [[-75, 0, 0, 1, 0, 0], [25, 0, 701, 0, 0, -1]])
    assert [names[v] for v in v1] == [
        "si_tracker", "vacuum_tube", "si_tracker", "em_calorimeter", "had_calorimeter",
        "sc_solenoid", "fe_muon_chambers", "world"]
    np.testing.assert_allclose(d1, [45, 60, 95, 50, 100, 100, 325, 300], rtol=1e-12)
    np.testing.assert_allclose(s1, [22.5, 700, 47.5, 25, 50, 50, 162.5, 150], rtol=1e-12)
    assert [names[v] for v in v2] == ["world", "vacuum_tube", "world"]
    np.testing.assert_allclose(d2, [1, 1400, 1300], rtol=1e-12)
    np.testing.assert_allclose(s2, [0.5, 5, 650], rtol=1e-12)


def test_simple_cms_trace_matches_oracle(amd):
    prob = problems.geometry_only("simple-cms.org.json")
    rays = _random_rays(prob["geometry"], 2000, seed=11)
    _trace_compare(amd, prob, rays, min_segments=3500)


def test_testem3_trace_matches_oracle(amd):
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    prob = wl.testem3_problem(physics=physics)
    rays = _random_rays(prob["geometry"], 1000, seed=12)
    # Rays along +x through every layer as well
    y = np.linspace(-19, 19, 50)
    along = np.stack([np.full(50, -23.0), y, -y / 2, np.ones(50), np.zeros(50),
                      np.zeros(50)], axis=1)
    _trace_compare(amd, prob, np.vstack([rays, along]), min_segments=25000)


def test_cylinder_trace_matches_oracle(amd):
    physics = amd_pkg.generate_physics(wl.cylinder_physics_spec())
    prob = wl.cylinder_problem(physics=physics)
    rays = _random_rays(prob["geometry"], 1000, seed=13)
    _trace_compare(amd, prob, rays, min_segments=17000)
