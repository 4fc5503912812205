"""Cylindrical detector in a solenoid field: geometry, tables and the
oracle loop (CPU only).

- A radial ray crosses every shell with the closed-form thicknesses.
- Element-selection CDFs of compounds are normalised and ordered, and the
  PbWO4 pair-production CDF follows the Z^2-weighted composition.
- A small event deposits most of an electron's energy in the calorimeters
  and the run empties; charged secondaries curl in the 3.8 T field.
"""

import math

import numpy as np
import pytest

import celeritas_amd as amd
from celeritas_amd import workloads as wl
from tests import oracle_bind as ob


@pytest.fixture(scope="module")
def physics():
    return amd.generate_physics(wl.cylinder_physics_spec())


@pytest.fixture(scope="module")
def problem(physics):
    return ob.Problem(wl.cylinder_problem(physics=physics, init_capacity=1 << 18))


def test_radial_trace(problem):
    geo, vol_mat, _ = wl.cylinder_geometry()
    names = geo["universes"][0]["cell_names"]
    vols, dists = problem.trace([0, 0, 0], [1, 0, 0], max_segments=200)
    # Expected shell boundaries along +x
    edges = [0.0, wl.CYL_PIPE_RADIUS]
    for a, b in wl.CYL_TRACKER:
        edges += [a, b]
    edges += list(wl.CYL_ECAL)
    edges.append(wl.CYL_HCAL[0][0])
    for a, b, c in wl.CYL_HCAL:
        edges += [b, c]
    expected = np.diff(edges)
    assert names[vols[0]] == "beam_pipe"
    np.testing.assert_allclose(dists[: len(expected)], expected, rtol=1e-12, atol=1e-12)
    assert names[vols[len(expected)]] == "world"
    # Then out of the world box at x = 400
    assert dists[len(expected)] == pytest.approx(wl.CYL_WORLD_HALF[0] - edges[-1])


def test_axial_trace_leaves_barrel(problem):
    geo, _, _ = wl.cylinder_geometry()
    names = geo["universes"][0]["cell_names"]
    vols, dists = problem.trace([0, 0, 0], [0, 0, 1], max_segments=10)
    assert names[vols[0]] == "beam_pipe"
    assert dists[0] == pytest.approx(wl.CYL_HALF_LENGTH)
    assert names[vols[1]] == "world"
    assert dists[1] == pytest.approx(wl.CYL_WORLD_HALF[2] - wl.CYL_HALF_LENGTH)


def test_element_cdfs(physics):
    mats = physics["materials"]
    conv = next(p for p in physics["processes"] if p["label"] == "conv")
    cdf = conv["models"][0]["element_cdf"]
    for m, entry in enumerate(cdf):
        if len(mats[m]["elements"]) <= 1:
            assert entry is None
            continue
        rows = np.array(entry["cdf"])
        assert rows.shape[0] == len(mats[m]["elements"])
        assert np.all(np.diff(rows, axis=0) >= -1e-15)
        live = rows[-1] > 0
        np.testing.assert_allclose(rows[-1][live], 1.0, rtol=1e-13)
    # PbWO4 at 1 GeV: pair production ~ Z(Z+1) weighted: Pb, W, 4 O
    rows = np.array(cdf[wl.PBWO4]["cdf"])
    emin, emax = cdf[wl.PBWO4]["emin"], cdf[wl.PBWO4]["emax"]
    e = np.exp(np.linspace(math.log(emin), math.log(emax), rows.shape[1]))
    i = np.searchsorted(e, 1e3)
    z = np.array([82.0, 74.0, 8.0]) * np.array([83.0, 75.0, 9.0]) * np.array([1, 1, 4])
    approx = np.cumsum(z) / z.sum()
    np.testing.assert_allclose(rows[:, i], approx, rtol=0.05)


def test_small_event(problem):
    prim = wl.cylinder_primaries(4, seed=1, emin=500, emax=2000)
    st = ob.State(problem, 4096)
    active, _ = ob.run(st, prim, max_iters=200000)
    assert len(active) < 200000
    edep = st.get_f64("calo_edep")
    total = prim["energy"].sum()
    assert 0.3 * total < edep.sum() < total
