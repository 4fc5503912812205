"""Oracle geometry against ORANGE golden traces (CPU)."""

import json
import os

import numpy as np
import pytest

from tests import oracle_bind as ob
from tests import problems

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "orange.json")))


@pytest.fixture(scope="module")
def testem3():
    doc = problems.load_geometry("testem3-flat.org.json")
    return doc, ob.Problem({"geometry": doc})


def test_testem3_trace(testem3):
    doc, prob = testem3
    g = GOLDEN["testem3_trace"]
    vols, dists, safe = prob.trace(g["pos"], g["dir"], safety=True)
    names = doc["universes"][0]["cell_names"]
    assert [names[v] for v in vols] == g["volumes"]
    np.testing.assert_allclose(dists, g["distances"], rtol=1e-12)
    np.testing.assert_allclose(safe, g["halfway_safety"], rtol=1e-12)


def test_testem3_transverse(testem3):
    # Closed form: from inside a layer along +y, exit through y = 20 into the
    # world, then world boundary y = 24.
    doc, prob = testem3
    vols, dists = prob.trace([-19.9, 0, 0], [0, 1, 0])
    names = doc["universes"][0]["cell_names"]
    assert [names[v] for v in vols] == ["gap_0", "world"]
    np.testing.assert_allclose(dists, [20.0, 4.0], rtol=1e-14)


def test_two_boxes_oblique():
    doc = problems.load_geometry("two-boxes.org.json")
    prob = ob.Problem({"geometry": doc})
    d = np.array([1.0, 1.0, 0.0]) / np.sqrt(2)
    vols, dists = prob.trace([-1, 0, 0], d)
    names = doc["universes"][0]["cell_names"]
    assert [names[v] for v in vols] == ["inner", "world"]
    # Exit inner box at y = 5 (distance 5*sqrt2), world at y = 500
    np.testing.assert_allclose(dists, [5 * np.sqrt(2), 495 * np.sqrt(2)], rtol=1e-13)
