"""Problem descriptions shared by tests, smoke() and bench.py.

Each function returns the JSON-able dict given to both the product
(celeritas_amd.CoreParams) and the oracle (tests/oracle_bind.Problem).
"""

import json
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_geometry(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


GAMMA = {"label": "gamma", "pdg": 22, "mass": 0.0, "charge": 0.0}
ELECTRON_HALF = {"label": "electron", "pdg": 11, "mass": 0.5, "charge": -1.0}


def two_boxes_track_init(stack_factor=8.0):
    """SimpleTestBase: two-boxes geometry, gamma/e- (SimpleTestBase.cc:32-200)."""
    return {
        "geometry": load_geometry("two-boxes.org.json"),
        "volume_to_material": [-1, 0, 1],
        "physics": {
            "particles": [GAMMA, ELECTRON_HALF],
            "options": {"secondary_stack_factor": stack_factor},
        },
        "track_init": {"capacity": 4096, "max_events": 4096, "track_order": "none"},
        "seed": 20220511,
    }


def simple_compton():
    """SimpleComptonTest: two boxes, Compton-only gamma physics with fake
    lambda tables (test/celeritas/SimpleTestBase.cc:32-200,
    test/celeritas/global/Stepper.test.cc:46-68)."""
    return {
        "geometry": load_geometry("two-boxes.org.json"),
        "volume_to_material": [-1, 0, 1],
        "physics": {
            "particles": [GAMMA, ELECTRON_HALF],
            "elements": [{"z": 13, "atomic_mass": 27.0}],
            "materials": [
                {"label": "Al", "elements": [{"element": 0, "fraction": 1.0}]},
                {"label": "hard vacuum", "elements": []},
            ],
            "cutoffs": {
                "gamma": [{"energy": 0.01, "range": 0.01}, {"energy": 100, "range": 100}],
                "electron": [{"energy": 1000, "range": 1000}, {"energy": 1000, "range": 1000}],
            },
            "options": {"secondary_stack_factor": 3.0},
            "processes": [
                {
                    "label": "Compton scattering",
                    "particle": "gamma",
                    "class": "compton",
                    "integral": False,
                    "models": [{"label": "scat-klein-nishina", "model": "klein_nishina",
                                "low": 1e-4, "high": 1e8}],
                    "lambda": [{"x": [1e-4, 1.0], "y": [1e1, 1e0]},
                               {"x": [1e-4, 1.0], "y": [1e-10, 1e-10]}],
                    "lambda_prim": [{"x": [1.0, 1e4, 1e8], "y": [1e0, 1e-2, 1e-4]},
                                    {"x": [1.0, 1e4, 1e8], "y": [1e-10, 1e-10, 1e-10]}],
                }
            ],
        },
        "along_step": "neutral",
        "track_init": {"capacity": 4096, "max_events": 4096, "track_order": "none"},
        "seed": 20220511,
    }


def simple_compton_primaries(n):
    import celeritas_amd

    return celeritas_amd.make_primaries(
        {"particle_id": 0, "energy": 100.0, "pos": [-22.0, 0, 0], "dir": [1.0, 0, 0]}
        for _ in range(n))


EMASS = 0.5109989461


def interactor_host():
    """InteractorHostTestBase particles (electron mass 0.5109989461 MeV) with
    Cu (Z=29) and K (Z=19) and a 1 keV electron cutoff."""
    return {
        "geometry": load_geometry("two-boxes.org.json"),
        "volume_to_material": [-1, 0, 0],
        "physics": {
            "particles": [
                {"label": "electron", "pdg": 11, "mass": EMASS, "charge": -1.0},
                {"label": "positron", "pdg": -11, "mass": EMASS, "charge": 1.0},
                {"label": "gamma", "pdg": 22, "mass": 0.0, "charge": 0.0},
            ],
            "elements": [{"z": 29, "atomic_mass": 63.546}, {"z": 19, "atomic_mass": 39.0983}],
            "materials": [
                {"label": "Cu", "number_density": 6.02214076e23, "electron_density": 29 * 6.02214076e23,
                 "elements": [{"element": 0, "fraction": 1.0}]},
                {"label": "K", "number_density": 6.02214076e23, "electron_density": 19 * 6.02214076e23,
                 "elements": [{"element": 1, "fraction": 1.0}]},
            ],
            "cutoffs": {"electron": [{"energy": 0.001}, {"energy": 0.001}]},
        },
    }


ELECTRON, POSITRON, GAMMA_ID = 0, 1, 2


def geometry_only(name):
    """A geometry with particle definitions and no physics, for tracking
    checks: every volume but the exterior maps to material 0."""
    geom = load_geometry(name)
    nvol = len(geom["universes"][0]["cell_names"])
    return {
        "geometry": geom,
        "volume_to_material": [-1] + [0] * (nvol - 1),
        "physics": {"particles": [GAMMA, ELECTRON_HALF]},
        "track_init": {"capacity": 1024, "max_events": 16, "track_order": "none"},
        "seed": 1,
    }
