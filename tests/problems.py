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
