"""Problem descriptions of the calorimeter workloads.

TestEm3 sampling calorimeter (geometry: data/testem3-flat.org.json, the
reference's test/geocel/data/testem3-flat.org.json): 50 layers along x of
2.3 mm Pb followed by 5.7 mm liquid argon, 40 x 40 cm transverse, from
x = -20 cm, inside a 48 cm vacuum cube. The reference's ORANGE file assigns
gap = lAr and absorber = Pb, while its GDML (testem3-flat.gdml:77-84) has
gap = Pb (2.3 mm) and absorber = lAr (5.7 mm); the GDML assignment is used.

Physics: standard EM (Compton, Livermore photoelectric, Bethe-Heitler
conversion, e+ annihilation, Moller-Bhabha ionisation, combined
Seltzer-Berger/relativistic bremsstrahlung) with tables from the built-in
generator. Production cuts are the energies Geant4 derives from a 0.7 mm
range cut in these materials (TestEm3's default); element shell data are
published binding energies with a synthetic photoelectric shell model.
"""

from __future__ import annotations

import json
import math
import os

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")

ELECTRON, POSITRON, GAMMA = 0, 1, 2


def load_orange(name: str) -> dict:
    with open(os.path.join(_DATA, name)) as f:
        return json.load(f)


def testem3_physics_spec() -> dict:
    """Compact physics description for the table generator."""
    elements = [
        {"z": 82, "atomic_mass": 207.2, "mean_exc_energy": 823e-6,
         "shells": [[88.0045e-3, 0.79], [15.861e-3, 0.05], [15.200e-3, 0.04],
                    [13.0352e-3, 0.06], [3.0e-3, 0.045], [0.5e-3, 0.012], [0.05e-3, 0.003]],
         "pe_reference": [0.1, 1790.0]},
        {"z": 18, "atomic_mass": 39.948, "mean_exc_energy": 188e-6,
         "shells": [[3.2059e-3, 0.85], [0.3263e-3, 0.05], [0.2506e-3, 0.03],
                    [0.2484e-3, 0.06], [0.0158e-3, 0.01]],
         "pe_reference": [0.01, 4140.0]},
        {"z": 1, "atomic_mass": 1.00794, "mean_exc_energy": 19.2e-6,
         "shells": [[13.6e-6, 1.0]],
         "pe_reference": [0.01, 1.0e-2]},
    ]
    materials = [
        {"label": "Pb", "density": 11.35, "elements": [{"element": 0, "fraction": 1.0}]},
        {"label": "lAr", "density": 1.396, "elements": [{"element": 1, "fraction": 1.0}]},
        {"label": "vacuum", "density": 1e-25, "gas": True,
         "elements": [{"element": 2, "fraction": 1.0}]},
    ]
    cutoffs = {
        "gamma": [0.101843, 0.00617835, 990e-6],
        "electron": [1.00386, 0.342545, 990e-6],
        "positron": [0.951321, 0.332994, 990e-6],
    }
    return {"elements": elements, "materials": materials, "cutoffs": cutoffs,
            "options": {"bins_per_decade": 7, "emin": 1e-4, "emax": 1e8}}


def testem3_problem(physics: dict | None = None, seed: int = 20220511,
                    along_step: str = "general_linear", init_capacity: int = 1 << 20,
                    max_events: int = 1 << 16, track_order: str = "none") -> dict:
    """TestEm3 problem; pass expanded ``physics`` to share tables."""
    geo = load_orange("testem3-flat.org.json")
    names = geo["universes"][0]["cell_names"]
    vol_to_mat, vol_to_det = [], []
    for name in names:
        if name.startswith("gap_"):
            vol_to_mat.append(0)
            vol_to_det.append(2 * int(name[4:]))
        elif name.startswith("absorber_"):
            vol_to_mat.append(1)
            vol_to_det.append(2 * int(name[9:]) + 1)
        elif name == "world":
            vol_to_mat.append(2)
            vol_to_det.append(-1)
        else:
            vol_to_mat.append(-1)
            vol_to_det.append(-1)
    return {
        "geometry": geo,
        "volume_to_material": vol_to_mat,
        "volume_to_detector": vol_to_det,
        "physics": physics if physics is not None else {"generate": testem3_physics_spec()},
        "along_step": along_step,
        "seed": seed,
        "track_init": {"capacity": init_capacity, "max_events": max_events,
                       "track_order": track_order},
    }


def testem3_primaries(n: int, seed: int = 12345, emin: float = 1e3, emax: float = 1e5,
                      cone_deg: float = 5.0, gamma_fraction: float = 0.5,
                      first_event: int = 0, per_event: int = 1) -> np.ndarray:
    """Primaries at (-22, 0, 0) cm: e-/gamma mix, log-uniform energies [MeV]
    and directions uniform in a cone about +x; ``per_event`` primaries share
    an event ID."""
    from . import PRIMARY_DTYPE

    rng = np.random.default_rng(seed)
    out = np.zeros(n, dtype=PRIMARY_DTYPE)
    out["particle_id"] = np.where(rng.random(n) < gamma_fraction, GAMMA, ELECTRON)
    out["event_id"] = first_event + np.arange(n) // per_event
    out["energy"] = np.exp(rng.uniform(math.log(emin), math.log(emax), n))
    cos_max = math.cos(math.radians(cone_deg))
    mu = rng.uniform(cos_max, 1.0, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    st = np.sqrt(1 - mu * mu)
    out["dir"][:, 0] = mu
    out["dir"][:, 1] = st * np.cos(phi)
    out["dir"][:, 2] = st * np.sin(phi)
    out["pos"][:, 0] = -22.0
    return out
