"""Readers for the G4EMLOW ASCII data files the reference's tests hold
(copied as fixtures under tests/golden/g4emlow/).

The Seltzer-Berger format follows the reference reader
(src/celeritas/io/SeltzerBergerReader.cc): a header "dummy ny nx", the ny
reduced photon energies, the nx log(E/MeV) values, then nx*ny values in
x-major order.
"""

from __future__ import annotations

import os

import numpy as np

DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "g4emlow")


def read_sb_table(path_or_z) -> dict:
    """Return {"x", "y", "values"} lists for a brNN file."""
    path = path_or_z
    if isinstance(path_or_z, int):
        path = os.path.join(DATA, f"br{path_or_z}")
    with open(path) as f:
        tok = f.read().split()
    _, ny, nx = (int(t) for t in tok[:3])
    vals = np.array(tok[3:], dtype=np.float64)
    if vals.size != ny + nx + nx * ny:
        raise ValueError(f"inconsistent SB file {path}")
    y = vals[:ny]
    x = vals[ny:ny + nx]
    v = vals[ny + nx:]
    return {"x": x.tolist(), "y": y.tolist(), "values": v.tolist()}
