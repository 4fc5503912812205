"""Independent numpy parsers of the G4EMLOW ASCII data files the
reference's tests hold (copied as fixtures under tests/golden/g4emlow/),
used to cross-check the library's own readers (celeritas_amd
read_seltzer_berger / read_livermore_pe).

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


def read_livermore_pe(z: int) -> dict:
    """Livermore photoelectric data for element z in the problem layout.

    Follows the reference reader (src/celeritas/io/LivermorePEReader.cc):
    pe-cs-Z (total above the K edge: "emin emax n n" then n pairs),
    pe-le-cs-Z (total below the K edge, may be empty), pe-low-Z and
    pe-high-Z ("n n thresh" then per shell: binding and 6 parameters) and
    pe-ss-cs-Z (per shell: "emin emax n id" then n pairs).
    """
    def tokens(name):
        with open(os.path.join(DATA, f"{name}-{z}.dat")) as f:
            return f.read().split()

    def pairs(tok, start, n):
        v = np.array(tok[start:start + 2 * n], dtype=np.float64).reshape(n, 2)
        return {"x": v[:, 0].tolist(), "y": v[:, 1].tolist()}

    out = {}
    t = tokens("pe-cs")
    out["xs_hi"] = pairs(t, 4, int(t[3]))
    t = tokens("pe-le-cs")
    if t:
        out["xs_lo"] = pairs(t, 4, int(t[3]))
    shells = []
    t = tokens("pe-low")
    nsh = int(t[1])
    out["thresh_lo"] = float(t[2])
    for i in range(nsh):
        row = t[3 + 7 * i: 3 + 7 * (i + 1)]
        shells.append({"binding_energy": float(row[0]),
                       "param_lo": [float(a) for a in row[1:]]})
    t = tokens("pe-high")
    out["thresh_hi"] = float(t[2])
    for i in range(nsh):
        row = t[3 + 7 * i: 3 + 7 * (i + 1)]
        assert float(row[0]) == shells[i]["binding_energy"]
        shells[i]["param_hi"] = [float(a) for a in row[1:]]
    t = tokens("pe-ss-cs")
    pos = 0
    for i in range(nsh):
        n = int(t[pos + 2])
        shells[i].update(pairs(t, pos + 4, n))
        pos += 4 + 2 * n
    out["shells"] = shells
    return out
