"""Track-order cases shared by the oracle and GPU tests.

The charge-partition strings are the ones the reference's
PartitionDataTest expects after extend-from-primaries + initialize-tracks
with TrackOrder::init_charge (test/celeritas/track/TrackSort.test.cc:
523-581): one character per slot, '_' inactive, 'N' neutral, 'C' charged.
Primaries are 1 GeV at (-22, 0, 0) cm along +x with particles cycling
e-, e+, gamma and event IDs 0, 1, 2, ... (TrackSort.test.cc:92-114). The
strings depend only on the particle types and the initialization order.
"""

import numpy as np

from celeritas_amd import workloads as wl

#: (num_slots, num_primaries, expected) -- host cases (:523-551)
PARTITION_INIT_HOST = [
    (32, 2, "______________________________CC"),
    (32, 16, "NNNNN________________CCCCCCCCCCC"),
    (32, 32, "NNNNNNNNNNCCCCCCCCCCCCCCCCCCCCCC"),
    (32, 64, "NNNNNNNNNNNCCCCCCCCCCCCCCCCCCCCC"),
]

#: device cases (:553-581)
PARTITION_INIT_DEVICE = [
    (8, 3, "N_____CC"),
    (16, 17, "NNNNNCCCCCCCCCCC"),
    (32, 31, "NNNNNNNNNN_CCCCCCCCCCCCCCCCCCCCC"),
]


def cycling_primaries(count: int) -> np.ndarray:
    """1 GeV e-, e+, gamma in turn along +x, one event per primary."""
    from celeritas_amd import PRIMARY_DTYPE

    p = np.zeros(count, dtype=PRIMARY_DTYPE)
    p["particle_id"] = np.array([wl.ELECTRON, wl.POSITRON, wl.GAMMA])[np.arange(count) % 3]
    p["event_id"] = np.arange(count)
    p["energy"] = 1000.0
    p["pos"] = (-22.0, 0.0, 0.0)
    p["dir"] = (1.0, 0.0, 0.0)
    return p


def charge_string(status, particle) -> str:
    """'_' for an inactive slot, 'N' for a photon, 'C' for e-/e+."""
    out = []
    for st, pid in zip(status, particle):
        if st == 0:
            out.append("_")
        elif pid == wl.GAMMA:
            out.append("N")
        else:
            out.append("C")
    return "".join(out)
