"""Oracle Urban energy-loss fluctuations pinned to the reference.

EnergyLossHelper.test.cc:238-285: a 100 MeV electron in argon (1 mol/cm^3,
I = 188 eV, electron cut 1 keV) with a 0.01 MeV mean loss: the helper's
maximum energy transfer and beta^2, then 10,000 samples from a
default-seeded mt19937 with mean 0.0099918954960280353 MeV after exactly
551,188 32-bit draws.
"""

import ctypes

import numpy as np
import pytest

from tests import oracle_bind as ob
from tests import problems_em as pem


def argon():
    prob = pem.single_element(18, 39.948, 1.0, 0.0, "Ar")
    prob["physics"]["particles"] = [
        {"label": "electron", "pdg": 11, "mass": 0.5109989461, "charge": -1.0}]
    prob["physics"]["materials"][0]["mean_exc_energy"] = 188e-6
    prob["physics"]["cutoffs"] = {"electron": [{"energy": 1e-3}]}
    prob["physics"]["options"] = {"eloss_fluctuation": True}
    return prob


def test_urban_fluctuation_golden():
    p = ob.Problem(argon())
    f = ob.load().orc_urban_fluct
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                  ctypes.c_uint32, ctypes.c_void_p]
    f.restype = ctypes.c_int
    out = np.zeros(4)
    ob._check(f(p.handle, 0, 100.0, 0.01, 10000, out.ctypes.data))
    assert out[0] == pytest.approx(0.001, rel=1e-12)
    assert out[1] == pytest.approx(0.99997415284006, rel=1e-12)
    assert out[2] == pytest.approx(0.0099918954960280353, rel=1e-12)
    assert out[3] == 551188
