"""Oracle Urban MSC step-limit algorithms ("minimal", "safety",
"safety_plus") against the reference's golden values (CPU only).

UrbanMsc.test.cc:296-468 (step_limit) samples 100 limits of an electron in
G4_STAINLESS-STEEL at five energies, each with a fresh MSC range and a
physical step equal to the range, with one default-seeded mt19937 that
continues through the four cases: minimal off a boundary, minimal on a
boundary (range factor 0.2 for both), safety and safety_plus on a boundary
(zero safety). The inputs the reference takes from its Geant4 tables are
recovered from its own golden values:

- the range at each energy is the minimal, off-boundary mean step (off a
  boundary on the first step the minimal limit is infinite, so every step
  is the physical step, the range);
- the MSC mean free path is the safety case's range_init at 10 and 100 MeV
  (there range_init = max(range, lambda) = lambda), and below from that
  case's limit_min through calc_limit_min, lambda = limit_min (2 + E (c0 +
  c1 E)) / scaled_zeff, with the steel coefficients pinned in
  test_oracle_msc.py. At 10 and 100 MeV both routes must agree, which
  checks calc_limit_min independently.

Everything else (range_init, range_factor, limit_min after the samples and
the mean sampled step, which fixes the Gaussian's mean, width, clamping and
the engine's draws) is compared with the golden values. Planted mistakes
(the safety_plus range_init from max(range, lambda), its 0.84 factor as
0.75, no step cap) must fail.
"""

import numpy as np
import pytest

from tests import oracle_bind as ob
from tests.oracle_brems import MtEngine
from tests.test_oracle_msc import steel_zeff

E = np.array([0.01, 0.1, 1.0, 10.0, 100.0])
INF = float("inf")

#: UrbanMsc.test.cc:378-467
GOLDEN = {
    ("minimal", False): {
        "mean_step": [5.4443402771743e-05, 0.0026634569415511, 0.07706894616868,
                      0.9059108153443, 8.8845468954557],
        "range_init": [INF] * 5, "range_factor": [0.2] * 5, "limit_min": [1e-08] * 5},
    ("minimal", True): {
        "mean_step": [1.0947964502125e-05, 0.0005291603863133, 0.015353327620128,
                      0.21986948123587, 8.8845468954555],
        "range_init": [1.0888680554349e-05, 0.00053269138831022, 0.015413789233736,
                       0.21762788543933, 15.553546812173],
        "range_factor": [0.2] * 5, "limit_min": [1e-08] * 5},
    ("safety", True): {
        "mean_step": [2.1774342857589e-06, 0.00010553666854323, 0.0030850804258273,
                      0.15209571378504, 8.8845468954555],
        "range_init": [5.4443402771744e-05, 0.0026634569415511, 0.07706894616868,
                       1.0881394271966, 77.767734060865],
        "range_factor": [0.04, 0.04, 0.04, 0.13881394271966, 7.8067734060865],
        "limit_min": [1.9688399316472e-06, 1.0522532283188e-05, 3.1432398888924e-05,
                      4.0583539826243e-05, 3.6094312868035e-05]},
    ("safety_plus", True): {
        "mean_step": [2.1762357027363e-06, 0.00010784095490268, 0.0030577621281947,
                      0.094075756521388, 3.1108913402487],
        "range_init": [5.4443402771744e-05, 0.0026634569415511, 0.07706894616868,
                       0.90591081534428, 8.8845468954556],
        "range_factor": [0.04, 0.04, 0.04, 0.10324092334058, 5.0107349798953],
        "limit_min": [1.9688399316472e-06, 1.0522532283188e-05, 3.1432398888924e-05,
                      4.0583539826243e-05, 3.6094312868035e-05]},
}
ORDER = [("minimal", False), ("minimal", True), ("safety", True), ("safety_plus", True)]


def inputs():
    rng = np.array(GOLDEN[("minimal", False)]["mean_step"])
    c = ob.msc_material_coeffs(steel_zeff())
    lmin = np.array(GOLDEN[("safety", True)]["limit_min"])
    lam = lmin * (2 + E * (c["stepmin0"] + c["stepmin1"] * E)) / c["scaled_zeff_e"]
    return rng, lam


def run_all(mutation=0):
    rng, lam = inputs()
    mt = MtEngine()
    out = {}
    ob.msc_set_mutation(mutation)
    try:
        for alg, bnd in ORDER:
            rf = 0.2 if alg == "minimal" else 0.04
            out[(alg, bnd)] = ob.msc_step_limit_sample(mt.h, alg, bnd, steel_zeff(), rf, E,
                                                       rng, lam, 100)
    finally:
        ob.msc_set_mutation(0)
    return out


def test_recovered_mfp_consistent():
    """calc_limit_min inverted at 10 and 100 MeV gives the MFP that the
    safety case's range_init shows directly."""
    _, lam = inputs()
    direct = np.array(GOLDEN[("safety", True)]["range_init"])[3:]
    np.testing.assert_allclose(lam[3:], direct, rtol=1e-10)
    # Below 10 MeV the MFP is below the range, as range_init = range shows
    rng, _ = inputs()
    assert np.all(lam[:3] < rng[:3])


def test_step_limit_golden():
    out = run_all()
    for key in ORDER:
        for field, expected in GOLDEN[key].items():
            np.testing.assert_allclose(out[key][field], expected, rtol=1e-10,
                                       err_msg=f"{key} {field}")


@pytest.mark.parametrize("mutation", [11, 12, 13])
def test_planted_step_limit_mistakes_fail(mutation):
    out = run_all(mutation)
    bad = False
    for key in ORDER:
        for field, expected in GOLDEN[key].items():
            if not np.allclose(out[key][field], expected, rtol=1e-10):
                bad = True
    assert bad, f"planted mistake {mutation} went unnoticed"
