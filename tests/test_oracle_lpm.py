"""Oracle LPM suppression pinned to the reference's golden values (CPU only).

- RelativisticBrem.test.cc:119-174: the relativistic DCS of a 25 GeV
  positron in Pb with and without LPM at ten photon energies.
- RelativisticBrem.test.cc:243-300: four interactor samples with LPM.
- RelativisticBrem.test.cc:302-347: mean photon energy and direction of
  1,000 samples with LPM.
- BetheHeitler.test.cc:208-287: the electron reduced-energy histogram of
  10,000 samples per energy in Cu (1 mol/cm^3), one engine through 1.5 MeV,
  100 MeV, 1 TeV (LPM active) and the threshold.

The material radiation length is derived from the elements as
MaterialParams.cc:496-519 does; each planted mistake in the oracle's LPM
functions (no dielectric suppression, no cubic xi correction, G and phi
exchanged) fails the DCS golden values.
"""

import ctypes

import numpy as np
import pytest

from tests import oracle_bind as ob
from tests import problems as pb
from tests import problems_em as pem

POSITRON, GAMMA_ID = pb.POSITRON, pb.GAMMA_ID
RTOL = 1e-12
K_POINTS = (1, 2, 5, 10, 20, 50, 100, 200, 500, 1000)
DXSEC_LPM = [3.15917865133079, 2.24073793752395, 1.7690465485807, 1.9393060296929,
             2.31528198148389, 2.89563604143589, 3.27222667999604, 3.5085408186385,
             3.48017265373814, 3.41228707554124]
DXSEC_NO_LPM = [3.55000253342095, 3.54986051043622, 3.54943449138746, 3.54872462599092,
                3.54730551901548, 3.54305318862896, 3.53598260644102, 3.52190382370326,
                3.48016652710843, 3.41226786120072]


def _with_lpm(prob):
    prob["physics"].setdefault("options", {})["lpm"] = True
    return prob


@pytest.fixture(scope="module")
def pb_lpm():
    return ob.Problem(_with_lpm(pem.rb_lead()))


@pytest.fixture(scope="module")
def cu_lpm():
    return ob.Problem(_with_lpm(pb.interactor_host()))


@pytest.fixture(autouse=True)
def _no_mutation():
    _mutate(0)
    yield
    _mutate(0)


def _mutate(m):
    f = ob.load().orc_lpm_set_mutation
    f.argtypes = [ctypes.c_int]
    f(m)


def _dxsec(problem, lpm, total, k):
    f = ob.load().orc_rb_dxsec_lpm
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                  ctypes.c_double, ctypes.c_double]
    f.restype = ctypes.c_double
    return f(problem.handle, 0, 0, int(lpm), total, k)


def _lpm_functions(problem, dielectric, gamma_energy, epsilon):
    f = ob.load().orc_lpm_functions
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                  ctypes.c_double, ctypes.c_double, ctypes.c_void_p]
    out = np.zeros(3)
    f(problem.handle, 0, 0, int(dielectric), gamma_energy, epsilon, out.ctypes.data)
    return out


def test_rb_dxsec_with_and_without_lpm(pb_lpm):
    total = 25000 + pem.EMASS
    with_lpm = [_dxsec(pb_lpm, True, total, k) for k in K_POINTS]
    without = [_dxsec(pb_lpm, False, total, k) for k in K_POINTS]
    np.testing.assert_allclose(with_lpm, DXSEC_LPM, rtol=RTOL)
    np.testing.assert_allclose(without, DXSEC_NO_LPM, rtol=RTOL)


def test_lpm_limits(pb_lpm):
    """Small suppression variable s -> phi ~ 6 s, G ~ 12 s (full suppression);
    large s -> phi, G -> 1 (no suppression); xi stays in [1, 2]."""
    # Huge photon energy and epsilon near 1/2: s' tiny
    xi, g, phi = _lpm_functions(pb_lpm, False, 1e12, 0.5)
    assert phi < 0.05 and g < 0.1 and 1 <= xi <= 2
    # Low photon energy: s' large, suppression vanishes
    xi, g, phi = _lpm_functions(pb_lpm, False, 1e-3, 0.5)
    assert abs(phi - 1) < 1e-6 and abs(g - 1) < 1e-6 and xi == pytest.approx(1 / phi)


def _angles_energies(rec):
    ang = np.einsum("ij,ij->i", rec[:, 2:5], rec[:, 9:12])
    return rec[:, 8], ang


def test_rb_interactor_with_lpm(pb_lpm):
    rec, _ = ob.interactor_sample(pb_lpm, "relativistic_brem", POSITRON, 25000.0,
                                  [0, 0, 1], 4, stack=4)
    energy, angle = _angles_energies(rec)
    np.testing.assert_allclose(energy, [18872.4157243063, 43.6117832245235,
                                        4030.31152398788, 217.621447606391], rtol=RTOL)
    np.testing.assert_allclose(angle, [0.999999971800136, 0.999999999587026,
                                       0.999999999683752, 0.999999999474844], rtol=RTOL)


def test_rb_interactor_stress_with_lpm(pb_lpm):
    rec, _ = ob.interactor_sample(pb_lpm, "relativistic_brem", POSITRON, 25000.0,
                                  [0, 0, 1], 1000, stack=1000)
    assert np.all(rec[:, 0] != 3)
    assert rec[:, 8].mean() == pytest.approx(2932.1072998587733, rel=RTOL)
    mean_dir = rec[:, 9:12].mean(axis=0)
    # EXPECT_SOFT_NEAR(..., 1e-8) for the transverse components
    assert mean_dir[0] == pytest.approx(3.3286986548662216e-06, rel=1e-8)
    assert mean_dir[1] == pytest.approx(1.3067055198983571e-06, rel=1e-8)
    assert mean_dir[2] == pytest.approx(0.99999999899845182, rel=RTOL)


def test_bethe_heitler_distributions_with_lpm(cu_lpm):
    n = 10000
    me = pb.EMASS
    energies = (1.5, 100.0, 1e6, 2 * me)
    items = [("bethe_heitler", GAMMA_ID, e, [0, 0, 1], n, 2 * n) for e in energies]
    rec = ob.interactor_sequence(cu_lpm, items)
    expected = [[0, 0, 0, 1911, 3054, 3142, 1893, 0, 0, 0],
                [754, 1109, 1054, 1055, 1010, 1010, 1024, 1055, 1090, 839],
                [1209, 1073, 911, 912, 844, 881, 903, 992, 1066, 1209],
                [0, 0, 0, 0, 0, 10000, 0, 0, 0, 0]]
    for i, e in enumerate(energies):
        eps = (rec[i * n:(i + 1) * n, 8] + me) / e
        bins = (eps * 10).astype(int)
        hist = np.bincount(bins[(bins >= 0) & (bins < 10)], minlength=10)
        np.testing.assert_array_equal(hist, expected[i], err_msg=f"E = {e} MeV")


@pytest.mark.parametrize("mutation", [1, 2, 3])
def test_planted_lpm_mistake_fails(pb_lpm, mutation):
    _mutate(mutation)
    total = 25000 + pem.EMASS
    got = np.array([_dxsec(pb_lpm, True, total, k) for k in K_POINTS])
    assert not np.allclose(got, DXSEC_LPM, rtol=RTOL)
