"""Oracle bremsstrahlung pinned to the reference's golden values.

SeltzerBerger.test.cc:142-441 (br29 table argmax, positron scaling, energy
distribution, four interactor samples, stress draw counts) and
RelativisticBrem.test.cc:119-241 (differential cross section without LPM,
four interactor samples).
"""

import numpy as np
import pytest

from tests import oracle_bind as ob
from tests import oracle_brems as obr
from tests import problems as pb
from tests import problems_em as pem

ELECTRON, POSITRON = pb.ELECTRON, pb.POSITRON


@pytest.fixture(scope="module")
def cu():
    return ob.Problem(pem.sb_copper())


@pytest.fixture(scope="module")
def pb_rb():
    return ob.Problem(pem.rb_lead())


def test_sb_table_argmax(cu):
    expected = [31, 31, 31, 30, 30, 7, 7, 6, 5, 4, 3] + [0] * 46
    np.testing.assert_array_equal(obr.sb_argmax(cu, 0, 57), expected)


def test_sb_positron_scaling():
    got = [obr.sb_positron_scale(29, pem.EMASS, 0.01, e, f * e)
           for e in (1.0, 10.0, 100.0, 1000.0) for f in (0.10001, 0.5, 0.9, 0.9999)]
    expected = [
        0.98771267862086, 0.88085886234621, 0.36375147691123, 2.6341925633236e-29,
        0.99965385757708, 0.99583269657665, 0.92157316225919, 2.1585790781929e-09,
        0.99999599590292, 0.99994914123134, 0.99844428624414, 0.0041293798201,
        0.99999995934326, 0.99999948043882, 0.99998298916928, 0.33428689072689]
    np.testing.assert_allclose(got, expected, atol=1e-11, rtol=0)


def test_sb_energy_distribution(cu):
    kc = 0.0009
    mt = obr.MtEngine()
    max_xs, frac, draws = [], [], []
    for e in (0.001, 0.0045, 0.567, 7.89, 89.0, 901.0):
        for corr in (0, 1):
            r = obr.sb_sample_many(cu, mt, e, pem.EMASS, kc, corr, 8192)
            frac.append(r[1])
            draws.append(r[2])
        max_xs.append(r[0])
    r = obr.sb_sample_many(cu, mt, 7.89, pem.EMASS, kc, 2, 8192)
    frac.append(r[1])
    draws.append(r[2])
    np.testing.assert_allclose(max_xs, [2.866525852195, 4.72696244794, 12.18911946078,
                                        13.93366489719, 13.85758694967, 13.3353235437],
                               rtol=1e-12)
    np.testing.assert_allclose(frac, [
        0.94912259860422, 0.90270157074556, 0.49736065674058, 0.27711716215819,
        0.081515129333292, 0.068559142299853, 0.065803331441324, 0.064344514250384,
        0.079512002547402, 0.077647502218254, 0.085615341879476, 0.086428313853775,
        0.065321200129584], rtol=1e-12)
    np.testing.assert_array_equal(draws, [
        4.0791015625, 137.044921875, 4.060546875, 15.74169921875, 5.103515625,
        5.26953125, 4.67333984375, 4.6572265625, 4.4306640625, 4.4638671875,
        4.35400390625, 4.349609375, 9.189453125])


def _angles_energies(rec):
    ang = np.einsum("ij,ij->i", rec[:, 2:5], rec[:, 9:12])
    return rec[:, 8], ang


def test_sb_interactor_basic(cu):
    rec, _ = ob.interactor_sample(cu, "seltzer_berger", ELECTRON, 1.0, [0, 0, 1], 5,
                                  stack=4)
    energy, angle = _angles_energies(rec[:4])
    np.testing.assert_allclose(energy, [0.0349225070114679, 0.0316182310804369,
                                        0.0838794010486177, 0.106195186929141], rtol=1e-12)
    np.testing.assert_allclose(angle, [0.959441513277674, 0.994350429950924,
                                       0.968866136008621, 0.961582855967571], rtol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(rec[:4, 2:5], axis=1), 1, rtol=1e-14)
    assert rec[4, 0] == 3  # out of secondary space -> failed


def test_sb_interactor_stress_draws(cu):
    got = []
    n = 10000
    dirs = ([0, 0, 1], [1, 0, 0], [1e-9, 0, 1], [1, 1, 1])
    mt = obr.MtEngine()
    for particle in (ELECTRON, POSITRON):
        for e in (1.5, 5.0, 10.0, 50.0, 100.0):
            count = sum(obr.interactor_stress(cu, mt, "seltzer_berger", particle, e, d, n)
                        for d in dirs)
            got.append(count / (n * len(dirs)))
    np.testing.assert_allclose(got, [14.088, 13.2402, 12.9641, 12.5832, 12.4988,
                                     14.2108, 13.254, 12.9431, 12.5952, 12.4888], rtol=1e-12)


def test_rb_dxsec_without_lpm(pb_rb):
    total = 25000 + pem.EMASS
    got = [obr.rb_dxsec(pb_rb, 0, total, k)
           for k in (1, 2, 5, 10, 20, 50, 100, 200, 500, 1000)]
    np.testing.assert_allclose(got, [
        3.55000253342095, 3.54986051043622, 3.54943449138746, 3.54872462599092,
        3.54730551901548, 3.54305318862896, 3.53598260644102, 3.52190382370326,
        3.48016652710843, 3.41226786120072], rtol=1e-12)


def test_rb_interactor_without_lpm(pb_rb):
    rec, count = ob.interactor_sample(pb_rb, "relativistic_brem", POSITRON, 25000.0,
                                      [0, 0, 1], 4, stack=4)
    energy, angle = _angles_energies(rec)
    assert count / 4 == 12
    np.testing.assert_allclose(energy, [9.7121539090503, 16.1109589071687,
                                        7.48863745059463, 8338.70226190511], rtol=1e-12)
    np.testing.assert_allclose(angle, [0.999999999858782, 0.999999999999921,
                                       0.999999976998416, 0.999999998137601], rtol=1e-12)
