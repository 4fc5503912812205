"""Oracle Livermore photoelectric pinned to the reference's golden values.

LivermorePE.test.cc:137-203 (four samples at 1 keV on K), :205-297 (stress:
draws per sample and mean photoelectron energy at five energies) and
:484-517 (macroscopic cross section over 1e-4 to 1e6 MeV), all with the
pe-*-19.dat files the reference's tests hold.
"""

import numpy as np
import pytest

from tests import oracle_bind as ob
from tests import oracle_brems as obr
from tests import problems as pb
from tests import problems_em as pem


@pytest.fixture(scope="module")
def k_problem():
    return ob.Problem(pem.pe_potassium())


def test_pe_basic(k_problem):
    rec, _ = ob.interactor_sample(k_problem, "livermore_pe", pb.GAMMA_ID, 0.001, [0, 0, 1], 5,
                                  stack=4)
    np.testing.assert_array_equal(rec[:4, 0], 1)  # absorbed
    np.testing.assert_allclose(rec[:4, 8], [0.00062884, 0.00062884, 0.00070136, 0.00069835],
                               rtol=1e-12)
    np.testing.assert_allclose(rec[:4, 11], [0.1217302869581, 0.8769397871407,
                                             -0.1414717733267, -0.2414106440617], rtol=1e-12)
    np.testing.assert_allclose(rec[:4, 5], [0.00037116, 0.00037116, 0.00029864, 0.00030165],
                               rtol=1e-12)
    assert rec[4, 0] == 3  # failed: out of secondary space


def test_pe_stress(k_problem):
    mt = obr.MtEngine()
    n = 8192
    draws, energy = [], []
    for e in (0.0001, 0.01, 1.0, 10.0, 1000.0):
        sums = np.zeros(3)
        count = 0
        for d in ([0, 0, 1], [1, 0, 0], [1e-9, 0, 1], [1, 1, 1]):
            count += obr.interactor_stress(k_problem, mt, "livermore_pe", pb.GAMMA_ID, e, d,
                                           n, sums=sums)
        draws.append(count / (4 * n))
        energy.append(sums[1] / sums[0])
        assert sums[0] == 4 * n
    np.testing.assert_allclose(draws, [15.99755859375, 16.09204101562, 13.79919433594,
                                       8.590209960938, 2], rtol=1e-12)
    np.testing.assert_allclose(energy, [7.287875885011e-05, 0.006708485731503,
                                        0.9967066970311, 9.996704339284, 999.9967069717],
                               rtol=1e-12)


def test_pe_macro_xs(k_problem):
    # Accumulated log energy, as the reference loop does
    got = []
    le = np.log(1e-4)
    delta = (np.log(1e6) - np.log(1e-4)) / 19
    for _ in range(20):
        got.append(obr.pe_macro_xs(k_problem, 0, float(np.exp(le))))
        le += delta
    np.testing.assert_allclose(got, [
        9.235615290944, 17.56658325086, 1.161217594282, 0.4108511065363,
        0.01515608909912, 0.0004000659204694, 9.083754758322e-06, 2.449452106704e-07,
        1.800625084911e-08, 3.188458732396e-09, 8.028833591133e-10, 2.2700912115e-10,
        6.653075041804e-11, 1.971081007251e-11, 5.85857761177e-12, 1.743005702864e-12,
        5.187166124179e-13, 1.543827005416e-13, 4.594922185898e-14, 1.367605938008e-14],
        rtol=1e-12)
