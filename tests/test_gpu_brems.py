"""Device bremsstrahlung against the oracle on identical uniform streams.

Seltzer-Berger with the br29 table (electrons and positrons), the
relativistic model in Pb (with and without LPM suppression) and the
combined model across its 1 GeV switch;
FP64 results must match to 1e-12 relative.
"""

import numpy as np
import pytest

from tests import gpu_compare
from tests import oracle_bind as ob
from tests import problems as pb
from tests import problems_em as pem

pytestmark = pytest.mark.gpu

N = 4096
STRIDE = 512

CASES = [
    ("seltzer_berger", "cu", pb.ELECTRON, (0.025, 5e3)),
    ("seltzer_berger", "cu", pb.POSITRON, (0.025, 5e3)),
    ("combined_brem", "cu", pb.ELECTRON, (0.025, 1e5)),
    ("combined_brem", "cu", pb.POSITRON, (0.025, 1e5)),
    ("relativistic_brem", "pb", pb.ELECTRON, (1e3, 1e6)),
    ("relativistic_brem", "pb", pb.POSITRON, (1e3, 1e6)),
    # LPM suppression (above ~0.5 GeV total energy in Pb)
    ("relativistic_brem", "pb_lpm", pb.ELECTRON, (1e2, 1e6)),
    ("relativistic_brem", "pb_lpm", pb.POSITRON, (1e2, 1e6)),
]


def _lpm(prob):
    prob["physics"]["options"] = {"lpm": True}
    return prob


@pytest.fixture(scope="module")
def setups(amd):
    out = {}
    for key, prob in (("cu", pem.sb_copper()), ("pb", pem.rb_lead()),
                      ("pb_lpm", _lpm(pem.rb_lead()))):
        params = amd.CoreParams(prob)
        out[key] = (params, amd.CoreState(params, N), ob.Problem(prob))
    return out


@pytest.mark.parametrize("model,mat,particle,erange", CASES)
def test_brems_parity(amd, setups, model, mat, particle, erange):
    params, state, oprob = setups[mat]
    rng = np.random.default_rng(7 + particle + 13 * len(model))
    energies = np.exp(rng.uniform(np.log(erange[0]), np.log(erange[1]), N))
    dirs = rng.normal(size=(N, 3))
    dirs[: N // 8] = [0, 0, 1]
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    uniforms = ob.mt_uniforms(99 + particle, N * STRIDE)
    # The oracle runs first: it raises if any sample needs more than STRIDE
    # uniforms, so the device never reads past its stretch of the buffer
    cpu = ob.interactor_replay_batch(oprob, model, particle, energies, dirs, uniforms,
                                     STRIDE, material=0)
    gpu = amd.verify_interactor(params, state, model, particle, energies, dirs, uniforms,
                                STRIDE, material=0)
    np.testing.assert_array_equal(gpu[:, 0], cpu[:, 0])
    np.testing.assert_array_equal(gpu[:, 6], cpu[:, 6])
    np.testing.assert_array_equal(gpu[:, 7], cpu[:, 7])
    gpu_compare.assert_records_close(gpu, cpu, rtol=1e-12, atol=1e-12)
    # Photon energies lie between the cutoff and the incident energy
    assert np.all(gpu[:, 8] > 0) and np.all(gpu[:, 8] < energies)
