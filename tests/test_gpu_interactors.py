"""Device interactors against the oracle on identical uniform streams.

Each sample gets its own stretch of uniforms from a seeded std::mt19937
(through std::generate_canonical, as in the reference's interactor tests);
the device result must match the oracle to 1e-12 relative (FP64).
"""

import zlib

import numpy as np
import pytest

from tests import gpu_compare
from tests import oracle_bind as ob
from tests import problems as pb

pytestmark = pytest.mark.gpu

N = 4096
STRIDE = 1024


def random_dirs(rng, n):
    v = rng.normal(size=(n, 3))
    v[: n // 8] = [0, 0, 1]  # exercise the near-axis branch of rotate()
    return v / np.linalg.norm(v, axis=1, keepdims=True)


CASES = [
    ("klein_nishina", pb.GAMMA_ID, (1e-3, 1e4), 0),
    ("bethe_heitler", pb.GAMMA_ID, (1.1, 1e5), 0),
    ("moller_bhabha", pb.ELECTRON, (0.0021, 1e5), 0),
    ("moller_bhabha", pb.POSITRON, (0.0011, 1e5), 0),
    ("eplusgg", pb.POSITRON, (1e-4, 1e5), 1),
]


@pytest.fixture(scope="module")
def setup(amd):
    prob = pb.interactor_host()
    params = amd.CoreParams(prob)
    state = amd.CoreState(params, N)
    return params, state, ob.Problem(prob)


@pytest.mark.parametrize("model,particle,erange,material", CASES)
def test_interactor_parity(amd, setup, model, particle, erange, material):
    params, state, oprob = setup
    rng = np.random.default_rng(zlib.crc32(model.encode()) + particle)
    energies = np.exp(rng.uniform(np.log(erange[0]), np.log(erange[1]), N))
    dirs = random_dirs(rng, N)
    uniforms = ob.mt_uniforms(1234 + particle, N * STRIDE)
    # Oracle first: it raises if a sample would read past its STRIDE uniforms
    cpu = ob.interactor_replay_batch(oprob, model, particle, energies, dirs, uniforms,
                                     STRIDE, material=material)
    gpu = amd.verify_interactor(params, state, model, particle, energies, dirs, uniforms,
                                STRIDE, material=material)
    np.testing.assert_array_equal(gpu[:, 0], cpu[:, 0])  # action
    np.testing.assert_array_equal(gpu[:, 6], cpu[:, 6])  # secondary count
    np.testing.assert_array_equal(gpu[:, 7], cpu[:, 7])  # secondary particle
    np.testing.assert_array_equal(gpu[:, 12], cpu[:, 12])
    gpu_compare.assert_records_close(gpu, cpu, rtol=1e-12, atol=1e-12)


@pytest.fixture(scope="module")
def setup_lpm(amd):
    prob = pb.interactor_host()
    prob["physics"]["options"] = {"lpm": True}
    params = amd.CoreParams(prob)
    state = amd.CoreState(params, N)
    return params, state, ob.Problem(prob)


def test_bethe_heitler_lpm_parity(amd, setup_lpm):
    """Bethe-Heitler with LPM suppression (active above 100 GeV) in Cu."""
    params, state, oprob = setup_lpm
    rng = np.random.default_rng(31)
    energies = np.exp(rng.uniform(np.log(1e4), np.log(1e7), N))
    dirs = random_dirs(rng, N)
    uniforms = ob.mt_uniforms(4321, N * STRIDE)
    cpu = ob.interactor_replay_batch(oprob, "bethe_heitler", pb.GAMMA_ID, energies, dirs,
                                     uniforms, STRIDE)
    gpu = amd.verify_interactor(params, state, "bethe_heitler", pb.GAMMA_ID, energies, dirs,
                                uniforms, STRIDE)
    np.testing.assert_array_equal(gpu[:, 0], cpu[:, 0])
    np.testing.assert_array_equal(gpu[:, 6], cpu[:, 6])
    gpu_compare.assert_records_close(gpu, cpu, rtol=1e-12, atol=1e-12)
