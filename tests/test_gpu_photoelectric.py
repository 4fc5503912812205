"""Device Livermore photoelectric against the oracle on identical uniform
streams: the reference's K (Z=19) data files and the generated Pb/Ar/H
shell-model data of the TestEm3 materials; FP64 to 1e-12 relative."""

import numpy as np
import pytest

import celeritas_amd as amd_pkg
from celeritas_amd import workloads as wl
from tests import gpu_compare
from tests import oracle_bind as ob
from tests import problems_em as pem

pytestmark = pytest.mark.gpu

N = 4096
STRIDE = 256


def _generated_problem():
    phys = amd_pkg.generate_physics(wl.testem3_physics_spec())
    prob = pem.single_element(19, 39.0983, 1e-5, 0.0, "K")
    prob["physics"] = phys
    prob["volume_to_material"] = [-1, 0, 1]
    return prob


@pytest.fixture(scope="module")
def setups(amd):
    out = {}
    for key, prob in (("k", pem.pe_potassium()), ("gen", _generated_problem())):
        params = amd.CoreParams(prob)
        out[key] = (params, amd.CoreState(params, N), ob.Problem(prob))
    return out


@pytest.mark.parametrize("key,material,erange", [
    ("k", 0, (1e-5, 1e3)),
    ("gen", 0, (1e-5, 1e3)),
    ("gen", 1, (1e-5, 1e2)),
    ("gen", 2, (1e-5, 1.0)),
])
def test_photoelectric_parity(amd, setups, key, material, erange):
    params, state, oprob = setups[key]
    gamma = 2
    rng = np.random.default_rng(11 + material)
    energies = np.exp(rng.uniform(np.log(erange[0]), np.log(erange[1]), N))
    dirs = rng.normal(size=(N, 3))
    dirs[: N // 8] = [0, 0, 1]
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    uniforms = ob.mt_uniforms(5 + material, N * STRIDE)
    cpu = ob.interactor_replay_batch(oprob, "livermore_pe", gamma, energies, dirs, uniforms,
                                     STRIDE, material=material)
    gpu = amd.verify_interactor(params, state, "livermore_pe", gamma, energies, dirs,
                                uniforms, STRIDE, material=material)
    np.testing.assert_array_equal(gpu[:, 0], cpu[:, 0])
    np.testing.assert_array_equal(gpu[:, 6], cpu[:, 6])
    np.testing.assert_array_equal(gpu[:, 7], cpu[:, 7])
    gpu_compare.assert_records_close(gpu, cpu, rtol=1e-12, atol=1e-12)
    # Energy conservation: photoelectron + deposited binding energy
    total = gpu[:, 5] + np.where(gpu[:, 6] > 0, gpu[:, 8], 0)
    np.testing.assert_allclose(total, energies, rtol=1e-12)
