"""Oracle TestEm3 transport on generated tables (CPU only).

- Positrons near the ionisation threshold in Pb: every interaction keeps
  energies non-negative and the run empties (a delta-ray cut taken from
  the positron instead of the electron cut once produced negative
  energies and tracks that never finished).
- A small electron shower deposits a plausible fraction of its energy in
  the calorimeter and empties.
"""

import numpy as np
import pytest

import celeritas_amd as amd
from celeritas_amd import workloads as wl
from tests import oracle_bind as ob


@pytest.fixture(scope="module")
def problem():
    phys = amd.generate_physics(wl.testem3_physics_spec())
    return ob.Problem(wl.testem3_problem(physics=phys, init_capacity=1 << 20,
                                         max_events=1 << 12))


def test_positrons_near_threshold(problem):
    n = 400
    prim = wl.testem3_primaries(n, seed=7, emin=0.5, emax=3.0, gamma_fraction=0.0)
    prim["particle_id"] = wl.POSITRON
    prim["pos"][:, 0] = -19.9  # inside the first Pb gap
    prim["event_id"] = np.arange(n)
    st = ob.State(problem, 4096)
    active, _ = ob.run(st, prim, max_iters=20000)
    assert active[-1] == 0 or len(active) < 20000
    assert np.all(st.get_f64("energy") >= 0)


def test_small_shower(problem):
    prim = wl.testem3_primaries(2, seed=3, emin=1e3, emax=1.0001e3, gamma_fraction=0.0)
    st = ob.State(problem, 4096)
    active, _ = ob.run(st, prim, max_iters=100000)
    edep = st.get_f64("calo_edep")
    assert len(active) < 100000
    # 1 GeV electrons in ~43 X0: most energy stays in the calorimeter; the
    # sampling fraction of the argon layers is a few percent
    assert 0.5 * 2e3 < edep.sum() < 2e3
    lar = edep[1::2].sum() / edep.sum()
    assert 0.02 < lar < 0.5
