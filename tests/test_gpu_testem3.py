"""TestEm3 calorimeter transport on the device against the oracle.

The tables come from the built-in generator (shared by both sides). Two
checks:

- Early steps: with the same seed and primaries, the slot states after
  each of the first steps match the oracle (integer fields exactly, reals
  to 1e-9: device and host transcendental functions may differ in the last
  bit, and trajectories diverge only after such a difference is amplified).
- Whole runs: per-detector energy deposition from independent GPU and
  oracle runs (different seeds) agree within statistical error: the chi^2
  of the batch means over detectors with signal is compared with its
  distribution (p > 1e-3).
"""

import numpy as np
import pytest
from scipy import stats

import celeritas_amd as amd_pkg
from celeritas_amd import workloads as wl
from tests import oracle_bind as ob

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def physics():
    return amd_pkg.generate_physics(wl.testem3_physics_spec())


def _problem(physics, seed):
    return wl.testem3_problem(physics=physics, seed=seed, init_capacity=1 << 18,
                              max_events=1 << 12)


def test_testem3_early_steps_match_oracle(amd, physics):
    prob = _problem(physics, 20220511)
    params = amd.CoreParams(prob)
    stepper = amd.Stepper(params, 4096)
    ost = ob.State(ob.Problem(prob), 4096)
    prim = wl.testem3_primaries(8, emin=1e3, emax=1e4, seed=3)
    for step in range(25):
        if step == 0:
            r = stepper(prim)
            ost.insert_primaries(prim)
        else:
            r = stepper()
        o = ob.step_once(ost)
        print(f"step {step}: device {dict(r)} oracle {o}", flush=True)
        assert dict(r) == o
        st = stepper.state
        np.testing.assert_array_equal(st.get("status"), ost.get_u32("status"))
        live = st.get("status") == 2
        np.testing.assert_array_equal(st.get("particle_id")[live],
                                      ost.get_u32("particle")[live])
        np.testing.assert_allclose(st.get("energy")[live], ost.get_f64("energy")[live],
                                   rtol=1e-9, atol=1e-12)
        pos = np.stack([st.get("pos_x"), st.get("pos_y"), st.get("pos_z")], axis=1)
        opos = ost.get_f64("pos").reshape(-1, 3)
        np.testing.assert_allclose(pos[live], opos[live], rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("alg", ["minimal", "safety_plus"])
def test_testem3_msc_step_limit_algorithms_match_oracle(amd, alg):
    """The other Urban MSC step-limit algorithms (UrbanMscMinimalStepLimit,
    UrbanMscSafetyStepLimit with safety_plus): early TestEm3 steps match
    the oracle, whose algorithms reproduce the reference's step_limit
    golden values (tests/test_oracle_msc_limit.py)."""
    spec = wl.testem3_physics_spec()
    spec["options"] = dict(spec["options"], msc_step_limit_algorithm=alg)
    physics = amd_pkg.generate_physics(spec)
    assert physics["msc"]["step_limit_algorithm"] == alg
    prob = _problem(physics, 4242)
    stepper = amd.Stepper(amd.CoreParams(prob), 4096)
    ost = ob.State(ob.Problem(prob), 4096)
    prim = wl.testem3_primaries(8, emin=1e3, emax=1e4, seed=7)
    for step in range(25):
        if step == 0:
            r = stepper(prim)
            ost.insert_primaries(prim)
        else:
            r = stepper()
        o = ob.step_once(ost)
        assert dict(r) == o, f"step {step}"
        st = stepper.state
        np.testing.assert_array_equal(st.get("status"), ost.get_u32("status"))
        live = st.get("status") == 2
        np.testing.assert_allclose(st.get("energy")[live], ost.get_f64("energy")[live],
                                   rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(st.get("step_length")[live],
                                   ost.get_f64("step_length")[live], rtol=1e-9, atol=1e-12)


def _batch_means(run_batch, nbatch):
    tallies = np.array([run_batch(b) for b in range(nbatch)])
    return tallies.mean(axis=0), tallies.std(axis=0, ddof=1) / np.sqrt(nbatch)


def test_testem3_edep_statistical_parity(amd, physics):
    nbatch, nprim, energy = 12, 6, 2e3
    gpu_params = amd.CoreParams(_problem(physics, 1001))
    gpu_stepper = amd.Stepper(gpu_params, 1 << 14)
    oprob = ob.Problem(_problem(physics, 2002))

    def gpu_batch(b):
        prim = wl.testem3_primaries(nprim, emin=energy, emax=energy * 1.0001, seed=100 + b,
                                    gamma_fraction=0.5)
        before = gpu_stepper.state.get("calo_edep", 100).copy()
        gpu_stepper.run(prim, max_steps=100_000)
        return gpu_stepper.state.get("calo_edep", 100) - before

    def oracle_batch(b):
        prim = wl.testem3_primaries(nprim, emin=energy, emax=energy * 1.0001, seed=500 + b,
                                    gamma_fraction=0.5)
        st = ob.State(oprob, 1 << 12, stream_id=b + 1)  # independent RNG streams per batch
        ob.run(st, prim)
        return st.get_f64("calo_edep")

    gm, gs = _batch_means(gpu_batch, nbatch)
    om, os_ = _batch_means(oracle_batch, nbatch)
    total_g, total_o = gm.sum(), om.sum()
    print(f"mean deposited per batch: device {total_g:.1f} MeV, oracle {total_o:.1f} MeV")
    assert total_g > 0.5 * nprim * energy * 0.1
    sel = (gs > 0) & (os_ > 0) & (gm + om > 1e-3 * (gm + om).max())
    chi2 = float((((gm - om) ** 2) / (gs**2 + os_**2))[sel].sum())
    ndf = int(sel.sum())
    p = stats.chi2.sf(chi2, ndf)
    print(f"chi2 = {chi2:.1f} for {ndf} detectors, p = {p:.3g}")
    assert ndf > 20
    assert p > 1e-3
