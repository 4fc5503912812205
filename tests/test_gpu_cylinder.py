"""Cylindrical detector in a 3.8 T solenoid field on the device against
the oracle (FP64, Dormand-Prince propagation, compounds with element
selection, separate Seltzer-Berger / relativistic brems models).

- Early steps: slot states after each of the first steps equal the
  oracle's (integers exactly, reals to 1e-9).
- Whole runs: per-layer energy deposition from independent device and
  oracle runs agrees within statistical error (chi^2 of batch means over
  layers with signal, p > 1e-3).
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
    return amd_pkg.generate_physics(wl.cylinder_physics_spec())


def _problem(physics, seed, order="none"):
    return wl.cylinder_problem(physics=physics, seed=seed, init_capacity=1 << 18,
                               max_events=1 << 12, track_order=order)


def test_cylinder_early_steps_match_oracle(amd, physics):
    prob = _problem(physics, 20240101)
    params = amd.CoreParams(prob)
    stepper = amd.Stepper(params, 4096)
    ost = ob.State(ob.Problem(prob), 4096)
    prim = wl.cylinder_primaries(8, emin=500, emax=5000, seed=3)
    for step in range(20):
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


def _batch_means(run_batch, nbatch):
    t = np.array([run_batch(b) for b in range(nbatch)])
    return t.mean(axis=0), t.std(axis=0, ddof=1) / np.sqrt(nbatch)


@pytest.mark.parametrize("order", ["none", "reindex_both_action"])
def test_cylinder_edep_statistical_parity(amd, physics, order):
    nbatch, nprim, energy = 10, 4, 5e3
    ndet = max(wl.cylinder_geometry()[2]) + 1
    stepper = amd.Stepper(amd.CoreParams(_problem(physics, 1001, order)), 1 << 14)
    oprob = ob.Problem(_problem(physics, 2002))

    def gpu_batch(b):
        prim = wl.cylinder_primaries(nprim, emin=energy, emax=energy * 1.0001, seed=100 + b)
        before = stepper.state.get("calo_edep", ndet).copy()
        stepper.run(prim, max_steps=200_000)
        return stepper.state.get("calo_edep", ndet) - before

    def oracle_batch(b):
        prim = wl.cylinder_primaries(nprim, emin=energy, emax=energy * 1.0001, seed=500 + b)
        st = ob.State(oprob, 1 << 12, stream_id=b + 1)  # independent RNG streams per batch
        ob.run(st, prim)
        return st.get_f64("calo_edep")[:ndet]

    gm, gs = _batch_means(gpu_batch, nbatch)
    om, os_ = _batch_means(oracle_batch, nbatch)
    print(f"deposited per batch: device {gm.sum():.1f} MeV, oracle {om.sum():.1f} MeV")
    assert gm.sum() > 0.2 * nprim * energy
    sel = (gs > 0) & (os_ > 0) & (gm + om > 1e-5 * (gm + om).max())
    chi2 = float((((gm[sel] - om[sel]) ** 2) / (gs[sel] ** 2 + os_[sel] ** 2)).sum())
    ndf = int(sel.sum())
    p = stats.chi2.sf(chi2, ndf)
    print(f"{order}: chi2 = {chi2:.1f} for {ndf} layers, p = {p:.3g}")
    assert ndf > 5
    assert p > 1e-3
