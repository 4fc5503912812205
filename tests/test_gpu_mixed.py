"""Mixed-precision transport against the FP64 oracle.

In the mixed mode the transcendental functions of sampling, cross-section
interpolation and MSC run in FP32 and uniform draws are FP32; positions,
safety, field integration, energies and the energy-deposition tally stay
FP64. Runs are compared with the FP64 oracle statistically: the chi^2 of
per-detector batch means (p > 1e-3), and the relative difference of the
total deposited energy is printed as the stated FP64-oracle error.
"""

import numpy as np
import pytest
from scipy import stats

import celeritas_amd as amd_pkg
from celeritas_amd import workloads as wl
from tests import oracle_bind as ob

pytestmark = pytest.mark.gpu


def _compare(gm, gs, om, os_, label):
    sel = (gs > 0) & (os_ > 0) & (gm + om > 1e-5 * (gm + om).max())
    chi2 = float((((gm[sel] - om[sel]) ** 2) / (gs[sel] ** 2 + os_[sel] ** 2)).sum())
    ndf = int(sel.sum())
    p = stats.chi2.sf(chi2, ndf)
    rel = (gm.sum() - om.sum()) / om.sum()
    err = np.sqrt((gs**2).sum() + (os_**2).sum()) / om.sum()
    print(f"{label}: total edep mixed/FP64-oracle - 1 = {rel:+.4f} +- {err:.4f}; "
          f"chi2 = {chi2:.1f} for {ndf} detectors, p = {p:.3g}")
    return ndf, p


def _batches(run_batch, nbatch):
    t = np.array([run_batch(b) for b in range(nbatch)])
    return t.mean(axis=0), t.std(axis=0, ddof=1) / np.sqrt(nbatch)


def test_testem3_mixed_vs_fp64_oracle(amd):
    physics = amd_pkg.generate_physics(wl.testem3_physics_spec())
    nbatch, nprim, energy = 12, 6, 2e3

    def problem(seed, precision):
        return wl.testem3_problem(physics=physics, seed=seed, init_capacity=1 << 18,
                                  max_events=1 << 12, precision=precision,
                                  track_order="reindex_both_action")

    stepper = amd.Stepper(amd.CoreParams(problem(1001, "mixed")), 1 << 14)
    oprob = ob.Problem(problem(2002, "fp64"))

    def gpu_batch(b):
        prim = wl.testem3_primaries(nprim, emin=energy, emax=energy * 1.0001, seed=100 + b)
        before = stepper.state.get("calo_edep", 100).copy()
        stepper.run(prim, max_steps=100_000)
        return stepper.state.get("calo_edep", 100) - before

    def oracle_batch(b):
        prim = wl.testem3_primaries(nprim, emin=energy, emax=energy * 1.0001, seed=500 + b)
        st = ob.State(oprob, 1 << 12, stream_id=b + 1)  # independent RNG streams per batch
        ob.run(st, prim)
        return st.get_f64("calo_edep")

    gm, gs = _batches(gpu_batch, nbatch)
    om, os_ = _batches(oracle_batch, nbatch)
    ndf, p = _compare(gm, gs, om, os_, "TestEm3")
    assert ndf > 20
    assert p > 1e-3


def test_cylinder_mixed_vs_fp64_oracle(amd):
    physics = amd_pkg.generate_physics(wl.cylinder_physics_spec())
    nbatch, nprim, energy = 10, 4, 5e3
    ndet = max(wl.cylinder_geometry()[2]) + 1

    def problem(seed, precision):
        return wl.cylinder_problem(physics=physics, seed=seed, init_capacity=1 << 18,
                                   max_events=1 << 12, precision=precision)

    stepper = amd.Stepper(amd.CoreParams(problem(1001, "mixed")), 1 << 14)
    oprob = ob.Problem(problem(2002, "fp64"))

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

    gm, gs = _batches(gpu_batch, nbatch)
    om, os_ = _batches(oracle_batch, nbatch)
    ndf, p = _compare(gm, gs, om, os_, "cylinder")
    assert ndf > 5
    assert p > 1e-3
