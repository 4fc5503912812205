"""Offload front end (LocalTransporter) on the device.

Mirrors the reference's accel behaviour (LocalTransporter.cc:190-300):
pushes buffer until auto_flush, Flush empties the buffer and transports to
completion, Finalize fails with buffered tracks, and max_steps either fails
or (nonfatal) kills the active tracks.
"""

import numpy as np
import pytest

import celeritas_amd as amd_pkg
from celeritas_amd import accel
from celeritas_amd import workloads as wl

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def params(amd):
    phys = amd_pkg.generate_physics(wl.testem3_physics_spec())
    return amd.CoreParams(wl.testem3_problem(physics=phys, init_capacity=1 << 18,
                                             max_events=16))


def test_buffer_and_flush(params):
    t = accel.LocalTransporter(params, 4096, auto_flush=16)
    t.InitializeEvent(3)
    prim = wl.testem3_primaries(40, seed=9, emin=100.0, emax=1000.0)
    t.Push(prim[:10])
    assert t.GetBufferSize() == 10
    with pytest.raises(amd_pkg.CeleritasError):
        t.Finalize()
    t.Push(prim[10:20])  # reaches 16: flushes everything pushed so far
    assert t.GetBufferSize() == 4
    t.Push(prim[20:])
    assert t.GetBufferSize() == 8
    t.Flush()
    assert t.GetBufferSize() == 0
    t.Finalize()
    edep = t.state.get("calo_edep", 100)
    assert 0.3 * prim["energy"].sum() < edep.sum() <= prim["energy"].sum() * (1 + 1e-9)


def test_max_steps(params):
    prim = wl.testem3_primaries(4, seed=10, emin=1e3, emax=2e3)
    t = accel.LocalTransporter(params, 4096, max_steps=5)
    t.InitializeEvent(0)
    t.Push(prim)
    with pytest.raises(amd_pkg.CeleritasError, match="exceeded"):
        t.Flush()
    t2 = accel.LocalTransporter(params, 4096, max_steps=5, nonfatal_flush=True)
    t2.InitializeEvent(0)
    t2.Push(prim)
    t2.Flush()  # kills the active tracks and finishes
    assert t2.GetBufferSize() == 0
    np.testing.assert_array_equal(t2.state.get("status") == 2, False)


def test_single_flush_equals_stepper(amd, params):
    """One flush of an event's tracks is the stepper's transport of the same
    primaries after the event's reseed (LocalTransporter.cc:199-300): the
    same kernels on the same RNG streams give the same deposition."""
    prim = wl.testem3_primaries(6, seed=21, emin=500.0, emax=2000.0)
    t = accel.LocalTransporter(params, 4096)
    t.InitializeEvent(7)
    t.Push(prim)
    t.Flush()
    t.Finalize()
    via_transporter = t.state.get("calo_edep", 100).copy()

    stepper = amd.Stepper(params, 4096)
    stepper.reseed(7)
    ev0 = prim.copy()
    ev0["event_id"] = 0  # the transporter offloads every track as event 0
    stepper.run(ev0, max_steps=1_000_000)
    direct = stepper.state.get("calo_edep", 100)
    assert via_transporter.sum() > 0
    np.testing.assert_allclose(via_transporter, direct, rtol=1e-12, atol=0)


def test_flushed_edep_vs_oracle(amd, params):
    """Per-detector deposition of events offloaded with auto-flush (several
    flushes per event) against independent oracle runs of the same
    primary distribution. Each detector's difference of batch means over
    its estimated error is a Welch t statistic; with 16 batches a side its
    tails are heavier than a normal's, so each is mapped to the normal
    quantile of its t probability before the squares are summed into a
    chi^2 over the detectors with signal (p > 1e-3)."""
    from scipy import stats

    from tests import oracle_bind as ob

    phys = amd_pkg.generate_physics(wl.testem3_physics_spec())
    oprob = ob.Problem(wl.testem3_problem(physics=phys, seed=4242, init_capacity=1 << 18,
                                          max_events=16))
    nbatch, nprim, energy = 16, 6, 2e3
    t = accel.LocalTransporter(params, 1 << 14, auto_flush=2)

    def transported(b):
        prim = wl.testem3_primaries(nprim, emin=energy, emax=energy * 1.0001, seed=700 + b)
        before = t.state.get("calo_edep", 100).copy()
        t.InitializeEvent(b)
        t.Push(prim)  # three auto-flushes of two tracks each
        assert t.GetBufferSize() == 0
        t.Flush()
        return t.state.get("calo_edep", 100) - before

    def oracle(b):
        prim = wl.testem3_primaries(nprim, emin=energy, emax=energy * 1.0001, seed=900 + b)
        st = ob.State(oprob, 1 << 12, stream_id=b + 1)  # independent RNG streams per batch
        ob.run(st, prim)
        return st.get_f64("calo_edep")

    g = np.array([transported(b) for b in range(nbatch)])
    o = np.array([oracle(b) for b in range(nbatch)])
    gm, gs = g.mean(0), g.std(0, ddof=1) / np.sqrt(nbatch)
    om, os_ = o.mean(0), o.std(0, ddof=1) / np.sqrt(nbatch)
    sel = (gs > 0) & (os_ > 0) & (gm + om > 1e-3 * (gm + om).max())
    tstat = ((gm - om) / np.sqrt(gs**2 + os_**2))[sel]
    z = stats.norm.isf(stats.t.sf(np.abs(tstat), 2 * nbatch - 2))
    chi2 = float((z**2).sum())
    ndf = int(sel.sum())
    p = stats.chi2.sf(chi2, ndf)
    print(f"transporter vs oracle: {gm.sum():.1f} / {om.sum():.1f} MeV per event, "
          f"chi2 {chi2:.1f}/{ndf}, p = {p:.3g}")
    # The totals agree within 3 sigma as well
    tot_err = np.sqrt((g.sum(1).std(ddof=1) ** 2 + o.sum(1).std(ddof=1) ** 2) / nbatch)
    assert abs(gm.sum() - om.sum()) < 3 * tot_err
    assert ndf > 20
    assert p > 1e-3
