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
