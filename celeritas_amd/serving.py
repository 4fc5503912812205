"""Continuous-batching transport of events that arrive over time.

Wraps the native serving loop (celer_server_*): events are admitted in
arrival order while the initializer queue has room, share the device with
those in flight, and complete when their device-resident live-track count
reaches zero; the device appends each completed event to a list that the
loop reads after each batch of steps.
"""

from __future__ import annotations

import ctypes
import math
from typing import Sequence

import numpy as np

from . import PRIMARY_DTYPE, CeleritasError, CoreParams, CoreState, _check, _load


class Server:
    """One track-slot state serving a stream of events."""

    def __init__(self, params: CoreParams, num_track_slots: int, stream_id: int = 0,
                 steps_per_sync: int = 4):
        lib = _load()
        lib.celer_server_create.restype = ctypes.c_void_p
        lib.celer_server_create.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32]
        lib.celer_server_destroy.argtypes = [ctypes.c_void_p]
        lib.celer_server_state.restype = ctypes.c_void_p
        lib.celer_server_state.argtypes = [ctypes.c_void_p]
        lib.celer_server_run.restype = ctypes.c_int
        lib.celer_server_run.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32] * 3 \
            + [ctypes.c_void_p] * 4
        self.params = params
        self._h = lib.celer_server_create(params.handle, num_track_slots, stream_id)
        if not self._h:
            raise CeleritasError(lib.celer_amd_last_error().decode())
        self.state = CoreState(params, _handle=lib.celer_server_state(self._h), _owner=self)
        lib.celer_server_set_steps_per_sync.restype = ctypes.c_int
        lib.celer_server_set_steps_per_sync.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        _check(lib.celer_server_set_steps_per_sync(self._h, steps_per_sync))

    def __del__(self):
        if getattr(self, "_h", None):
            _load().celer_server_destroy(self._h)
            self._h = None

    def run(self, events: Sequence[np.ndarray], arrivals: Sequence[float],
            max_queued: int = 0, max_steps: int = 0) -> dict:
        """Serve ``events`` (arrays of primaries) arriving at ``arrivals``
        [s after the call]; returns per-event admission and completion
        times, latencies and deposited energies plus the run totals."""
        n = len(events)
        if len(arrivals) != n:
            raise CeleritasError("one arrival time per event is required")
        offsets = np.zeros(n + 1, np.uint32)
        offsets[1:] = np.cumsum([len(e) for e in events])
        prim = np.ascontiguousarray(np.concatenate(events) if n else
                                    np.zeros(0, PRIMARY_DTYPE), dtype=PRIMARY_DTYPE)
        arr = np.ascontiguousarray(arrivals, np.float64)
        admit = np.zeros(n)
        done = np.zeros(n)
        edep = np.zeros(n)
        totals = np.zeros(3)
        _check(_load().celer_server_run(self._h, prim.ctypes.data, offsets.ctypes.data,
                                        arr.ctypes.data, n, max_queued, max_steps,
                                        admit.ctypes.data, done.ctypes.data, edep.ctypes.data,
                                        totals.ctypes.data))
        latency = done - arr
        return {"admit": admit, "done": done, "latency": latency, "edep": edep,
                "wall_time": totals[0], "steps": int(totals[1]),
                "track_steps": int(totals[2]),
                "events_per_s": n / totals[0] if totals[0] > 0 else math.inf,
                "p50_latency": float(np.percentile(latency, 50)) if n else 0.0,
                "p99_latency": float(np.percentile(latency, 99)) if n else 0.0}


def poisson_arrivals(n: int, rate_hz: float, seed: int = 0) -> np.ndarray:
    """Arrival times of a Poisson process (rate_hz = inf: all at zero)."""
    if not math.isfinite(rate_hz):
        return np.zeros(n)
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_hz, n))
