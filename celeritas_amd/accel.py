"""Offload front end: :class:`LocalTransporter` (accel/LocalTransporter.hh).

Buffers tracks handed over by a host application (``Push``), transports
them when the buffer reaches ``auto_flush`` or on ``Flush``, and reseeds the
track RNG streams per event (``InitializeEvent``). The Geant4 ``G4Track``
conversion is the caller's: tracks are given as primaries
(``PRIMARY_DTYPE`` records).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import PRIMARY_DTYPE, CeleritasError, CoreParams, CoreState, _check, _load


class LocalTransporter:
    def __init__(self, params: CoreParams, max_num_tracks: int, auto_flush: int = 0,
                 max_steps: int = 1_000_000, stream_id: int = 0, nonfatal_flush: bool = False,
                 action_times: bool = False):
        lib = _load()
        u32 = ctypes.c_uint32
        sigs = {
            "celer_transporter_create": (ctypes.c_void_p, [ctypes.c_void_p, u32, u32, u32, u32,
                                                           ctypes.c_int, ctypes.c_int]),
            "celer_transporter_destroy": (None, [ctypes.c_void_p]),
            "celer_transporter_initialize_event": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
            "celer_transporter_push": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, u32]),
            "celer_transporter_flush": (ctypes.c_int, [ctypes.c_void_p]),
            "celer_transporter_finalize": (ctypes.c_int, [ctypes.c_void_p]),
            "celer_transporter_buffer_size": (u32, [ctypes.c_void_p]),
            "celer_transporter_state": (ctypes.c_void_p, [ctypes.c_void_p]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        self._lib = lib
        self.params = params
        self._h = lib.celer_transporter_create(params.handle, max_num_tracks, auto_flush,
                                               max_steps, stream_id, int(nonfatal_flush),
                                               int(action_times))
        if not self._h:
            raise CeleritasError(lib.celer_amd_last_error().decode())
        self.state = CoreState(params, _handle=lib.celer_transporter_state(self._h), _owner=self)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.celer_transporter_destroy(self._h)
            self._h = None

    def InitializeEvent(self, event_id: int):  # noqa: N802 (reference name)
        _check(self._lib.celer_transporter_initialize_event(self._h, event_id))

    def Push(self, tracks: np.ndarray):  # noqa: N802
        p = np.ascontiguousarray(np.atleast_1d(tracks), dtype=PRIMARY_DTYPE)
        _check(self._lib.celer_transporter_push(self._h, p.ctypes.data, len(p)))

    def Flush(self):  # noqa: N802
        _check(self._lib.celer_transporter_flush(self._h))

    def Finalize(self):  # noqa: N802
        _check(self._lib.celer_transporter_finalize(self._h))

    def GetBufferSize(self) -> int:  # noqa: N802
        return int(self._lib.celer_transporter_buffer_size(self._h))
