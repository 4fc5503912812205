"""celeritas_amd: MI355X-native per-step particle transport loop.

Python mirror of the host API over the C ABI in ``include/celeritas_amd.h``
(``libceleritas_amd.so``). The classes keep the reference's names and
argument meanings: :class:`CoreParams` (celeritas/global/CoreParams.hh),
:class:`CoreState` (celeritas/global/CoreState.hh), :class:`Stepper`
(celeritas/global/Stepper.hh) and :class:`LocalTransporter`
(accel/LocalTransporter.hh). There is no CPU execution path: every
transport call runs HIP kernels and raises if the library or a GPU is
missing.
"""

from __future__ import annotations

import ctypes
import json
import os
from typing import Iterable, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libceleritas_amd.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

INVALID = 0xFFFFFFFF


class CeleritasError(RuntimeError):
    """Error raised by the native library (RuntimeError in the reference)."""


class CPrimary(ctypes.Structure):
    _fields_ = [
        ("particle_id", ctypes.c_uint32),
        ("event_id", ctypes.c_uint32),
        ("energy", ctypes.c_double),
        ("pos", ctypes.c_double * 3),
        ("dir", ctypes.c_double * 3),
        ("time", ctypes.c_double),
    ]


class CStepperResult(ctypes.Structure):
    _fields_ = [
        ("generated", ctypes.c_uint32),
        ("queued", ctypes.c_uint32),
        ("active", ctypes.c_uint32),
        ("alive", ctypes.c_uint32),
    ]


class CCounters(ctypes.Structure):
    _fields_ = [
        (name, ctypes.c_uint32)
        for name in (
            "num_initializers",
            "num_vacancies",
            "num_active",
            "num_secondaries",
            "num_alive",
            "num_generated",
            "num_new_tracks",
            "init_base",
            "vacancy_base",
            "prev_secondaries",
            "secondary_stack_size",
            "error_flags",
            "steps_taken",
            "num_new_neutral",
            "num_sampling_failures",
            "num_completed",
        )
    ] + [("track_steps", ctypes.c_uint64)]


# numpy dtypes matching the C structs
PRIMARY_DTYPE = np.dtype(
    [
        ("particle_id", "<u4"),
        ("event_id", "<u4"),
        ("energy", "<f8"),
        ("pos", "<f8", (3,)),
        ("dir", "<f8", (3,)),
        ("time", "<f8"),
    ]
)
INITIALIZER_DTYPE = np.dtype(
    [
        ("track_id", "<u4"),
        ("parent_id", "<u4"),
        ("event_id", "<u4"),
        ("particle_id", "<u4"),
        ("time", "<f8"),
        ("pos", "<f8", (3,)),
        ("dir", "<f8", (3,)),
        ("energy", "<f8"),
        ("volume", "<u4"),
        ("reserved", "<u4"),
    ]
)
SECONDARY_DTYPE = np.dtype(
    [
        ("particle_id", "<u4"),
        ("pad", "<u4"),
        ("energy", "<f8"),
        ("dir", "<f8", (3,)),
    ]
)

# Per-slot fields: name -> (enum value, numpy dtype, length kind)
FIELDS = {
    "track_id": (0, np.uint32, "slot"),
    "parent_id": (1, np.uint32, "slot"),
    "event_id": (2, np.uint32, "slot"),
    "num_steps": (3, np.uint32, "slot"),
    "time": (4, np.float64, "slot"),
    "status": (5, np.uint8, "slot"),
    "step_length": (6, np.float64, "slot"),
    "post_step_action": (7, np.uint8, "slot"),
    "along_step_action": (8, np.uint8, "slot"),
    "particle_id": (9, np.uint8, "slot"),
    "energy": (10, np.float64, "slot"),
    "pos_x": (11, np.float64, "slot"),
    "pos_y": (12, np.float64, "slot"),
    "pos_z": (13, np.float64, "slot"),
    "dir_x": (14, np.float64, "slot"),
    "dir_y": (15, np.float64, "slot"),
    "dir_z": (16, np.float64, "slot"),
    "volume": (17, np.uint32, "slot"),
    "surface": (18, np.uint32, "slot"),
    "interaction_mfp": (19, np.float64, "slot"),
    "sec_start": (20, np.uint32, "slot"),
    "sec_count": (21, np.uint32, "slot"),
    "energy_deposition": (22, np.float64, "slot"),
    "rng_xs0": (23, np.uint32, "slot"),
    "rng_xs1": (24, np.uint32, "slot"),
    "rng_xs2": (25, np.uint32, "slot"),
    "rng_xs3": (26, np.uint32, "slot"),
    "rng_xs4": (27, np.uint32, "slot"),
    "rng_weyl": (28, np.uint32, "slot"),
    "vacancies": (29, np.uint32, "slot"),
    "secondary_counts": (30, np.uint32, "slot+1"),
    "track_counters": (31, np.uint32, "events"),
    "initializers": (32, INITIALIZER_DTYPE, "capacity"),
    "secondaries": (33, SECONDARY_DTYPE, "stack"),
    "calo_edep": (34, np.float64, "detectors"),
    "macro_xs": (35, np.float64, "slot"),
    "dedx_range": (36, np.float64, "slot"),
    "track_slots": (37, np.uint32, "slot"),
    "action_offsets": (38, np.uint32, "actions"),
    "next_step": (39, np.float64, "slot"),
    "geo_failed": (40, np.uint8, "slot"),
}

TRACK_STATUS = {"inactive": 0, "initializing": 1, "alive": 2, "errored": 3, "killed": 4}

_lib = None


def _load():
    """Load the native library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CeleritasError(
            f"native library not built: {LIB_PATH} (run __graft_entry__.build())"
        )
    lib = ctypes.CDLL(LIB_PATH)
    c_void_p, c_char_p, c_uint32, c_int = (
        ctypes.c_void_p,
        ctypes.c_char_p,
        ctypes.c_uint32,
        ctypes.c_int,
    )
    sigs = {
        "celer_amd_version": (c_char_p, []),
        "celer_amd_device_available": (c_int, []),
        "celer_amd_last_error": (c_char_p, []),
        "celer_params_create": (c_void_p, [c_char_p]),
        "celer_params_destroy": (None, [c_void_p]),
        "celer_params_num_actions": (c_uint32, [c_void_p]),
        "celer_params_action_label": (c_int, [c_void_p, c_uint32, c_char_p, ctypes.c_size_t]),
        "celer_params_find_action": (c_uint32, [c_void_p, c_char_p]),
        "celer_params_action_order": (c_int, [c_void_p, c_uint32]),
        "celer_params_find_particle": (c_uint32, [c_void_p, c_int]),
        "celer_params_num_volumes": (c_uint32, [c_void_p]),
        "celer_state_create": (c_void_p, [c_void_p, c_uint32, c_uint32]),
        "celer_state_destroy": (None, [c_void_p]),
        "celer_state_size": (c_uint32, [c_void_p]),
        "celer_state_insert_primaries": (c_int, [c_void_p, c_void_p, c_uint32]),
        "celer_state_execute_action": (c_int, [c_void_p, c_uint32]),
        "celer_state_counters": (c_int, [c_void_p, ctypes.POINTER(CCounters)]),
        "celer_state_set_counters": (c_int, [c_void_p, ctypes.POINTER(CCounters)]),
        "celer_state_copy_field": (c_int, [c_void_p, c_int, c_int, c_void_p, ctypes.c_size_t]),
        "celer_state_reseed": (c_int, [c_void_p, ctypes.c_uint64]),
        "celer_state_rng_initialize": (c_int, [c_void_p, c_void_p, c_void_p]),
        "celer_state_rng_draw": (c_int, [c_void_p, c_uint32, c_int, c_void_p]),
        "celer_state_reset": (c_int, [c_void_p]),
        "celer_state_synchronize": (c_int, [c_void_p]),
        "celer_stepper_create": (c_void_p, [c_void_p, c_uint32, c_uint32, c_int]),
        "celer_stepper_destroy": (None, [c_void_p]),
        "celer_stepper_step": (c_int, [c_void_p, c_void_p, c_uint32, ctypes.POINTER(CStepperResult)]),
        "celer_stepper_warm_up": (c_int, [c_void_p]),
        "celer_stepper_kill_active": (c_int, [c_void_p]),
        "celer_stepper_reseed": (c_int, [c_void_p, ctypes.c_uint64]),
        "celer_stepper_state": (c_void_p, [c_void_p]),
        "celer_stepper_action_sequence": (c_int, [c_void_p, c_char_p, ctypes.c_size_t]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def lib():
    return _load()


def _check(rc: int):
    if rc != 0:
        raise CeleritasError(_load().celer_amd_last_error().decode())


def device_available() -> bool:
    return bool(_load().celer_amd_device_available())


def make_primaries(records: Iterable[dict]) -> np.ndarray:
    """Build a primary array from dicts with particle_id, energy, pos, dir."""
    recs = list(records)
    arr = np.zeros(len(recs), dtype=PRIMARY_DTYPE)
    for i, r in enumerate(recs):
        arr[i]["particle_id"] = r["particle_id"]
        arr[i]["event_id"] = r.get("event_id", 0)
        arr[i]["energy"] = r["energy"]
        arr[i]["pos"] = r["pos"]
        arr[i]["dir"] = r["dir"]
        arr[i]["time"] = r.get("time", 0.0)
    return arr


class CoreParams:
    """Shared problem data built from a JSON description."""

    def __init__(self, problem: dict | str):
        text = problem if isinstance(problem, str) else json.dumps(problem)
        self._text = text
        self._h = _load().celer_params_create(text.encode())
        if not self._h:
            raise CeleritasError(_load().celer_amd_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            _load().celer_params_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def problem_json(self) -> str:
        return self._text

    def num_actions(self) -> int:
        return _load().celer_params_num_actions(self._h)

    def action_label(self, action_id: int) -> str:
        buf = ctypes.create_string_buffer(256)
        _check(_load().celer_params_action_label(self._h, action_id, buf, 256))
        return buf.value.decode()

    def action_labels(self) -> list[str]:
        return [self.action_label(i) for i in range(self.num_actions())]

    def find_action(self, label: str) -> int:
        return _load().celer_params_find_action(self._h, label.encode())

    def find_particle(self, pdg: int) -> int:
        return _load().celer_params_find_particle(self._h, pdg)

    def num_volumes(self) -> int:
        return _load().celer_params_num_volumes(self._h)


class CoreState:
    """Per-stream device state (standalone or owned by a Stepper)."""

    def __init__(self, params: CoreParams, num_slots: int = 0, stream_id: int = 0,
                 _handle=None, _owner=None):
        self.params = params
        self._owner = _owner
        if _handle is not None:
            self._h = _handle
            self._owned = False
        else:
            self._h = _load().celer_state_create(params.handle, stream_id, num_slots)
            if not self._h:
                raise CeleritasError(_load().celer_amd_last_error().decode())
            self._owned = True
        self.size = _load().celer_state_size(self._h)

    def __del__(self):
        if getattr(self, "_owned", False) and getattr(self, "_h", None):
            _load().celer_state_destroy(self._h)
            self._h = None

    def insert_primaries(self, primaries: np.ndarray):
        primaries = np.ascontiguousarray(primaries, dtype=PRIMARY_DTYPE)
        _check(_load().celer_state_insert_primaries(
            self._h, primaries.ctypes.data, len(primaries)))

    def execute(self, action: int | str):
        aid = self.params.find_action(action) if isinstance(action, str) else action
        if aid == INVALID:
            raise CeleritasError(f"no action '{action}'")
        _check(_load().celer_state_execute_action(self._h, aid))

    def counters(self) -> dict:
        c = CCounters()
        _check(_load().celer_state_counters(self._h, ctypes.byref(c)))
        return {name: getattr(c, name) for name, _ in CCounters._fields_}

    def set_counters(self, values: dict):
        c = CCounters()
        cur = self.counters()
        for name, _ in CCounters._fields_:
            setattr(c, name, int(values.get(name, cur[name])))
        _check(_load().celer_state_set_counters(self._h, ctypes.byref(c)))

    def _field_len(self, kind: str) -> int:
        if kind == "slot":
            return self.size
        if kind == "slot+1":
            return self.size + 1
        raise CeleritasError(f"length of field kind '{kind}' must be given")

    def get(self, name: str, count: Optional[int] = None) -> np.ndarray:
        fid, dtype, kind = FIELDS[name]
        n = count if count is not None else self._field_len(kind)
        out = np.zeros(n, dtype=dtype)
        _check(_load().celer_state_copy_field(self._h, fid, 0, out.ctypes.data, out.nbytes))
        return out

    def set(self, name: str, values: np.ndarray):
        fid, dtype, _ = FIELDS[name]
        arr = np.ascontiguousarray(values, dtype=dtype)
        _check(_load().celer_state_copy_field(self._h, fid, 1, arr.ctypes.data, arr.nbytes))

    def reseed(self, event_id: int):
        _check(_load().celer_state_reseed(self._h, event_id))

    def set_block_size(self, block_size: int):
        """Workgroup size of the step kernels (64, 128 or 256)."""
        f = _load().celer_state_set_block_size
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        f.restype = ctypes.c_int
        _check(f(self._h, block_size))

    def rng_initialize(self, subsequence: Sequence[int], offset: Sequence[int]):
        sub = np.ascontiguousarray(subsequence, dtype=np.uint64)
        off = np.ascontiguousarray(offset, dtype=np.uint64)
        _check(_load().celer_state_rng_initialize(self._h, sub.ctypes.data, off.ctypes.data))

    def rng_draw(self, draws: int, kind: str = "u32") -> np.ndarray:
        k = {"u32": 0, "f64": 1, "f32": 2}[kind]
        dtype = {0: np.uint32, 1: np.float64, 2: np.float32}[k]
        out = np.zeros((self.size, draws), dtype=dtype)
        _check(_load().celer_state_rng_draw(self._h, draws, k, out.ctypes.data))
        return out

    def rng_words(self) -> np.ndarray:
        """RNG state as (size, 6) words x0..x4, d."""
        names = ["rng_xs0", "rng_xs1", "rng_xs2", "rng_xs3", "rng_xs4", "rng_weyl"]
        return np.stack([self.get(n) for n in names], axis=1)

    def reset(self):
        _check(_load().celer_state_reset(self._h))

    def synchronize(self):
        _check(_load().celer_state_synchronize(self._h))


class StepperResult(dict):
    def __bool__(self):
        return self["queued"] > 0 or self["alive"] > 0


class Stepper:
    """Step iterations over one device state."""

    def __init__(self, params: CoreParams, num_track_slots: int, stream_id: int = 0,
                 action_times: bool = False):
        self.params = params
        self._h = _load().celer_stepper_create(
            params.handle, stream_id, num_track_slots, int(action_times))
        if not self._h:
            raise CeleritasError(_load().celer_amd_last_error().decode())
        self.state = CoreState(params, _handle=_load().celer_stepper_state(self._h),
                               _owner=self)

    def __del__(self):
        if getattr(self, "_h", None):
            _load().celer_stepper_destroy(self._h)
            self._h = None

    def __call__(self, primaries: Optional[np.ndarray] = None) -> StepperResult:
        r = CStepperResult()
        if primaries is not None and len(primaries) > 0:
            p = np.ascontiguousarray(primaries, dtype=PRIMARY_DTYPE)
            _check(_load().celer_stepper_step(self._h, p.ctypes.data, len(p), ctypes.byref(r)))
        else:
            _check(_load().celer_stepper_step(self._h, None, 0, ctypes.byref(r)))
        return StepperResult(generated=r.generated, queued=r.queued, active=r.active,
                             alive=r.alive)

    def warm_up(self):
        _check(_load().celer_stepper_warm_up(self._h))

    def capture_step(self):
        """Replay later steps without primaries from a captured hipGraph."""
        f = _load().celer_stepper_capture_step
        f.argtypes = [ctypes.c_void_p]
        f.restype = ctypes.c_int
        _check(f(self._h))

    def kill_active(self):
        _check(_load().celer_stepper_kill_active(self._h))

    def reseed(self, event_id: int):
        _check(_load().celer_stepper_reseed(self._h, event_id))

    def action_sequence(self) -> list[str]:
        buf = ctypes.create_string_buffer(8192)
        _check(_load().celer_stepper_action_sequence(self._h, buf, 8192))
        return [s for s in buf.value.decode().split("\n") if s]

    def diagnostics(self, enable: Optional[bool] = None, reset: bool = False) -> dict:
        """Per-kernel work counters of the stepper's state (see
        celer_state_diagnostics); ``enable`` turns them on or off."""
        f = _load().celer_state_diagnostics
        f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                      ctypes.c_int]
        f.restype = ctypes.c_int
        flag = -1 if enable is None else int(bool(enable))
        if enable is False:
            _check(f(self.state._h, 0, None, 0, 0))
            return {}
        out = np.zeros(8, np.uint64)
        _check(f(self.state._h, flag, out.ctypes.data, 8, int(reset)))
        names = ("pre_step", "along_charged", "along_neutral", "boundary", "discrete",
                 "interaction", "discrete_in_charged")
        return {k: int(v) for k, v in zip(names, out)}

    def action_times(self, reset: bool = False) -> dict:
        """Accumulated device seconds and timed-step counts per step action
        (stepper created with ``action_times=True``)."""
        labels = self.action_sequence()
        n = len(labels)
        secs = np.zeros(n, np.float64)
        counts = np.zeros(n, np.uint64)
        f = _load().celer_stepper_action_times
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                      ctypes.c_int]
        f.restype = ctypes.c_int
        _check(f(self._h, secs.ctypes.data, counts.ctypes.data, n, int(reset)))
        return {lab: (float(s), int(c)) for lab, s, c in zip(labels, secs, counts)}

    def run_steps(self, count: int) -> StepperResult:
        """Take ``count`` steps without primaries with one counter readback."""
        f = _load().celer_stepper_run_steps
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        f.restype = ctypes.c_int
        r = CStepperResult()
        _check(f(self._h, count, ctypes.byref(r)))
        return StepperResult(generated=r.generated, queued=r.queued, active=r.active,
                             alive=r.alive)

    def run(self, primaries: np.ndarray, max_steps: int = 100_000,
            steps_per_sync: int = 1) -> dict:
        """Transport primaries to completion (celer-sim's Transporter loop:
        Runner.cc/Transporter.cc); returns step iterations and the sum of
        active tracks over iterations (track-steps). With steps_per_sync > 1
        the steps after the first are launched in batches with one counter
        readback each (the counts then come from the device's step and
        track-step counters; a batch may end with a few empty steps)."""
        if steps_per_sync > 1:
            c0 = self.state.counters()
            result = self(primaries)
            launched = 1
            while result:
                if launched >= max_steps:
                    raise CeleritasError(f"transport did not finish in {max_steps} steps")
                result = self.run_steps(steps_per_sync)
                launched += steps_per_sync
            c = self.state.counters()
            return {"iterations": c["steps_taken"] - c0["steps_taken"],
                    "track_steps": int(c["track_steps"] - c0["track_steps"])}
        result = self(primaries)
        iters, track_steps = 1, result["active"]
        while result:
            if iters >= max_steps:
                raise CeleritasError(f"transport did not finish in {max_steps} steps")
            result = self()
            iters += 1
            track_steps += result["active"]
        return {"iterations": iters, "track_steps": int(track_steps)}


MODEL_CLASSES = {
    "klein_nishina": 0,
    "livermore_pe": 1,
    "bethe_heitler": 2,
    "moller_bhabha": 3,
    "seltzer_berger": 4,
    "relativistic_brem": 5,
    "eplusgg": 6,
    "combined_brem": 7,
}


def verify_interactor(params: CoreParams, state: CoreState, model: str, particle: int,
                      energies, dirs, uniforms, stride: int, material: int = 0,
                      element: int = 0) -> np.ndarray:
    """Run a device interactor on replayed uniforms (17 doubles per sample)."""
    lib_ = _load()
    f = lib_.celer_verify_interactor
    f.argtypes = [ctypes.c_void_p] * 2 + [ctypes.c_uint32] * 2 + [ctypes.c_void_p] * 2 \
        + [ctypes.c_uint32] * 3 + [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                                   ctypes.c_void_p]
    f.restype = ctypes.c_int
    lib_.celer_params_device_ref.restype = ctypes.c_void_p
    lib_.celer_params_device_ref.argtypes = [ctypes.c_void_p]
    lib_.celer_state_device_ref.restype = ctypes.c_void_p
    lib_.celer_state_device_ref.argtypes = [ctypes.c_void_p]
    lib_.celer_state_stream.restype = ctypes.c_void_p
    lib_.celer_state_stream.argtypes = [ctypes.c_void_p]
    e = np.ascontiguousarray(energies, np.float64)
    d = np.ascontiguousarray(dirs, np.float64)
    u = np.ascontiguousarray(uniforms, np.float64)
    n = len(e)
    out = np.zeros((n, 17), np.float64)
    rc = f(lib_.celer_params_device_ref(params.handle), lib_.celer_state_device_ref(state._h),
           MODEL_CLASSES[model], particle, e.ctypes.data, d.ctypes.data, material, element, n,
           u.ctypes.data, stride, out.ctypes.data, lib_.celer_state_stream(state._h))
    if rc != 0:
        raise CeleritasError(f"celer_verify_interactor failed with HIP error {rc}")
    return out


def verify_field_propagate(params: CoreParams, tracks, steps,
                           max_substeps: int = 0) -> np.ndarray:
    """Uniform-field propagation on the device: ``tracks`` rows are (pos[3],
    dir[3], momentum, charge, field[3]); ``steps`` is (ntracks, nsteps) of
    requests (0 crosses, -1 restarts the propagator). Returns
    (ntracks, nsteps, 10): distance, boundary, looping, pos, dir, volume."""
    lib_ = _load()
    f = lib_.celer_verify_field_propagate
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                  ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    lib_.celer_params_device_ref.restype = ctypes.c_void_p
    lib_.celer_params_device_ref.argtypes = [ctypes.c_void_p]
    t = np.ascontiguousarray(tracks, np.float64).reshape(-1, 11)
    s = np.ascontiguousarray(steps, np.float64).reshape(len(t), -1)
    out = np.zeros((len(t), s.shape[1], 10), np.float64)
    rc = f(lib_.celer_params_device_ref(params.handle), max_substeps, t.ctypes.data, len(t),
           s.ctypes.data, s.shape[1], out.ctypes.data, None)
    if rc != 0:
        raise CeleritasError(f"celer_verify_field_propagate failed with HIP error {rc}")
    return out


def verify_geo_locate(params: CoreParams, points) -> np.ndarray:
    """Volume of each point (rows of pos[3], dir[3]) as found by track
    initialization on the device; 0xffffffff where it fails."""
    lib_ = _load()
    f = lib_.celer_verify_geo_locate
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    lib_.celer_params_device_ref.restype = ctypes.c_void_p
    lib_.celer_params_device_ref.argtypes = [ctypes.c_void_p]
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 6)
    out = np.zeros(len(p), np.uint32)
    rc = f(lib_.celer_params_device_ref(params.handle), p.ctypes.data, len(p),
           out.ctypes.data, None)
    if rc != 0:
        raise CeleritasError(f"celer_verify_geo_locate failed with HIP error {rc}")
    return out


def verify_geo_trace(params: CoreParams, rays, max_segments: int = 64):
    """Trace rays (rows of pos[3], dir[3]) through the geometry on the
    device until each leaves the world: lists of (volumes, distances,
    midpoint safeties) per ray."""
    lib_ = _load()
    f = lib_.celer_verify_geo_trace
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    lib_.celer_params_device_ref.restype = ctypes.c_void_p
    lib_.celer_params_device_ref.argtypes = [ctypes.c_void_p]
    r = np.ascontiguousarray(rays, np.float64).reshape(-1, 6)
    n, m = len(r), int(max_segments)
    counts = np.zeros(n, np.uint32)
    vols = np.zeros(n * m, np.uint32)
    dists = np.zeros(n * m, np.float64)
    safe = np.zeros(n * m, np.float64)
    rc = f(lib_.celer_params_device_ref(params.handle), r.ctypes.data, n, m,
           counts.ctypes.data, vols.ctypes.data, dists.ctypes.data, safe.ctypes.data, None)
    if rc != 0:
        raise CeleritasError(f"celer_verify_geo_trace failed with HIP error {rc}")
    return [(vols[i * m:i * m + c], dists[i * m:i * m + c], safe[i * m:i * m + c])
            for i, c in enumerate(counts)]


def generate_physics(spec: dict | str) -> dict:
    """Expand a compact physics description into full tables (host only).

    The result is the "physics" section of a problem, with cross-section,
    stopping-power, range, Seltzer-Berger and Livermore tables.
    """
    lib_ = _load()
    f = lib_.celer_generate_physics
    f.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t,
                  ctypes.POINTER(ctypes.c_size_t)]
    f.restype = ctypes.c_int
    text = (spec if isinstance(spec, str) else json.dumps(spec)).encode()
    needed = ctypes.c_size_t(0)
    _check(f(text, None, 0, ctypes.byref(needed)))
    buf = ctypes.create_string_buffer(needed.value)
    _check(f(text, buf, needed.value, ctypes.byref(needed)))
    return json.loads(buf.value.decode())


def _read_json_text(fn, *args) -> dict:
    """Call a C entry point that fills a caller buffer with JSON text."""
    fn.restype = ctypes.c_int
    needed = ctypes.c_size_t(0)
    _check(fn(*args, None, 0, ctypes.byref(needed)))
    buf = ctypes.create_string_buffer(needed.value)
    _check(fn(*args, buf, needed.value, ctypes.byref(needed)))
    return json.loads(buf.value.decode())


def kernel_attributes() -> list:
    """Attributes of the step kernels launched so far in this process (one
    entry per kernel and launch shape): registers, scratch, LDS, resident
    workgroups per CU, waves per SIMD and occupancy, as the runtime reports
    them (the reference's KernelRegistry, corecel/sys/KernelAttributes.hh)."""
    f = _load().celer_kernel_attributes_json
    f.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
    return _read_json_text(f)


def kernel_timing(enable: bool) -> None:
    """Time every following step-kernel launch with events attached to the
    dispatch itself (hipExtLaunchKernel), or stop timing."""
    f = _load().celer_kernel_timing
    f.argtypes = [ctypes.c_int]
    f.restype = ctypes.c_int
    _check(f(int(bool(enable))))


def kernel_times(reset: bool = False) -> list:
    """Device seconds and launch counts per kernel and launch shape of the
    timed launches (waits for them to end)."""
    f = _load().celer_kernel_times_json
    f.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                  ctypes.POINTER(ctypes.c_size_t)]
    return _read_json_text(f, int(bool(reset)))


def read_seltzer_berger(z: int, directory: str | None = None) -> dict:
    """Seltzer-Berger table of element ``z`` from a G4EMLOW ``brem_SB``
    folder (``$G4LEDATA/brem_SB`` when ``directory`` is None), as the
    element's ``sb_table`` ({"x", "y", "values"}).

    Mirrors ``SeltzerBergerReader::operator()(AtomicNumber)``
    (io/SeltzerBergerReader.cc:54-102).
    """
    f = _load().celer_read_seltzer_berger
    f.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                  ctypes.POINTER(ctypes.c_size_t)]
    return _read_json_text(f, directory.encode() if directory else None, int(z))


def read_livermore_pe(z: int, directory: str | None = None) -> dict:
    """Livermore photoelectric data of element ``z`` from a G4EMLOW
    ``livermore/phot_epics2014`` folder (from ``$G4LEDATA`` when
    ``directory`` is None), as the element's ``livermore_pe`` object.

    Mirrors ``LivermorePEReader::operator()(AtomicNumber)``
    (io/LivermorePEReader.cc:50-209).
    """
    f = _load().celer_read_livermore_pe
    f.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t,
                  ctypes.POINTER(ctypes.c_size_t)]
    return _read_json_text(f, directory.encode() if directory else None, int(z))
