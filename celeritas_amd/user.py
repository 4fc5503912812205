"""User actions over the C ABI: diagnostics and detector step points.

Python mirror of the reference's user-side observers of the step loop:
:class:`ActionDiagnostic` (celeritas/user/ActionDiagnostic.hh),
:class:`StepDiagnostic` (celeritas/user/StepDiagnostic.hh) and
:class:`StepCollector` with :class:`StepSelection`
(celeritas/user/StepCollector.hh, StepData.hh, DetectorSteps.hh). Each is
created on the :class:`~celeritas_amd.CoreParams` before the stepper whose
steps it observes, like the reference's ``make_and_insert``; results
accumulate on the device until read.
"""

from __future__ import annotations

import ctypes
import json
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import INVALID, CeleritasError, CoreParams, _check, _load

_POINT_FIELDS = ("time", "pos", "dir", "volume_id", "energy")
_STEP_FIELDS = ("event_id", "parent_id", "track_step_count", "action_id", "step_length",
                "particle", "energy_deposition")


class _CPoint(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint8) for f in _POINT_FIELDS]


class _CSelection(ctypes.Structure):
    _fields_ = [("points", _CPoint * 2)] + [(f, ctypes.c_uint8) for f in _STEP_FIELDS]


@dataclass
class StepPointSelection:
    """Properties gathered at the pre- or post-step point."""

    time: bool = False
    pos: bool = False
    dir: bool = False
    volume_id: bool = False
    energy: bool = False

    @classmethod
    def all(cls):
        return cls(True, True, True, True, True)


@dataclass
class StepSelection:
    """Properties gathered every step (StepData.hh:64-120)."""

    pre: StepPointSelection = field(default_factory=StepPointSelection)
    post: StepPointSelection = field(default_factory=StepPointSelection)
    event_id: bool = False
    parent_id: bool = False
    track_step_count: bool = False
    action_id: bool = False
    step_length: bool = False
    particle: bool = False
    energy_deposition: bool = False

    @classmethod
    def all(cls):
        return cls(StepPointSelection.all(), StepPointSelection.all(),
                   *([True] * len(_STEP_FIELDS)))

    def _c(self) -> _CSelection:
        c = _CSelection()
        for i, p in enumerate((self.pre, self.post)):
            for f in _POINT_FIELDS:
                setattr(c.points[i], f, int(getattr(p, f)))
        for f in _STEP_FIELDS:
            setattr(c, f, int(getattr(self, f)))
        return c


def particle_labels(params: CoreParams) -> list[str]:
    """Particle labels of a problem, in particle-ID order."""
    doc = json.loads(params.problem_json)
    phys = doc.get("physics", {})
    if "particles" in phys:
        return [p["label"] for p in phys["particles"]]
    # Generated physics: the generator's particle list
    return ["electron", "positron", "gamma"]


def _sig(name, res, args):
    f = getattr(_load(), name)
    f.restype = res
    f.argtypes = args
    return f


class ActionDiagnostic:
    """Post-step action counts by particle (ActionDiagnostic.hh)."""

    label = "action-diagnostic"

    def __init__(self, params: CoreParams):
        self.params = params
        h = ctypes.c_void_p()
        _check(_sig("celer_action_diagnostic_create", ctypes.c_int,
                    [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)])(
                        params.handle, ctypes.byref(h)))
        self._h = h.value

    def __del__(self):
        if getattr(self, "_h", None):
            _sig("celer_action_diagnostic_destroy", None, [ctypes.c_void_p])(self._h)
            self._h = None

    def calc_actions(self) -> np.ndarray:
        """Counts [particle][action] summed over streams."""
        npart, nact = ctypes.c_uint32(), ctypes.c_uint32()
        _check(_sig("celer_action_diagnostic_shape", ctypes.c_int,
                    [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p])(
                        self._h, ctypes.byref(npart), ctypes.byref(nact)))
        out = np.zeros((npart.value, nact.value), dtype=np.uint64)
        if out.size:
            _check(_sig("celer_action_diagnostic_counts", ctypes.c_int,
                        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t])(
                            self._h, out.ctypes.data, out.size))
        return out

    def calc_actions_map(self) -> dict:
        """Nonzero counts keyed by '<action label> <particle label>'
        (ActionDiagnostic.cc:139-165), in sorted key order."""
        counts = self.calc_actions()
        labels = particle_labels(self.params)
        actions = self.params.action_labels()
        result = {}
        for p in range(counts.shape[0]):
            for a in range(counts.shape[1]):
                if counts[p, a]:
                    result[f"{actions[a]} {labels[p]}"] = int(counts[p, a])
        return dict(sorted(result.items()))

    def clear(self):
        _check(_sig("celer_action_diagnostic_clear", ctypes.c_int, [ctypes.c_void_p])(self._h))


class StepDiagnostic:
    """Steps per killed track by particle (StepDiagnostic.hh)."""

    label = "step-diagnostic"

    def __init__(self, params: CoreParams, max_step_bin: int):
        self.params = params
        h = ctypes.c_void_p()
        _check(_sig("celer_step_diagnostic_create", ctypes.c_int,
                    [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)])(
                        params.handle, max_step_bin, ctypes.byref(h)))
        self._h = h.value

    def __del__(self):
        if getattr(self, "_h", None):
            _sig("celer_step_diagnostic_destroy", None, [ctypes.c_void_p])(self._h)
            self._h = None

    def calc_steps(self) -> np.ndarray:
        """Counts [particle][num_steps] (last bin: overflow)."""
        npart, nbins = ctypes.c_uint32(), ctypes.c_uint32()
        _check(_sig("celer_step_diagnostic_shape", ctypes.c_int,
                    [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p])(
                        self._h, ctypes.byref(npart), ctypes.byref(nbins)))
        out = np.zeros((npart.value, nbins.value), dtype=np.uint64)
        _check(_sig("celer_step_diagnostic_counts", ctypes.c_int,
                    [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t])(
                        self._h, out.ctypes.data, out.size))
        return out

    def clear(self):
        _check(_sig("celer_step_diagnostic_clear", ctypes.c_int, [ctypes.c_void_p])(self._h))


#: Columns of DetectorStepOutput: name -> (dtype, values per hit)
_COLUMNS = {
    "detector": (np.uint32, 1), "track_id": (np.uint32, 1), "event_id": (np.uint32, 1),
    "parent_id": (np.uint32, 1), "track_step_count": (np.uint32, 1),
    "action_id": (np.uint32, 1), "particle": (np.uint32, 1),
    "step_length": (np.float64, 1), "energy_deposition": (np.float64, 1),
}
for _p in ("pre", "post"):
    _COLUMNS.update({f"{_p}.time": (np.float64, 1), f"{_p}.pos": (np.float64, 3),
                     f"{_p}.dir": (np.float64, 3), f"{_p}.volume_id": (np.uint32, 1),
                     f"{_p}.energy": (np.float64, 1)})


class StepCollector:
    """Detector step points gathered on the device (StepCollector.hh)."""

    def __init__(self, params: CoreParams, selection: StepSelection,
                 volume_to_detector: Optional[Sequence[int]] = None,
                 nonzero_energy_deposition: bool = False, hit_capacity: int = 1 << 20):
        self.params = params
        self.selection = selection
        c = selection._c()
        vmap = None
        nvol = 0
        if volume_to_detector is not None:
            vmap = np.ascontiguousarray(
                [INVALID if d is None or d < 0 else d for d in volume_to_detector],
                dtype=np.uint32)
            nvol = len(vmap)
        h = ctypes.c_void_p()
        _check(_sig("celer_step_collector_create", ctypes.c_int,
                    [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                     ctypes.c_int, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)])(
                        params.handle, ctypes.byref(c),
                        None if vmap is None else vmap.ctypes.data, nvol,
                        int(nonzero_energy_deposition), hit_capacity, ctypes.byref(h)))
        self._h = h.value
        self._vmap = vmap

    def __del__(self):
        if getattr(self, "_h", None):
            _sig("celer_step_collector_destroy", None, [ctypes.c_void_p])(self._h)
            self._h = None

    def _selected(self, name: str) -> bool:
        if name in ("detector", "track_id"):
            return True
        if "." in name:
            point, f = name.split(".")
            return bool(getattr(getattr(self.selection, point), f))
        return bool(getattr(self.selection, name))

    def read(self, stream_id: int = 0) -> dict:
        """Copy and clear the hits of a stream: a dict of numpy columns for
        the selected fields, plus 'overflow_steps' and 'dropped_hits'."""
        h = ctypes.c_void_p()
        _check(_sig("celer_step_collector_read", ctypes.c_int,
                    [ctypes.c_void_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_void_p)])(
                        self._h, stream_id, ctypes.byref(h)))
        try:
            n = _sig("celer_detector_steps_size", ctypes.c_uint32, [ctypes.c_void_p])(h)
            ov, dr = ctypes.c_uint32(), ctypes.c_uint32()
            _check(_sig("celer_detector_steps_overflow", ctypes.c_int,
                        [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p])(
                            h, ctypes.byref(ov), ctypes.byref(dr)))
            out = {"overflow_steps": ov.value, "dropped_hits": dr.value}
            col = _sig("celer_detector_steps_column", ctypes.c_int,
                       [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_size_t])
            for name, (dtype, width) in _COLUMNS.items():
                if not self._selected(name):
                    continue
                arr = np.zeros((n, width) if width > 1 else n, dtype=dtype)
                if n:
                    _check(col(h, name.encode(), arr.ctypes.data, arr.nbytes))
                out[name] = arr
            return out
        finally:
            _sig("celer_detector_steps_destroy", None, [ctypes.c_void_p])(h)


__all__ = ["ActionDiagnostic", "StepDiagnostic", "StepCollector", "StepSelection",
           "StepPointSelection", "particle_labels", "CeleritasError"]
