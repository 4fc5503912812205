"""ctypes binding of the oracle (oracle/_build/liboracle.so).

The oracle is test infrastructure: a CPU restatement of the reference
algorithms used as the checker for the HIP path. Only tests/, the smoke
check and bench.py's cpu_baseline leg load it.
"""

from __future__ import annotations

import ctypes
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "liboracle.so")

_lib = None


class OracleError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OracleError(f"oracle not built: {LIB_PATH}")
    lib = ctypes.CDLL(LIB_PATH)
    vp, u32, i32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_ulonglong
    sigs = {
        "orc_last_error": (ctypes.c_char_p, []),
        "orc_rng_seeded_states": (None, [u32, u32, u32, vp]),
        "orc_rng_initialize": (None, [u32, u64, u64, vp]),
        "orc_rng_discard": (None, [vp, u64]),
        "orc_rng_draw": (None, [vp, u32, i32, vp]),
        "orc_problem_create": (vp, [ctypes.c_char_p]),
        "orc_problem_destroy": (None, [vp]),
        "orc_problem_action_label": (i32, [vp, u32, ctypes.c_char_p, ctypes.c_size_t]),
        "orc_problem_num_actions": (u32, [vp]),
        "orc_state_create": (vp, [vp, u32, u32]),
        "orc_state_destroy": (None, [vp]),
        "orc_state_insert_primaries": (i32, [vp, vp, u32]),
        "orc_state_action": (i32, [vp, ctypes.c_char_p]),
        "orc_state_mock_interact": (i32, [vp, vp, vp]),
        "orc_state_counters": (None, [vp, vp]),
        "orc_state_get_u32": (u32, [vp, ctypes.c_char_p, vp, u32]),
        "orc_state_get_f64": (u32, [vp, ctypes.c_char_p, vp, u32]),
        "orc_state_set_rng": (None, [vp, vp]),
        "orc_geo_trace": (u32, [vp, vp, vp, u32, vp, vp, vp]),
    }
    for name, (res, args) in sigs.items():
        if not hasattr(lib, name):
            continue
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise OracleError(load().orc_last_error().decode())


# --------------------------------------------------------------------------- #
# RNG
# --------------------------------------------------------------------------- #
def rng_seeded_states(seed: int, stream: int, n: int) -> np.ndarray:
    out = np.zeros((n, 6), dtype=np.uint32)
    load().orc_rng_seeded_states(seed, stream, n, out.ctypes.data)
    return out


def rng_initialize(seed: int, subsequence: int, offset: int) -> np.ndarray:
    out = np.zeros(6, dtype=np.uint32)
    load().orc_rng_initialize(seed, subsequence, offset, out.ctypes.data)
    return out


def rng_discard(state6: np.ndarray, count: int) -> np.ndarray:
    s = np.array(state6, dtype=np.uint32)
    load().orc_rng_discard(s.ctypes.data, count)
    return s


def rng_draw(state6: np.ndarray, n: int, kind: str = "u32"):
    """Return (values, new_state)."""
    k = {"u32": 0, "f64": 1, "f32": 2}[kind]
    dtype = {0: np.uint32, 1: np.float64, 2: np.float32}[k]
    s = np.array(state6, dtype=np.uint32)
    out = np.zeros(n, dtype=dtype)
    load().orc_rng_draw(s.ctypes.data, n, k, out.ctypes.data)
    return out, s


# --------------------------------------------------------------------------- #
# PROBLEM / STATE
# --------------------------------------------------------------------------- #
class Problem:
    def __init__(self, problem):
        text = problem if isinstance(problem, str) else json.dumps(problem)
        self._h = load().orc_problem_create(text.encode())
        if not self._h:
            raise OracleError(load().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            load().orc_problem_destroy(self._h)

    @property
    def handle(self):
        return self._h

    def action_labels(self):
        n = load().orc_problem_num_actions(self._h)
        result = []
        for i in range(n):
            buf = ctypes.create_string_buffer(256)
            _check(load().orc_problem_action_label(self._h, i, buf, 256))
            result.append(buf.value.decode())
        return result

    def trace(self, pos, dir, max_segments=1000, safety=False):
        p = np.asarray(pos, dtype=np.float64)
        d = np.asarray(dir, dtype=np.float64)
        vols = np.zeros(max_segments, dtype=np.uint32)
        dists = np.zeros(max_segments, dtype=np.float64)
        safe = np.zeros(max_segments, dtype=np.float64) if safety else None
        n = load().orc_geo_trace(self._h, p.ctypes.data, d.ctypes.data, max_segments,
                                 vols.ctypes.data, dists.ctypes.data,
                                 safe.ctypes.data if safety else None)
        if safety:
            return vols[:n], dists[:n], safe[:n]
        return vols[:n], dists[:n]

    def locate(self, points):
        """Volume of each point (rows of pos[3], dir[3]); 0xffffffff where
        initialization fails."""
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 6)
        out = np.zeros(len(p), dtype=np.uint32)
        f = load().orc_geo_locate
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        f.restype = None
        f(self._h, p.ctypes.data, len(p), out.ctypes.data)
        return out


class State:
    def __init__(self, problem: Problem, num_slots: int, stream_id: int = 0):
        self.problem = problem
        self.size = num_slots
        self._h = load().orc_state_create(problem.handle, num_slots, stream_id)
        if not self._h:
            raise OracleError(load().orc_last_error().decode())

    def __del__(self):
        if getattr(self, "_h", None):
            load().orc_state_destroy(self._h)

    def insert_primaries(self, primaries: np.ndarray):
        from celeritas_amd import PRIMARY_DTYPE

        p = np.ascontiguousarray(primaries, dtype=PRIMARY_DTYPE)
        _check(load().orc_state_insert_primaries(self._h, p.ctypes.data, len(p)))

    def action(self, label: str):
        _check(load().orc_state_action(self._h, label.encode()))

    def mock_interact(self, num_secondaries, alive):
        n = np.ascontiguousarray(num_secondaries, dtype=np.uint32)
        a = np.ascontiguousarray(alive, dtype=np.uint8)
        _check(load().orc_state_mock_interact(self._h, n.ctypes.data, a.ctypes.data))

    def counters(self) -> dict:
        out = np.zeros(6, dtype=np.uint32)
        load().orc_state_counters(self._h, out.ctypes.data)
        names = ["num_initializers", "num_vacancies", "num_active", "num_secondaries",
                 "num_alive", "num_generated"]
        return dict(zip(names, (int(v) for v in out)))

    def get_u32(self, name: str, max_count: int = 1 << 22) -> np.ndarray:
        out = np.zeros(max_count, dtype=np.uint32)
        n = load().orc_state_get_u32(self._h, name.encode(), out.ctypes.data, max_count)
        return out[:n].copy()

    def get_f64(self, name: str, max_count: int = 1 << 22) -> np.ndarray:
        out = np.zeros(max_count, dtype=np.float64)
        n = load().orc_state_get_f64(self._h, name.encode(), out.ctypes.data, max_count)
        return out[:n].copy()

    def set_rng(self, words: np.ndarray):
        w = np.ascontiguousarray(words, dtype=np.uint32)
        load().orc_state_set_rng(self._h, w.ctypes.data)

    # User observers (oracle/user.hh)
    def enable_user(self, num_actions: int = 0, step_bins: int = 0, collect: bool = False,
                    volume_to_detector=None, nonzero_edep: bool = False):
        f = load().orc_state_enable_user
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                      ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int]
        f.restype = ctypes.c_int
        vmap = None
        if volume_to_detector is not None:
            vmap = np.ascontiguousarray(
                [0xFFFFFFFF if d is None or d < 0 else d for d in volume_to_detector],
                dtype=np.uint32)
        _check(f(self._h, num_actions, step_bins, int(collect),
                 None if vmap is None else vmap.ctypes.data,
                 0 if vmap is None else len(vmap), int(nonzero_edep)))
        self._num_actions = num_actions
        self._step_bins = step_bins
        self._vmap = vmap

    def _user_counts(self, which: int, width: int) -> np.ndarray:
        f = load().orc_state_user_counts
        f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32]
        f.restype = ctypes.c_int
        out = np.zeros(64 * width, dtype=np.uint64)
        _check(f(self._h, which, out.ctypes.data, len(out)))
        return out

    def action_counts(self, num_particles: int) -> np.ndarray:
        out = self._user_counts(0, self._num_actions)
        return out[: num_particles * self._num_actions].reshape(num_particles, -1)

    def step_counts(self, num_particles: int) -> np.ndarray:
        out = self._user_counts(1, self._step_bins)
        return out[: num_particles * self._step_bins].reshape(num_particles, -1)

    def take_hits(self) -> dict:
        """Copy and clear the collected hits as DetectorStepOutput-like
        columns (every field; the caller picks the selected ones)."""
        n = load().orc_state_num_hits
        n.argtypes = [ctypes.c_void_p]
        n.restype = ctypes.c_uint32
        count = n(self._h)
        u = np.zeros((count, 9), dtype=np.uint32)
        d = np.zeros((count, 18), dtype=np.float64)
        f = load().orc_state_take_hits
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
        f.restype = ctypes.c_int
        _check(f(self._h, u.ctypes.data if count else None, d.ctypes.data if count else None,
                 count))
        out = {"detector": u[:, 0], "track_id": u[:, 1], "event_id": u[:, 2],
               "parent_id": u[:, 3], "track_step_count": u[:, 4], "action_id": u[:, 5],
               "particle": u[:, 6], "pre.volume_id": u[:, 7], "post.volume_id": u[:, 8],
               "step_length": d[:, 0], "energy_deposition": d[:, 1]}
        for i, p in enumerate(("pre", "post")):
            q = d[:, 2 + 8 * i: 10 + 8 * i]
            out[f"{p}.time"] = q[:, 0]
            out[f"{p}.pos"] = q[:, 1:4]
            out[f"{p}.dir"] = q[:, 4:7]
            out[f"{p}.energy"] = q[:, 7]
        return out


def rng_draw_batch(states: np.ndarray, draws: int, kind: str = "u32"):
    k = {"u32": 0, "f64": 1, "f32": 2}[kind]
    dtype = {0: np.uint32, 1: np.float64, 2: np.float32}[k]
    st = np.ascontiguousarray(states, dtype=np.uint32).copy()
    out = np.zeros((len(st), draws), dtype=dtype)
    f = load().orc_rng_draw_batch
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p]
    f(st.ctypes.data, len(st), draws, k, out.ctypes.data)
    return out, st


def rng_initialize_batch(seed: int, subsequence, offset) -> np.ndarray:
    sub = np.ascontiguousarray(subsequence, dtype=np.uint64)
    off = np.ascontiguousarray(offset, dtype=np.uint64)
    out = np.zeros((len(sub), 6), dtype=np.uint32)
    f = load().orc_rng_initialize_batch
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_void_p]
    f(seed, len(sub), sub.ctypes.data, off.ctypes.data, out.ctypes.data)
    return out


#: The oracle's initializer records (the reference's TrackInitializer fields)
ORACLE_INITIALIZER_DTYPE = np.dtype(
    [
        ("track_id", "<u4"),
        ("parent_id", "<u4"),
        ("event_id", "<u4"),
        ("particle_id", "<u4"),
        ("time", "<f8"),
        ("pos", "<f8", (3,)),
        ("dir", "<f8", (3,)),
        ("energy", "<f8"),
    ]
)


def state_initializers(state: "State") -> np.ndarray:
    n = state.counters()["num_initializers"]
    out = np.zeros(max(n, 1), dtype=ORACLE_INITIALIZER_DTYPE)
    f = load().orc_state_initializers
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
    f.restype = ctypes.c_uint32
    f(state._h, out.ctypes.data, len(out))
    return out[:n]


def run(state: "State", primaries: np.ndarray, max_iters: int = 1_000_000):
    """StepperTestBase::run on the oracle: (active, queued) per step."""
    from celeritas_amd import PRIMARY_DTYPE

    p = np.ascontiguousarray(primaries, dtype=PRIMARY_DTYPE)
    act = np.zeros(max_iters, np.uint32)
    q = np.zeros(max_iters, np.uint32)
    f = load().orc_run
    f.restype = ctypes.c_uint32
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                  ctypes.c_void_p, ctypes.c_void_p]
    n = f(state._h, p.ctypes.data, len(p), max_iters, act.ctypes.data, q.ctypes.data)
    if n == 0:
        raise OracleError(load().orc_last_error().decode())
    return act[:n].copy(), q[:n].copy()


def emptying_step(active):
    """StepperTestBase::RunResult::calc_emptying_step (StepperTestBase.cc:120-145)."""
    result = 0
    prev = int(active[0])
    max_cap = 0
    for i in range(1, len(active)):
        max_cap = max(prev, max_cap)
        if prev == max_cap and int(active[i]) < max_cap:
            result = i
        prev = int(active[i])
    return result


def step_once(state: "State") -> dict:
    out = np.zeros(4, np.uint32)
    f = load().orc_step_once
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    _check(f(state._h, out.ctypes.data))
    return dict(zip(("generated", "queued", "active", "alive"), (int(v) for v in out)))


def track_counters(state: "State") -> np.ndarray:
    out = np.zeros(1 << 16, np.uint32)
    f = load().orc_state_track_counters
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
    f.restype = ctypes.c_uint32
    n = f(state._h, out.ctypes.data, len(out))
    return out[:n].copy()


INTERACTION_FIELDS = 17


def interactor_sample(problem, model, particle, energy, direction, n, material=0,
                      element=0, stack=None, uniforms=None, stride=0):
    """Sample n interactions; returns (records[n, 17], mt_draw_count)."""
    out = np.zeros((n, INTERACTION_FIELDS), np.float64)
    d = np.asarray(direction, np.float64)
    cnt = ctypes.c_ulonglong(0)
    f = load().orc_interactor_sample
    f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_double,
                  ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                  ctypes.c_uint32, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint32,
                  ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulonglong)]
    f.restype = ctypes.c_int
    kind = 0 if uniforms is None else 1
    u = None if uniforms is None else np.ascontiguousarray(uniforms, np.float64)
    _check(f(problem.handle, model.encode(), particle, energy, d.ctypes.data, material,
             element, n, stack if stack is not None else 2 * n, kind,
             None if u is None else u.ctypes.data, stride, out.ctypes.data,
             ctypes.byref(cnt)))
    return out, cnt.value


def mt_uniforms(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float64)
    f = load().orc_mt_uniforms
    f.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_void_p]
    f(seed, n, out.ctypes.data)
    return out


def interactor_sequence(problem, items, count=False):
    """Sample items [(model, particle, energy, dir[, repeat, stack_reset])]
    from one default-seeded mt19937 stream; returns records (and draws)."""
    models = b"".join(it[0].encode() + b"\0" for it in items)
    n = len(items)
    part = np.array([it[1] for it in items], np.uint32)
    en = np.array([it[2] for it in items], np.float64)
    dirs = np.array([it[3] for it in items], np.float64).reshape(n, 3)
    mat = np.array([it[6] if len(it) > 6 else 0 for it in items], np.uint32)
    el = np.zeros(n, np.uint32)
    rep = np.array([it[4] if len(it) > 4 else 1 for it in items], np.uint32)
    reset = np.array([it[5] if len(it) > 5 else 0 for it in items], np.uint32)
    if reset[0] == 0:
        reset[0] = max(2 * int(rep.sum()), 2)
    out = np.zeros((int(rep.sum()), INTERACTION_FIELDS), np.float64)
    cnt = ctypes.c_ulonglong(0)
    f = load().orc_interactor_sequence
    f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32] + [ctypes.c_void_p] * 7 \
        + [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulonglong)]
    f.restype = ctypes.c_int
    _check(f(problem.handle, models, n, part.ctypes.data, en.ctypes.data, dirs.ctypes.data,
             mat.ctypes.data, el.ctypes.data, rep.ctypes.data, reset.ctypes.data,
             out.ctypes.data, ctypes.byref(cnt)))
    return (out, cnt.value) if count else out


def interactor_replay_batch(problem, model, particle, energies, dirs, uniforms, stride,
                            material=0, element=0):
    e = np.ascontiguousarray(energies, np.float64)
    d = np.ascontiguousarray(dirs, np.float64)
    u = np.ascontiguousarray(uniforms, np.float64)
    n = len(e)
    out = np.zeros((n, INTERACTION_FIELDS), np.float64)
    f = load().orc_interactor_replay_batch
    f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_void_p,
                  ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                  ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    f.restype = ctypes.c_int
    _check(f(problem.handle, model.encode(), particle, e.ctypes.data, d.ctypes.data,
             material, element, n, u.ctypes.data, stride, out.ctypes.data))
    return out


def field_propagate(problem: "Problem", pos, direction, momentum: float, charge: float,
                    field, steps, max_substeps: int = 0) -> np.ndarray:
    """Successive uniform-field propagations of one track through one
    propagator; a step of 0 crosses the boundary the track is on and starts
    a new propagator, -1 only starts a new propagator.

    Returns rows of (distance, boundary, looping, x, y, z, u, v, w, volume,
    cumulative Dormand-Prince steps)."""
    lib = load()
    f = lib.orc_field_propagate
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                  ctypes.c_double, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                  ctypes.c_uint32, ctypes.c_void_p]
    p = np.ascontiguousarray(pos, dtype=np.float64)
    d = np.ascontiguousarray(direction, dtype=np.float64)
    b = np.ascontiguousarray(field, dtype=np.float64)
    s = np.ascontiguousarray(steps, dtype=np.float64)
    out = np.zeros((len(s), 11))
    _check(f(problem.handle, p.ctypes.data, d.ctypes.data, momentum, charge, b.ctypes.data,
             max_substeps, s.ctypes.data, len(s), out.ctypes.data))
    return out


# --------------------------------------------------------------------------- #
# Urban MSC pieces (oracle/msc.hh)
# --------------------------------------------------------------------------- #
def msc_set_mutation(m: int) -> None:
    """Plant a deliberate mistake in the oracle's MSC code (0: none)."""
    f = load().orc_msc_set_mutation
    f.argtypes = [ctypes.c_int]
    f(m)


def msc_material_coeffs(zeff: float) -> dict:
    f = load().orc_msc_material_coeffs
    f.argtypes = [ctypes.c_double, ctypes.c_void_p]
    out = np.zeros(11)
    f(zeff, out.ctypes.data)
    keys = ("stepmin0", "stepmin1", "theta0", "theta1", "tail0", "tail1", "tail2",
            "tail_corr", "scaled_zeff_e", "scaled_zeff_p", "d_over_r")
    return dict(zip(keys, out.tolist()))


def urban_positron_correction(zeff: float, y: float) -> float:
    f = load().orc_urban_positron_correction
    f.argtypes = [ctypes.c_double, ctypes.c_double]
    f.restype = ctypes.c_double
    return f(zeff, y)


def msc_step_conversion(problem: "Problem", material: int, par: int, energy: float,
                        tstep: float, gsteps) -> tuple:
    """(range, lambda, geometric path, alpha, end energy), true paths of gsteps."""
    f = load().orc_msc_step_conversion
    f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_double,
                  ctypes.c_double, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p,
                  ctypes.c_void_p]
    f.restype = ctypes.c_int
    g = np.ascontiguousarray(gsteps, dtype=np.float64)
    info = np.zeros(5)
    out = np.zeros(len(g))
    _check(f(problem.handle, material, par, energy, tstep, g.ctypes.data, len(g),
             info.ctypes.data, out.ctypes.data))
    return tuple(info.tolist()), out


def msc_simple_scattering(xmean: float, x2mean: float, n: int) -> np.ndarray:
    f = load().orc_msc_simple_scattering
    f.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_uint32, ctypes.c_void_p]
    out = np.zeros(n)
    f(xmean, x2mean, n, out.ctypes.data)
    return out


def msc_sample_cos_theta(zeff: float, tau: float, theta0: float, true_path: float,
                         limit_min: float, lam: float, rad_length: float,
                         n: int) -> np.ndarray:
    f = load().orc_msc_sample_cos_theta
    f.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
    inp = np.array([tau, theta0, true_path, limit_min, lam, rad_length])
    out = np.zeros(n)
    f(zeff, inp.ctypes.data, n, out.ctypes.data)
    return out


def msc_tail_parts(zeff: float, tau: float, theta0: float, true_path: float,
                   limit_min: float, lam: float, rad_length: float) -> dict:
    f = load().orc_msc_tail_parts
    f.argtypes = [ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    inp = np.array([tau, theta0, true_path, limit_min, lam, rad_length])
    out = np.zeros(9)
    f(zeff, inp.ctypes.data, out.ctypes.data)
    keys = ("xsi", "x", "c", "d", "prob", "qprob", "xmean_1", "xmean_2", "ea")
    return dict(zip(keys, out.tolist()))


def run_events_parallel(problem: "Problem", primaries: np.ndarray, num_slots: int,
                        num_threads: int, num_calo: int = 0) -> dict:
    """Event-parallel transport on OpenMP threads (one state per thread)."""
    from celeritas_amd import PRIMARY_DTYPE

    f = load().orc_run_events_parallel
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                  ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32]
    f.restype = ctypes.c_int
    p = np.ascontiguousarray(primaries, dtype=PRIMARY_DTYPE)
    out = np.zeros(4)
    calo = np.zeros(max(num_calo, 1))
    _check(f(problem.handle, p.ctypes.data, len(p), num_slots, num_threads, out.ctypes.data,
             calo.ctypes.data if num_calo else None, num_calo))
    return {"track_steps": int(out[0]), "iterations": int(out[1]), "seconds": out[2],
            "threads": int(out[3]), "calo_edep": calo[:num_calo]}


def msc_step_limit_sample(mt, alg: str, on_boundary: bool, zeff: float, range_factor: float,
                          energies, ranges, lambdas, nsamples: int) -> dict:
    """UrbanMsc.test.cc's step-limit sampling through the oracle: per
    energy, the mean limit of nsamples and the final MSC range."""
    algs = {"minimal": 0, "safety": 1, "safety_plus": 2}
    f = load().orc_msc_step_limit_sample
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                  ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                  ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
    f.restype = ctypes.c_int
    e = np.ascontiguousarray(energies, np.float64)
    r = np.ascontiguousarray(ranges, np.float64)
    lam = np.ascontiguousarray(lambdas, np.float64)
    out = np.zeros((len(e), 4))
    _check(f(mt, algs[alg], int(on_boundary), zeff, range_factor, e.ctypes.data,
             r.ctypes.data, lam.ctypes.data, len(e), nsamples, out.ctypes.data))
    return {"mean_step": out[:, 0], "range_init": out[:, 1], "range_factor": out[:, 2],
            "limit_min": out[:, 3]}
