"""ctypes bindings of the oracle's bremsstrahlung pieces (test only)."""

from __future__ import annotations

import ctypes

import numpy as np

from tests import oracle_bind as ob

_done = False


def _lib():
    global _done
    lib = ob.load()
    if not _done:
        vp, u32, i32, f64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int, ctypes.c_double
        sigs = {
            "orc_mt_create": (vp, []),
            "orc_mt_destroy": (None, [vp]),
            "orc_sb_argmax": (i32, [vp, u32, vp, u32]),
            "orc_sb_positron_scale": (f64, [i32, f64, f64, f64, f64]),
            "orc_sb_sample_many": (i32, [vp, vp, u32, u32, f64, f64, f64, i32, u32, vp]),
            "orc_interactor_stress": (i32, [vp, vp, ctypes.c_char_p, u32, f64, vp, u32, u32,
                                            ctypes.POINTER(ctypes.c_ulonglong), vp]),
            "orc_pe_macro_xs": (f64, [vp, u32, f64]),
            "orc_rb_dxsec": (f64, [vp, u32, f64, f64]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _done = True
    return lib


class MtEngine:
    """A default-seeded std::mt19937 kept alive across oracle calls."""

    def __init__(self):
        self.h = _lib().orc_mt_create()

    def __del__(self):
        if getattr(self, "h", None):
            _lib().orc_mt_destroy(self.h)
            self.h = None


def sb_argmax(problem: ob.Problem, element: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint32)
    ob._check(_lib().orc_sb_argmax(problem.handle, element, out.ctypes.data, n))
    return out


def sb_positron_scale(z, mass, kc, inc, exit_energy) -> float:
    return _lib().orc_sb_positron_scale(z, mass, kc, inc, exit_energy)


def sb_sample_many(problem, mt: MtEngine, energy, mass, kc, corrector, n,
                   material=0, element=0):
    out = np.zeros(3)
    ob._check(_lib().orc_sb_sample_many(problem.handle, mt.h, material, element, energy, mass,
                                        kc, corrector, n, out.ctypes.data))
    return out


def interactor_stress(problem, mt: MtEngine, model, particle, energy, direction, n,
                      material=0, sums=None) -> int:
    """Run n interactions; ``sums`` (3 doubles) accumulates the secondary
    count, energy and cosine with ``direction``."""
    d = np.ascontiguousarray(direction, dtype=np.float64)
    count = ctypes.c_ulonglong(0)
    ob._check(_lib().orc_interactor_stress(problem.handle, mt.h, model.encode(), particle, energy,
                                           d.ctypes.data, material, n, ctypes.byref(count),
                                           None if sums is None else sums.ctypes.data))
    return count.value


def pe_macro_xs(problem, material, energy) -> float:
    return _lib().orc_pe_macro_xs(problem.handle, material, energy)


def rb_dxsec(problem, element, total_energy, gamma_energy) -> float:
    return _lib().orc_rb_dxsec(problem.handle, element, total_energy, gamma_energy)
