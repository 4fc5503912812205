"""Uniform-field propagation cases in the two-boxes geometry.

Golden values are the reference's FieldPropagator tests
(test/celeritas/field/FieldPropagator.test.cc): inner box +-5 cm inside a
+-500 cm world, field along +z, particle mass 0.5109989461 MeV
(:95-110). Each case lists the track, the field and a sequence of step
requests (0 crosses the boundary); the checks live in the tests.
"""

import math

import numpy as np

from tests import problems

MASS = 0.5109989461
#: Field giving a 10 MeV e-/e+ a 1 cm radius (FieldPropagator.test.cc:170)
UNIT_RADIUS_FIELD = 3.5019461121752274
#: Lorentz coefficient: curvature [1/cm] per tesla for p = 1 MeV/c, q = 1 e
LORENTZ = 2.99792458
#: coarse tolerance of the reference's tests (corecel/math/SoftEqual.hh)
COARSE_EPS = 1e-3


def momentum(energy, mass=MASS):
    return math.sqrt(energy * (energy + 2 * mass))


def two_boxes():
    return problems.two_boxes_track_init()


def radius_10mev_step():
    """Quarter-turn splits of the electron_interior case (:177-269)."""
    radius = 3.8085385437789383
    quarter = 0.5 * math.pi * radius
    steps = [1e-2] + [(quarter - 1e-2) / 25] * 25 + [quarter, math.pi * radius]
    return radius, steps
