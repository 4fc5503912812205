"""Oracle Urban MSC pieces against the reference's golden values, its own
step-conversion invariants and closed forms (CPU only).

- Steel material coefficients: UrbanMsc.test.cc:139-167, for the effective
  Z of G4_STAINLESS-STEEL computed here from four-steel-slabs.gdml (mass
  fractions at :65-72, isotope abundances and molar masses at :5-59).
- Positron correction: UrbanMsc.test.cc:55-80 (Z = 1 and Z = 74).
- Step conversion: the invariants of UrbanMsc.test.cc:188-283 (geometric
  step no longer than the true step and positive; the true step of every
  geometric step in [gstep, pstep]; true -> geometric -> true to 1e-10, or
  1e-3 where 1 - alpha pstep < 1e-8) for e- and e+ at the reference's ten
  energies, on the table generator's steel.
- simple_scattering: its mixture of (1 + mu)^a and a flat density is built
  so that <mu> = e^-tau and <mu^2> = (1 + 2 e^-2.5 tau)/3; both moments are
  checked on 2e6 samples.
- sample_cos_theta (tail model): the reference's xmean_2 is the exact mean
  of g2 (checked against quadrature of g2's inverse CDF), and the sample
  mean equals the mixture's: qprob (prob m1 + (1 - prob) m2), with m1 the
  mean of g1 from its inverse CDF. (The reference's own estimate xmean_1 =
  1 - x (1 + xsi ea / (1 - ea)), UrbanMscScatter.hh:438, is not g1's mean,
  1 - x (1 - xsi ea / (1 - ea)), so <mu> is not exactly e^-tau in this
  branch; the oracle keeps the reference's formula.)

Each pin has a planted-mistake twin: a deliberate slip switched on inside
the oracle (ob.msc_set_mutation) must make the same check fail.
"""

import math

import numpy as np
import pytest

import celeritas_amd as amd
from celeritas_amd import workloads as wl
from tests import oracle_bind as ob

# --------------------------------------------------------------------------- #
# G4_STAINLESS-STEEL from four-steel-slabs.gdml
# --------------------------------------------------------------------------- #
#: (abundance, molar mass g/mol) of each isotope, four-steel-slabs.gdml:5-59
ISOTOPES = {
    26: [(0.05845, 53.9396), (0.91754, 55.9349), (0.02119, 56.9354), (0.00282, 57.9333)],
    24: [(0.04345, 49.946), (0.83789, 51.9405), (0.09501, 52.9407), (0.02365, 53.9389)],
    28: [(0.680769, 57.9353), (0.262231, 59.9308), (0.011399, 60.9311),
         (0.036345, 61.9283), (0.009256, 63.928)],
}
#: Mass fractions, four-steel-slabs.gdml:65-72
MASS_FRACTIONS = {26: 0.746212874621521, 24: 0.169001044311525, 28: 0.0847860810669534}


def steel_number_fractions():
    molar = {z: sum(f * a for f, a in iso) / sum(f for f, _ in iso)
             for z, iso in ISOTOPES.items()}
    n = {z: MASS_FRACTIONS[z] / molar[z] for z in MASS_FRACTIONS}
    tot = sum(n.values())
    return {z: v / tot for z, v in n.items()}, molar


def steel_zeff():
    frac, _ = steel_number_fractions()
    return sum(z * f for z, f in frac.items())


#: UrbanMsc.test.cc:150-166 (steel material data and e- d_over_r)
STEEL_GOLDEN = {
    "stepmin0": 1e3 * 4.4449610414595817,
    "stepmin1": 1e3 * 1.5922149179564158,
    "theta0": 0.97326969977637379,
    "theta1": 0.044188139325421663,
    "tail0": 1.6889578380303167,
    "tail1": 2.745018223507488,
    "tail2": -2.2531516772497562,
    "tail_corr": 0.052696806851297018,
    "d_over_r": 0.64474963087322135,
}

#: UrbanMsc.test.cc:61-77
POSITRON_Y = [1e-3, 1e-2, 0.4, 0.5, 0.6, 1.0, 1.5, 2., 10., 1e2, 1e3, 1e6]
POSITRON_H = [1.378751990475, 1.3787519983432, 1.3813527280086, 1.3825378340463,
              1.3834564182635, 1.3856807011387, 1.3865656925136, 1.3865681880571,
              1.3876210627429, 1.3882415266217, 1.3882507402225, 1.3882508352478]
POSITRON_W = [0.21482671339734, 0.4833017838367, 0.70738388881252, 0.70471228941815,
              0.7026415135041, 0.69762728474033, 0.69563878645763, 0.69577660924627,
              0.75392431413533, 0.78819102317998, 0.78869986791365, 0.78870511592834]


@pytest.fixture(autouse=True)
def no_mutation():
    ob.msc_set_mutation(0)
    yield
    ob.msc_set_mutation(0)


def rel(a, b):
    return abs(a - b) / abs(b)


# --------------------------------------------------------------------------- #
def test_steel_zeff_from_gdml():
    # 74/18/8 atom percent steel
    assert steel_zeff() == pytest.approx(25.8, rel=1e-13)


def test_steel_coefficients():
    c = ob.msc_material_coeffs(steel_zeff())
    for key, expected in STEEL_GOLDEN.items():
        assert rel(c[key], expected) < 1e-12, key


def test_planted_coefficient_mistakes_fail():
    # A sign slip in the theta0 correction's Z term
    ob.msc_set_mutation(1)
    c = ob.msc_material_coeffs(steel_zeff())
    assert rel(c["theta1"], STEEL_GOLDEN["theta1"]) > 1e-3
    ob.msc_set_mutation(0)
    # Zeff from mass fractions instead of atom fractions
    zmass = sum(z * w for z, w in MASS_FRACTIONS.items())
    c = ob.msc_material_coeffs(zmass)
    assert rel(c["stepmin0"], STEEL_GOLDEN["stepmin0"]) > 1e-6


@pytest.mark.parametrize("z,expected", [(1.0, POSITRON_H), (74.0, POSITRON_W)])
def test_positron_correction(z, expected):
    got = [ob.urban_positron_correction(z, y) for y in POSITRON_Y]
    for g, e in zip(got, expected):
        assert rel(g, e) < 1e-12


def test_planted_positron_mistake_fails():
    ob.msc_set_mutation(2)
    got = [ob.urban_positron_correction(74.0, y) for y in POSITRON_Y]
    # The pin (1e-12) rejects it
    assert max(rel(g, e) for g, e in zip(got, POSITRON_W)) > 1e-9


# --------------------------------------------------------------------------- #
# Step conversion on the generated steel tables
# --------------------------------------------------------------------------- #
def steel_physics_spec():
    spec = wl.testem3_physics_spec()
    frac, molar = steel_number_fractions()
    base = len(spec["elements"])
    shells = {26: 7.112e-3, 24: 5.989e-3, 28: 8.333e-3}
    for z in (26, 24, 28):
        spec["elements"].append({"z": z, "atomic_mass": molar[z],
                                 "mean_exc_energy": {26: 286e-6, 24: 257e-6, 28: 311e-6}[z],
                                 "shells": [[shells[z], 0.8], [shells[z] / 8, 0.2]],
                                 "pe_reference": [0.1, 100.0]})
    steel = {"label": "G4_STAINLESS-STEEL", "density": 8.0,
             "mean_exc_energy": 282.97693634122e-6,
             "elements": [{"element": base + i, "fraction": frac[z]}
                          for i, z in enumerate((26, 24, 28))]}
    spec["materials"] = [steel] + spec["materials"][1:]
    spec["cutoffs"] = {k: [v[0]] + v[1:] for k, v in spec["cutoffs"].items()}
    return spec


@pytest.fixture(scope="module")
def steel():
    phys = amd.generate_physics(steel_physics_spec())
    return ob.Problem(wl.testem3_problem(physics=phys))


ENERGIES = [99.999, 51.0231, 10.0564, 5.05808, 1.01162, 0.501328, 0.102364, 0.0465336,
            0.00708839, 1e-5]
LIMIT_MIN_FIX = 1e-9  # [cm]


def log_points(lo, hi, n):
    """LogInterp({0, lo}, {n, hi}) at 0..n (UrbanMsc.test.cc:199-200)."""
    return [math.exp(math.log(lo) + i / n * (math.log(hi) - math.log(lo))) for i in range(n + 1)]


def check_conversion(steel, par, energy):
    """Assertions of the reference's test_one lambda; returns the worst
    relative round-trip error seen in each regime."""
    (rng, lam, _, _, _), _ = ob.msc_step_conversion(steel, 0, par, energy, 1e-3, [])
    worst = {}
    pts = log_points(0.9 * LIMIT_MIN_FIX, rng, 8)
    pts[-1] = rng
    for pstep in pts:
        (_, _, gp, alpha, _), _ = ob.msc_step_conversion(steel, 0, par, energy, pstep, [])
        assert 0 < gp <= pstep
        gsteps = [min(g, pstep) for g in log_points(0.9 * LIMIT_MIN_FIX, gp, 8)] + [gp]
        _, true_steps = ob.msc_step_conversion(steel, 0, par, energy, pstep, gsteps)
        for g, t in zip(gsteps, true_steps):
            assert t <= pstep
            assert t >= g
        tol = 1e-3 if 1 - alpha * pstep < 1e-8 else 1e-10
        err = rel(true_steps[-1], pstep)
        assert err <= tol, (energy, pstep, alpha)
        regime = "tiny" if pstep < 1e-7 else ("small" if alpha == 0 else "linear")
        worst[regime] = max(worst.get(regime, 0), err)
    return worst


@pytest.mark.parametrize("par", [0, 1])
def test_step_conversion_invariants(steel, par):
    regimes = set()
    for energy in ENERGIES:
        regimes |= set(check_conversion(steel, par, energy))
    # Every regime of MscStepToGeo is exercised
    assert regimes == {"tiny", "small", "linear"}


def test_small_step_closed_form(steel):
    """Constant-MFP regime: z = lambda (1 - exp(-t/lambda))."""
    energy = 10.0564
    (rng, lam, _, _, _), _ = ob.msc_step_conversion(steel, 0, 0, energy, 1e-3, [])
    for t in (2e-7, 1e-5, 1e-3 * rng, 0.04 * rng):
        (_, _, gp, alpha, _), _ = ob.msc_step_conversion(steel, 0, 0, energy, t, [])
        assert alpha == 0
        assert rel(gp, -lam * math.expm1(-t / lam)) < 1e-14


def test_planted_closed_form_mistake_fails(steel):
    ob.msc_set_mutation(7)
    with pytest.raises(AssertionError):
        test_small_step_closed_form(steel)


@pytest.mark.parametrize("mutation", [3, 4])
def test_planted_conversion_mistakes_fail(steel, mutation):
    ob.msc_set_mutation(mutation)
    with pytest.raises(AssertionError):
        for par in (0, 1):
            for energy in ENERGIES:
                check_conversion(steel, par, energy)


# --------------------------------------------------------------------------- #
# Angular sampling
# --------------------------------------------------------------------------- #
def moments_ok(samples, m1, m2, nsigma=5):
    n = len(samples)
    mean = samples.mean()
    sq = (samples ** 2).mean()
    s1 = samples.std() / math.sqrt(n)
    s2 = (samples ** 2).std() / math.sqrt(n)
    return abs(mean - m1) < nsigma * s1 and abs(sq - m2) < nsigma * s2


@pytest.mark.parametrize("tau", [0.05, 0.3, 1.0, 3.0])
def test_simple_scattering_moments(tau):
    xm = math.exp(-tau)
    x2 = (1 + 2 * math.exp(-2.5 * tau)) / 3
    s = ob.msc_simple_scattering(xm, x2, 2_000_000)
    assert np.all(np.abs(s) <= 1)
    assert moments_ok(s, xm, x2)


def test_planted_simple_scattering_mistake_fails():
    ob.msc_set_mutation(5)
    tau = 0.3
    xm = math.exp(-tau)
    x2 = (1 + 2 * math.exp(-2.5 * tau)) / 3
    assert not moments_ok(ob.msc_simple_scattering(xm, x2, 2_000_000), xm, x2)


#: Tail-model inputs (tau, theta0, true path [cm], limit_min [cm], lambda
#: [cm], radiation length [cm]) in steel and lead-like regimes
TAIL_CASES = [
    (0.01, 0.05, 0.01, 1e-4, 1.0, 1.76),
    (0.002, 0.02, 1e-3, 1e-4, 0.5, 0.56),
    (0.05, 0.12, 0.05, 1e-3, 1.0, 1.76),
]


def g1_mean(x, ea):
    """Mean of mu = 1 + x ln u, u uniform on [ea, 1]."""
    # int_ea^1 ln u du = -1 - ea ln(ea) + ea
    return 1 + x * (-1 - ea * math.log(ea) + ea) / (1 - ea)


def g2_mean(x, xsi, c, d):
    """Mean of mu = x (c - xsi - c (v + d)^(-1/(c-1))) + 1, v uniform on
    [0, 1 - d], by Gauss-Legendre quadrature of the inverse CDF."""
    nodes, weights = np.polynomial.legendre.leggauss(200)
    w = 0.5 * (1 + nodes)  # (v + d) in [d, 1] through t: v + d = d^(1 - w)
    vd = d ** (1 - w)
    jac = vd * (-math.log(d))  # d(v+d)/dw
    mu = x * (c - xsi - c * vd ** (-1 / (c - 1))) + 1
    return float(0.5 * np.sum(weights * mu * jac)) / (1 - d)


@pytest.mark.parametrize("case", TAIL_CASES)
def test_tail_model(case):
    zeff = steel_zeff()
    parts = ob.msc_tail_parts(zeff, *case)
    xm = math.exp(-case[0])
    # The tail branch is the one taken for these inputs
    assert parts["xmean_1"] > 0.999 * xm
    m2 = g2_mean(parts["x"], parts["xsi"], parts["c"], parts["d"])
    assert rel(parts["xmean_2"], m2) < 1e-10
    m1 = g1_mean(parts["x"], parts["ea"])
    q, p = parts["qprob"], parts["prob"]
    expected = q * (p * m1 + (1 - p) * m2)
    s = ob.msc_sample_cos_theta(zeff, *case, 2_000_000)
    assert np.all(np.abs(s) <= 1)
    sigma = s.std() / math.sqrt(len(s))
    assert abs(s.mean() - expected) < 5 * sigma


def test_planted_tail_mistake_fails():
    ob.msc_set_mutation(6)
    zeff = steel_zeff()
    case = TAIL_CASES[0]
    parts = ob.msc_tail_parts(zeff, *case)
    m2 = g2_mean(parts["x"], parts["xsi"], parts["c"], parts["d"])
    m1 = g1_mean(parts["x"], parts["ea"])
    q, p = parts["qprob"], parts["prob"]
    expected = q * (p * m1 + (1 - p) * m2)
    s = ob.msc_sample_cos_theta(zeff, *case, 2_000_000)
    assert abs(s.mean() - expected) > 5 * s.std() / math.sqrt(len(s))
