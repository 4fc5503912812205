//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file em/LPM.hh
//! Landau-Pomeranchuk-Migdal suppression functions xi(s), G(s), phi(s).
//!
//! Restates em/xs/LPMCalculator.hh (Stanev et al., Phys. Rev. D 25 (1982)
//! 1291, Eqs. 14-21; Geant4 Physics Reference Manual section 10.2.2): the
//! suppression variable s' from the material's LPM energy (radiation
//! length times the LPM constant), Stanev's xi(s'), optional dielectric
//! (Migdal) suppression, and the piecewise approximations of phi and G
//! evaluated by Horner's rule with fma like the reference's PolyEvaluator.
//! The relativistic bremsstrahlung DCS uses them with dielectric
//! suppression (em/xs/RBDiffXsCalculator.hh:186-205), Bethe-Heitler
//! without (em/interactor/BetheHeitlerInteractor.hh:237-273).
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Precision.hh"
#include "phys/PhysicsData.hh"
#include "phys/PhysicsImpl.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
namespace lpm
{
//---------------------------------------------------------------------------//
/*!
 * alpha (m_e c^2)^2 / (2 h c) [MeV/cm] (PhysicsConstants.hh:46-55).
 *
 * Evaluated in CGS from the electron mass in grams, as the reference does:
 * that rest energy differs from 0.51099895 MeV by 7.5e-12 relative, which
 * the LPM DCS resolves at the 1e-12 level.
 */
CELER_FORCEINLINE_FUNCTION double lpm_constant()
{
    constexpr double c_light = 2.99792458e10;  // [cm/s]
    constexpr double h_planck = 6.62607015e-27;  // [erg s]
    constexpr double electron_mass = 9.1093837015e-28;  // [g]
    constexpr double mev = 1.602176634e-6;  // [erg]
    constexpr double alpha = 7.2973525693e-3;
    double const emc2 = electron_mass * (c_light * c_light);
    return alpha * (emc2 * emc2) / (2 * h_planck * c_light) / mev;
}

//! 4 pi r_e lambdabar_e^2 [cm^3]
CELER_FORCEINLINE_FUNCTION double migdal_constant()
{
    constexpr double lambdabar_electron = 3.8615926796e-11;  // [cm]
    return 4 * phys::pi * phys::r_electron * (lambdabar_electron * lambdabar_electron);
}

struct Functions
{
    double xi;
    double g;
    double phi;
};

CELER_FORCEINLINE_FUNCTION double horner2(double c0, double c1, double c2, double x)
{
    return std::fma(x, std::fma(x, c2, c1), c0);
}

CELER_FUNCTION double calc_phi(double s)
{
    if (s < 0.01)
        return s * std::fma(s, -6 * phys::pi, 6.0);
    if (s < 1.55)
    {
        double const a = horner2(0.623, 0.796, 0.658, s);
        double const b = horner2(-6.0, -6 * (3 - phys::pi), 1 / a, s);
        return 1 - mth::exp(s * b);
    }
    double const s2 = s * s;
    return 1 - 0.01190476 / (s2 * s2);
}

CELER_FUNCTION double calc_g(double s, double phi)
{
    if (s < 0.01)
        return std::fma(s, 12.0, -2 * phi);
    if (s < 0.415827)
    {
        double const a = std::fma(
            s, std::fma(s, std::fma(s, std::fma(s, 7.5, -0.05), 4.97), 3.936), 1.0);
        double const b = std::fma(s, -8 / a, -4.0);
        double const psi = 1 - mth::exp(s * b);
        return 3 * psi - 2 * phi;
    }
    if (s < 1.9156)
    {
        double const p = std::fma(
            s,
            std::fma(s, std::fma(s, std::fma(s, -0.120772, 0.672827), -1.798138), 3.755030),
            -0.160723);
        return mth::tanh(p);
    }
    double const s2 = s * s;
    return 1 - 0.0230655 / (s2 * s2);
}

/*!
 * Suppression functions at epsilon = (lepton energy) / (photon energy) for
 * a photon of gamma_energy in a material with the given LPM energy.
 */
CELER_FUNCTION Functions calc_functions(double cbrt_z,
                                        double electron_density,
                                        double lpm_energy,
                                        bool dielectric_suppression,
                                        double gamma_energy,
                                        double epsilon)
{
    double const s_prime = mth::sqrt(
        lpm_energy / (8 * epsilon * gamma_energy * std::fabs(epsilon - 1)));
    double const r = cbrt_z / 184.15;
    double const s1 = r * r;
    constexpr double sqrt_two = 1.41421356237309504880;
    double xi = 2;
    if (s_prime > 1)
    {
        xi = 1;
    }
    else if (s_prime > sqrt_two * s1)
    {
        double const log_s1 = mth::log(sqrt_two * s1);
        double const h = mth::log(s_prime) / log_s1;
        xi = 1 + h - 0.08 * (1 - h) * h * (2 - h) / log_s1;
    }
    double s = s_prime / mth::sqrt(xi);
    if (dielectric_suppression)
    {
        double const ek = epsilon * gamma_energy;
        double const k_p_sq = electron_density * migdal_constant() * (ek * ek);
        s *= (1 + k_p_sq / (gamma_energy * gamma_energy));
        xi = 2;
        if (s > 1)
            xi = 1;
        else if (s > s1)
            xi = 1 + mth::log(s) / mth::log(s1);
    }
    Functions f;
    f.phi = calc_phi(s);
    if (xi * f.phi > 1 || s > 0.57)
        xi = 1 / f.phi;
    f.xi = xi;
    f.g = calc_g(s, f.phi);
    return f;
}

//---------------------------------------------------------------------------//
}  // namespace lpm
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
