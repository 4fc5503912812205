//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file em/StandardEm.hh
//! Pair production, ionisation and positron annihilation interactors.
//!
//! Sampling methods follow the reference interactors, which implement the
//! Geant4 Physics Reference Manual algorithms:
//! - Bethe-Heitler pair production with the Tsai screening functions and
//!   Coulomb correction (em/interactor/BetheHeitlerInteractor.hh:165-321),
//!   with LPM suppression above 100 GeV when the physics enables it.
//! - Moller (e-e-) and Bhabha (e+e-) delta-ray production above the cut
//!   (em/interactor/MollerBhabhaInteractor.hh:119-157,
//!   em/distribution/{Moller,Bhabha}EnergyDistribution.hh,
//!   em/interactor/detail/IoniFinalStateHelper.hh).
//! - Two-photon positron annihilation, Heitler cross section
//!   (em/interactor/EPlusGGInteractor.hh:99-170,
//!   em/xs/EPlusGGMacroXsCalculator.hh).
//! The order of uniform draws is the same as the reference's, so a shared
//! stream reproduces its outputs.
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Precision.hh"
#include "phys/PhysicsImpl.hh"
#include "random/XorwowEngine.hh"
#include "InteractorCommon.hh"
#include "LPM.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
//---------------------------------------------------------------------------//

//! Polar cosine of a pair product (TsaiUrbanDistribution)
template<class Engine>
CELER_FUNCTION double sample_tsai_urban(double energy, double mass, Engine& rng)
{
    double const umax = 2 * (1 + energy / mass);
    double u;
    unsigned int iter = 0;
    do
    {
        double const a = draw_canonical(rng);
        double const b = draw_canonical(rng);
        double const uu = -mth::log(a * b);
        u = uu * (draw_canonical(rng) < 0.25 ? 1.6 : 1.6 / 3);
    } while (u > umax && within_budget(iter, rng));
    double const r = u / umax;
    return 1 - 2 * r * r;
}

//! Direction at polar cosine mu about dir with uniform azimuth
template<class Engine>
CELER_FUNCTION void
sample_exiting_direction(double mu, double const* dir, Engine& rng, double* out)
{
    double const phi = std::fma(2 * phys::pi, draw_canonical(rng), 0.0);
    double local[3];
    from_spherical(mu, phi, local);
    rotate(local, dir, out);
}

//! Momentum-conserving exiting direction (calc_exiting_direction)
CELER_FUNCTION void momentum_direction(double p_in,
                                       double const* d_in,
                                       double p_out,
                                       double const* d_out,
                                       double* out)
{
    for (int i = 0; i < 3; ++i)
        out[i] = d_in[i] * p_in - d_out[i] * p_out;
    make_unit(out);
}

//---------------------------------------------------------------------------//
// BETHE-HEITLER
//---------------------------------------------------------------------------//
CELER_FORCEINLINE_FUNCTION double bh_screening_f1(double d)
{
    return d > 1.4 ? 42.038 - 8.29 * mth::log(d + 0.958)
                   : 42.184 - d * (7.444 - 1.623 * d);
}
CELER_FORCEINLINE_FUNCTION double bh_screening_f2(double d)
{
    return d > 1.4 ? 42.038 - 8.29 * mth::log(d + 0.958)
                   : 41.326 - d * (5.848 - 0.902 * d);
}

//! Screening functions Phi_1 and Phi_2 (BetheHeitlerInteractor.hh:338-356)
CELER_FORCEINLINE_FUNCTION void bh_screening_phi(double d, double* phi1, double* phi2)
{
    if (d > 1.4)
    {
        *phi1 = 21.0190 - 4.145 * mth::log(d + 0.958);
        *phi2 = *phi1;
    }
    else
    {
        *phi1 = 20.806 - d * (3.190 - 0.5710 * d);
        *phi2 = 20.234 - d * (2.126 - 0.0903 * d);
    }
}

//! Photon energy above which Bethe-Heitler applies LPM suppression [MeV]
inline constexpr double bh_lpm_threshold = 1e5;

template<class Engine>
CELER_FUNCTION Interaction sample_bethe_heitler(PhysicsParamsRef const& p,
                                                CoreStateRef const& s,
                                                MaterialRecord const& mat,
                                                ElementRecord const& el,
                                                double energy,
                                                double const* dir,
                                                Engine& rng)
{
    Interaction r;
    std::uint32_t const start = allocate_secondaries(s, 2);
    if (start == invalid_id)
    {
        r.action = Interaction::failed;
        return r;
    }
    double const me = p.model.electron_mass;
    double const eps0 = me / energy;
    constexpr double half = 0.5;
    double eps;
    if (energy < 2)
    {
        eps = std::fma(half - eps0, draw_canonical(rng), eps0);
    }
    else
    {
        double const delta_min = 4 * 136 / el.cbrt_z * eps0;
        double f_z = 8.0 / 3.0 * el.log_z;
        if (energy > 50)
            f_z += 8 * el.coulomb_correction;
        double const delta_max = mth::exp((42.038 - f_z) / 8.29) - 0.958;
        double const eps1 = half - half * mth::sqrt(1 - delta_min / delta_max);
        double const eps_min = eps0 > eps1 ? eps0 : eps1;
        double const f10 = bh_screening_f1(delta_min) - f_z;
        double const f20 = bh_screening_f2(delta_min) - f_z;
        double const st = (half - eps_min) * (half - eps_min) * f10;
        double const sf = 1.5 * f20;
        double const p_f1 = st / (st + sf);
        // LPM without dielectric suppression (BetheHeitlerData.hh:51-54)
        bool const use_lpm = p.scalars.enable_lpm && energy > bh_lpm_threshold;
        double const lpm_energy = mat.rad_length * lpm::lpm_constant();
        double g;
        unsigned int iter = 0;
        do
        {
            if (draw_canonical(rng) < p_f1)
            {
                eps = half - (half - eps_min) * mth::cbrt(draw_canonical(rng));
                double const delta = 136 / el.cbrt_z * eps0 / (eps * (1 - eps));
                if (use_lpm)
                {
                    double phi1, phi2;
                    bh_screening_phi(delta, &phi1, &phi2);
                    lpm::Functions const f = lpm::calc_functions(
                        el.cbrt_z, mat.electron_density, lpm_energy, false, energy, eps);
                    g = f.xi * ((2 * f.phi + f.g) * phi1 - f.g * phi2 - f.phi * f_z) / f10;
                }
                else
                {
                    g = (bh_screening_f1(delta) - f_z) / f10;
                }
            }
            else
            {
                eps = eps_min + (half - eps_min) * draw_canonical(rng);
                double const delta = 136 / el.cbrt_z * eps0 / (eps * (1 - eps));
                if (use_lpm)
                {
                    double phi1, phi2;
                    bh_screening_phi(delta, &phi1, &phi2);
                    lpm::Functions const f = lpm::calc_functions(
                        el.cbrt_z, mat.electron_density, lpm_energy, false, energy, eps);
                    g = f.xi
                        * ((f.phi + half * f.g) * phi1 + half * f.g * phi2
                           - half * (f.g + f.phi) * f_z)
                        / f20;
                }
                else
                {
                    g = (bh_screening_f2(delta) - f_z) / f20;
                }
            }
        } while (g < draw_canonical(rng) && within_budget(iter, rng));
    }
    r.action = Interaction::absorbed;
    r.energy = 0;
    r.sec_start = start;
    r.num_secondaries = 2;
    Secondary* sec = s.secondaries.storage + start;
    double e0 = (1 - eps) * energy - me;
    double e1 = eps * energy - me;
    if (draw_canonical(rng) < half)
    {
        double const t = e0;
        e0 = e1;
        e1 = t;
    }
    sec[0].particle_id = p.electron_id;
    sec[1].particle_id = p.positron_id;
    sec[0].energy = e0;
    sec[1].energy = e1;
    double const phi = std::fma(2 * phys::pi, draw_canonical(rng), 0.0);
    double local[3];
    double cost = sample_tsai_urban(e0, me, rng);
    from_spherical(cost, phi, local);
    rotate(local, dir, sec[0].dir);
    cost = sample_tsai_urban(e1, me, rng);
    from_spherical(cost, phi + phys::pi, local);
    rotate(local, dir, sec[1].dir);
    return r;
}

//---------------------------------------------------------------------------//
// MOLLER-BHABHA
//---------------------------------------------------------------------------//
CELER_FORCEINLINE_FUNCTION double moller_g(double gamma, double eps)
{
    double const tg = (2 * gamma - 1) / (gamma * gamma);
    double const comp = 1 - eps;
    return 1 - tg * eps + eps * eps * (1 - tg + (1 - tg * comp) / (comp * comp));
}

CELER_FORCEINLINE_FUNCTION double bhabha_g(double gamma, double emin, double emax)
{
    double const y = 1 / (1 + gamma);
    double const y_sq = y * y;
    double const omy = 1 - 2 * y;
    double const b1 = 2 - y_sq;
    double const b2 = omy * (3 + y_sq);
    double const b4 = omy * omy * omy;
    double const b3 = omy * omy + b4;
    double const beta_sq = 1 - 1 / (gamma * gamma);
    double const em2 = emax * emax;
    return 1 + (em2 * em2 * b4 - emin * emin * emin * b3 + em2 * b2 - emin * b1) * beta_sq;
}

template<class Engine>
CELER_FUNCTION Interaction sample_moller_bhabha(PhysicsParamsRef const& p,
                                                CoreStateRef const& s,
                                                std::uint32_t particle,
                                                std::uint32_t mat,
                                                double energy,
                                                double const* dir,
                                                Engine& rng)
{
    Interaction r;
    std::uint32_t const start = allocate_secondaries(s, 1);
    if (start == invalid_id)
    {
        r.action = Interaction::failed;
        return r;
    }
    double const me = p.model.electron_mass;
    double const cut = p.cutoff_energy[p.electron_id * p.num_materials + mat];
    double const min_frac = cut / energy;
    double const gamma = 1 + energy / me;
    double eps;
    unsigned int iter = 0;
    if (particle == p.electron_id)
    {
        double const g_den = moller_g(gamma, 0.5);
        double const a = 1 / 0.5;
        double const delta = 1 / min_frac - a;
        do
        {
            eps = 1 / std::fma(delta, draw_canonical(rng), a);
        } while (g_den * draw_canonical(rng) > moller_g(gamma, eps)
                 && within_budget(iter, rng));
    }
    else
    {
        double const g_den = bhabha_g(gamma, min_frac, 1.0);
        double const a = 1.0;
        double const delta = 1 / min_frac - a;
        do
        {
            eps = 1 / std::fma(delta, draw_canonical(rng), a);
        } while (g_den * draw_canonical(rng) > bhabha_g(gamma, eps, eps)
                 && within_budget(iter, rng));
    }
    double const e_sec = energy * eps;
    double const inc_momentum = mth::sqrt(energy * energy + 2 * me * energy);
    double const momentum = mth::sqrt(e_sec * (e_sec + 2 * me));
    double const costheta = e_sec * (energy + me + me) / (momentum * inc_momentum);
    Secondary& sec = s.secondaries.storage[start];
    sec.energy = e_sec;
    sample_exiting_direction(costheta, dir, rng, sec.dir);
    sec.particle_id = p.electron_id;
    r.energy = energy - e_sec;
    momentum_direction(inc_momentum, dir, momentum, sec.dir, r.dir);
    r.sec_start = start;
    r.num_secondaries = 1;
    return r;
}

//---------------------------------------------------------------------------//
// POSITRON ANNIHILATION
//---------------------------------------------------------------------------//
template<class Engine>
CELER_FUNCTION Interaction sample_eplusgg(PhysicsParamsRef const& p,
                                          CoreStateRef const& s,
                                          double energy,
                                          double const* dir,
                                          Engine& rng)
{
    Interaction r;
    std::uint32_t const start = allocate_secondaries(s, 2);
    if (start == invalid_id)
    {
        r.action = Interaction::failed;
        return r;
    }
    double const me = p.model.electron_mass;
    r.action = Interaction::absorbed;
    r.energy = 0;
    r.sec_start = start;
    r.num_secondaries = 2;
    Secondary* sec = s.secondaries.storage + start;
    sec[0].particle_id = sec[1].particle_id = p.gamma_id;
    if (energy == 0)
    {
        sec[0].energy = sec[1].energy = me;
        double const costheta = std::fma(2.0, draw_canonical(rng), -1.0);
        double const phi = std::fma(2 * phys::pi, draw_canonical(rng), 0.0);
        from_spherical(costheta, phi, sec[0].dir);
        for (int i = 0; i < 3; ++i)
            sec[1].dir[i] = -sec[0].dir[i];
        return r;
    }
    constexpr double half = 0.5;
    double const tau = energy / me;
    double const tau2 = tau + 2;
    double const sqgrate = mth::sqrt(tau / tau2) * half;
    double const lo = half - sqgrate;
    double const logratio = mth::log((1 / lo) * (half + sqgrate));
    double epsil;
    unsigned int iter = 0;
    do
    {
        epsil = lo * mth::exp(logratio * draw_canonical(rng));
    } while (draw_canonical(rng)
                 < epsil - (2 * (tau + 1) * epsil - 1) / (epsil * tau2 * tau2)
             && within_budget(iter, rng));
    double const cost = (epsil * tau2 - 1) / (epsil * mth::sqrt(tau * tau2));
    double const total = energy + 2 * me;
    double const eg = epsil * total;
    double const eplus_moment = mth::sqrt(energy * total);
    sec[0].energy = eg;
    sample_exiting_direction(cost, dir, rng, sec[0].dir);
    sec[1].energy = total - eg;
    momentum_direction(eplus_moment, dir, energy, dir, sec[1].dir);
    return r;
}

//---------------------------------------------------------------------------//
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
