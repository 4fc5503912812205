//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file em/KleinNishina.hh
//! Compton scattering with the Klein-Nishina cross section.
//!
//! Sampling follows the reference interactor
//! (celeritas/em/interactor/KleinNishinaInteractor.hh:104-190), which is the
//! Geant4 Physics Reference Manual method: epsilon = E'/E drawn from the
//! mixture 1/eps on [eps0, 1] (weight -ln eps0) and eps on [eps0, 1]
//! (weight (1 - eps0^2)/2), accepted with probability
//! 1 - eps sin^2(theta) / (1 + eps^2). The same sequence of uniform draws
//! is consumed, so a shared RNG stream gives identical outputs.
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Precision.hh"
#include "phys/PhysicsImpl.hh"
#include "random/XorwowEngine.hh"
#include "InteractorCommon.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
//---------------------------------------------------------------------------//
//! Electron kinetic energy below which the recoil is deposited locally
inline constexpr double kn_secondary_cutoff = 1e-4;  // [MeV]

template<class Engine>
CELER_FUNCTION Interaction sample_klein_nishina(PhysicsParamsRef const& p,
                                                CoreStateRef const& s,
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
    double const k = energy * p.model.inv_electron_mass;
    double const eps0 = 1 / (1 + 2 * k);
    double const a = -mth::log(eps0);
    double const b = 0.5 * (1 - eps0 * eps0);
    double const p_f1 = a / (a + b);
    double const logratio = mth::log(eps0);
    double const f2_lo = eps0 * eps0;
    double const f2_delta = 1 - f2_lo;
    double eps, omc;
    unsigned int iter = 0;
    while (true)
    {
        double eps_sq;
        if (draw_canonical(rng) < p_f1)
        {
            eps = mth::exp(logratio * draw_canonical(rng));
            eps_sq = eps * eps;
        }
        else
        {
            eps_sq = std::fma(f2_delta, draw_canonical(rng), f2_lo);
            eps = mth::sqrt(eps_sq);
        }
        omc = (1 - eps) / (eps * k);
        double const sin2 = omc * (2 - omc);
        double const reject = eps * sin2 / (1 + eps_sq);
        if (!(draw_canonical(rng) < reject) || !within_budget(iter, rng))
            break;
    }
    r.energy = eps * energy;
    double const phi = std::fma(2 * phys::pi, draw_canonical(rng), 0.0);
    double local[3];
    from_spherical(1 - omc, phi, local);
    rotate(local, dir, r.dir);
    r.sec_start = start;
    r.num_secondaries = 1;
    Secondary& sec = s.secondaries.storage[start];
    double const e_electron = energy - r.energy;
    if (e_electron < kn_secondary_cutoff)
    {
        r.energy_deposition = e_electron;
        return r;
    }
    sec.energy = e_electron;
    sec.particle_id = p.electron_id;
    double v[3];
    for (int i = 0; i < 3; ++i)
        v[i] = dir[i] * energy - r.dir[i] * r.energy;
    make_unit(v);
    for (int i = 0; i < 3; ++i)
        sec.dir[i] = v[i];
    return r;
}

//---------------------------------------------------------------------------//
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
