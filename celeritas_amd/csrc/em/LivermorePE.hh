//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file em/LivermorePE.hh
//! Livermore photoelectric interactor (atomic relaxation off).
//!
//! Follows em/interactor/LivermorePEInteractor.hh:145-358 and
//! em/executor/LivermorePEExecutor.hh: for a material with several elements
//! and no preselected component, the element is drawn in proportion to its
//! micro cross section (mat/ElementSelector.hh); the subshell is sampled
//! from the cumulative subshell cross sections; the photoelectron takes
//! E - binding and a Sauter-Gavrila direction; the binding energy is
//! deposited locally.
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Precision.hh"
#include "phys/LivermoreXs.hh"
#include "StandardEm.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
//---------------------------------------------------------------------------//
//! Element component by micro cross section (ElementSelector::operator())
template<class Engine>
CELER_FUNCTION std::uint32_t
select_pe_element(PhysicsParamsRef const& p, std::uint32_t mat, double energy, Engine& rng)
{
    MaterialRecord const& m = p.materials[mat];
    ElementComponent const* comps = p.components + m.elem_begin;
    double material_xs = 0;
    for (std::uint32_t i = 0; i < m.num_elements; ++i)
        material_xs += comps[i].fraction
                       * phys::livermore_pe_micro_xs(p, comps[i].element, energy);
    double accum = -material_xs * draw_canonical(rng);
    std::uint32_t i = 0;
    std::uint32_t const imax = m.num_elements - 1;
    for (; i != imax; ++i)
    {
        accum += comps[i].fraction
                 * phys::livermore_pe_micro_xs(p, comps[i].element, energy);
        if (accum > 0)
            break;
    }
    return i;
}

//! Sauter-Gavrila photoelectron direction (sample_direction)
template<class Engine>
CELER_FUNCTION void
sample_sauter_gavrila(double inv_electron_mass, double inc_energy,
                      double const* inc_dir, Engine& rng, double* out)
{
    if (inc_energy > 100.)
    {
        out[0] = inc_dir[0];
        out[1] = inc_dir[1];
        out[2] = inc_dir[2];
        return;
    }
    double const energy_per_mecsq = (inc_energy > 1e-6 ? inc_energy : 1e-6)
                                    * inv_electron_mass;
    double const gamma = energy_per_mecsq + 1;
    double const beta = mth::sqrt(energy_per_mecsq * (gamma + 1)) / gamma;
    double const a = (1 - beta) / beta;
    double const b = 0.5 * beta * gamma * energy_per_mecsq * (gamma - 2);
    double const g_max = 2 * (1 / a + b);
    double g, nu;
    unsigned int iter = 0;
    do
    {
        double const u = draw_canonical(rng);
        nu = 2 * a * (2 * u + (a + 2) * mth::sqrt(u)) / ((a + 2) * (a + 2) - 4 * u);
        g = (2 - nu) * (1 / (a + nu) + b);
    } while (g < g_max * draw_canonical(rng) && within_budget(iter, rng));
    sample_exiting_direction(1 - nu, inc_dir, rng, out);
}

//! Photoelectric absorption of a photon on one element
template<class Engine>
CELER_FUNCTION Interaction sample_livermore_pe(PhysicsParamsRef const& p,
                                               CoreStateRef const& s,
                                               std::uint32_t element,
                                               double inc_energy,
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
    LivermoreElement const& el = p.pe_elements[element];
    LivermoreShell const* shells = p.pe_shells + el.shell_begin;
    std::uint32_t const num_shells = el.num_shells;
    double const inv_energy = 1 / inc_energy;

    // Subshell (sample_subshell)
    std::uint32_t shell_id = 0;
    double const cutoff = draw_canonical(rng)
                          * phys::livermore_pe_micro_xs(p, element, inc_energy);
    if (inc_energy < el.thresh_lo)
    {
        double xs = 0;
        double const inv_cube = inv_energy * inv_energy * inv_energy;
        for (; shell_id < num_shells; ++shell_id)
        {
            LivermoreShell const& sh = shells[shell_id];
            if (inc_energy < sh.binding_energy)
                continue;
            xs += inv_cube * phys::generic_calc(sh.xs, p.reals, inc_energy);
            if (xs > cutoff)
                break;
        }
    }
    else
    {
        int const pidx = inc_energy < el.thresh_hi ? 0 : 1;
        for (; shell_id < num_shells - 1; ++shell_id)
        {
            double const xs
                = inv_energy * phys::poly5(shells[shell_id].param[pidx], inv_energy);
            if (xs > cutoff)
                break;
        }
    }
    r.action = Interaction::absorbed;
    r.energy = 0;
    r.sec_start = start;
    r.num_secondaries = 1;
    if (CELER_UNLIKELY(shell_id == num_shells))
    {
        // No shell accessible: deposit everything, no photoelectron
        r.energy_deposition = inc_energy;
        r.num_secondaries = 0;
        return r;
    }
    double const binding = shells[shell_id].binding_energy;
    Secondary& e = s.secondaries.storage[start];
    e.particle_id = p.electron_id;
    e.energy = inc_energy - binding;
    sample_sauter_gavrila(p.model.inv_electron_mass, inc_energy, dir, rng, e.dir);
    r.energy_deposition = binding;
    return r;
}

//---------------------------------------------------------------------------//
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
