//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file em/Interactors.hh
//! Model dispatch, element selection and the post-interaction update.
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Precision.hh"
#include "geo/OrangeLiteTracker.hh"
#include "Brems.hh"
#include "InteractorCommon.hh"
#include "KleinNishina.hh"
#include "LivermorePE.hh"
#include "StandardEm.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
//---------------------------------------------------------------------------//
//! Everything an interactor may need besides the RNG
struct InteractionContext
{
    PhysicsParamsRef const& params;
    CoreStateRef const& state;
    size_type slot;
    std::uint32_t material;
    std::uint32_t element;  //!< component index in the material
    std::uint32_t particle;
};

/*!
 * Sample an element component for a multi-element material.
 *
 * Single-element materials need no draw (PhysicsStepUtils.hh:290-300).
 */
template<class Engine>
CELER_FUNCTION std::uint32_t sample_element(PhysicsParamsRef const& p,
                                            ProcessRecord const& proc,
                                            std::uint32_t model_index,
                                            std::uint32_t mat,
                                            double energy,
                                            Engine& rng)
{
    if (p.num_materials == 0 || p.materials == nullptr)
        return 0;
    MaterialRecord const& m = p.materials[mat];
    std::uint32_t const table = proc.models[model_index].elem_cdf;
    if (m.num_elements <= 1 || table == invalid_id)
        return 0;
    std::uint32_t const first = p.elem_cdf_index[table + mat];
    if (first == invalid_id)
        return 0;
    // TabulatedElementSelector::operator() (TabulatedElementSelector.hh:
    // 75-95): the first element whose normalised cumulative XS exceeds u
    double const u = draw_canonical(rng);
    std::uint32_t i = 0;
    for (; i + 1 < m.num_elements; ++i)
    {
        if (phys::calc_grid(p.grids[first + i], p.grid_reals, energy) > u)
            break;
    }
    return i;
}

//! Dispatch to the model's interactor
template<class Engine>
CELER_FUNCTION Interaction sample_interaction(ModelClass model,
                                              InteractionContext const& ctx,
                                              double energy,
                                              double const* dir,
                                              Engine& rng)
{
    PhysicsParamsRef const& p = ctx.params;
    switch (model)
    {
        case ModelClass::klein_nishina:
            return sample_klein_nishina(p, ctx.state, energy, dir, rng);
        case ModelClass::bethe_heitler: {
            MaterialRecord const& m = p.materials[ctx.material];
            ElementRecord const& el
                = p.elements[p.components[m.elem_begin + ctx.element].element];
            return sample_bethe_heitler(p, ctx.state, m, el, energy, dir, rng);
        }
        case ModelClass::moller_bhabha:
            return sample_moller_bhabha(p, ctx.state, ctx.particle, ctx.material,
                                        energy, dir, rng);
        case ModelClass::eplusgg:
            return sample_eplusgg(p, ctx.state, energy, dir, rng);
        case ModelClass::livermore_pe: {
            // No element table: multi-element materials select here
            MaterialRecord const& m = p.materials[ctx.material];
            std::uint32_t comp = ctx.element;
            if (m.num_elements > 1)
                comp = select_pe_element(p, ctx.material, energy, rng);
            return sample_livermore_pe(p, ctx.state,
                                       p.components[m.elem_begin + comp].element,
                                       energy, dir, rng);
        }
        case ModelClass::seltzer_berger:
            return sample_brems(p, ctx.state, ctx.particle, ctx.material,
                                ctx.element, energy, dir, 0, rng);
        case ModelClass::relativistic_brem:
            return sample_brems(p, ctx.state, ctx.particle, ctx.material,
                                ctx.element, energy, dir, 1, rng);
        case ModelClass::combined_brem:
            return sample_brems(p, ctx.state, ctx.particle, ctx.material,
                                ctx.element, energy, dir, 2, rng);
        default:
            break;
    }
    Interaction r;
    r.action = Interaction::unchanged;
    return r;
}

/*!
 * Apply a sampled interaction to the track (InteractionApplier.hh:106-171).
 */
CELER_FUNCTION void apply_interaction(CoreParamsRef const& params,
                                      CoreStateRef const& state,
                                      size_type slot,
                                      std::uint32_t mat,
                                      Interaction const& r)
{
    PhysicsParamsRef const& pp = params.phys;
    if (CELER_UNLIKELY(r.action == Interaction::failed))
    {
        if (0 < state.sim.step_length[slot])
        {
            state.sim.step_length[slot] = 0;
            state.sim.post_step_action[slot] = pp.scalars.failure_action;
        }
        return;
    }
    if (r.action == Interaction::unchanged)
        return;
#if CELER_PREC_IS_MIXED
    // Reduced-precision sampling must not leave a non-physical energy
    if (CELER_UNLIKELY(!(r.energy >= 0 && r.energy < HUGE_VAL)))
    {
        state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::errored);
        state.sim.post_step_action[slot] = params.scalars.tracking_cut_action;
        return;
    }
#endif
    state.particle.energy[slot] = r.energy;
    if (r.action != Interaction::absorbed)
    {
        geo::GeoTrack g = geo::GeoTrack::load(state.geo, slot);
        geo::set_dir(params.geo, g, r.dir);
        g.store(state.geo, slot);
    }
    else
    {
        state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::killed);
    }
    double deposition = r.energy_deposition;
    if (pp.scalars.apply_post_interaction_cut)
    {
        for (std::uint32_t k = 0; k < r.num_secondaries; ++k)
        {
            Secondary& sec = state.secondaries.storage[r.sec_start + k];
            if (sec.particle_id != invalid_id
                && sec.energy < pp.cutoff_energy[sec.particle_id * pp.num_materials + mat])
            {
                deposition += sec.energy;
                if (sec.particle_id == pp.positron_id)
                    deposition += 2 * pp.mass[sec.particle_id];
                sec.particle_id = invalid_id;
                sec.energy = 0;
            }
        }
    }
    state.phys.energy_deposition[slot] += deposition;
    state.phys.sec_start[slot] = r.sec_start;
    state.phys.sec_count[slot] = r.num_secondaries;
}

//---------------------------------------------------------------------------//
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
