//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/StepExecutors.hh
//! Per-track bodies of the step actions (one call per track slot).
//!
//! The kernels in StepKernels.hip map a thread to a slot, filter by status
//! and action, count work and call these; the same functions compile for
//! the host, where the track-level loop can be driven slot by slot. They
//! restate the reference executors: PreStepExecutor
//! (phys/detail/PreStepExecutor.hh:46-115), the neutral along-step
//! (global/alongstep/AlongStep.hh:38-48 with PropagationApplier,
//! TimeUpdater, TrackUpdater), DiscreteSelectExecutor
//! (phys/detail/DiscreteSelectExecutor.hh:38-63), InteractionApplier
//! (phys/InteractionApplier.hh:106-171), BoundaryExecutor
//! (geo/detail/BoundaryExecutor.hh:42-84) and TrackingCutExecutor
//! (phys/detail/TrackingCutExecutor.hh:49-84).
//---------------------------------------------------------------------------//
#pragma once

#include <cmath>

#include "corecel/Precision.hh"
#include "em/Interactors.hh"
#include "geo/OrangeLiteTracker.hh"
#include "phys/PhysicsImpl.hh"
#include "track/TrackInitImpl.hh"
#include "CoreData.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
namespace exec
{
//---------------------------------------------------------------------------//
CELER_FORCEINLINE_FUNCTION TrackStatus status_of(CoreStateRef const& s, size_type i)
{
    return static_cast<TrackStatus>(s.sim.status[i]);
}

CELER_FORCEINLINE_FUNCTION std::uint32_t material_of(CoreParamsRef const& p,
                                                     CoreStateRef const& s,
                                                     size_type i)
{
    return p.volume_to_material[s.geo.volume[i]];
}

//---------------------------------------------------------------------------//
//! Pre-step of an inactive slot: no step, no actions
CELER_FORCEINLINE_FUNCTION void pre_step_inactive(CoreStateRef const& state, size_type slot)
{
    state.sim.step_length[slot] = HUGE_VAL;
    state.sim.post_step_action[slot] = invalid_id8;
    state.sim.along_step_action[slot] = invalid_id8;
}

/*!
 * Pre-step of an active track: reset the step state, sample the number of
 * interaction lengths if needed, compute the per-process and total cross
 * sections and the physics step limit, and choose the along-step action.
 */
CELER_FUNCTION void pre_step_track(CoreParamsRef const& params,
                                   CoreStateRef const& state,
                                   size_type slot,
                                   TrackStatus status)
{
    PhysicsStateRef const& ps = state.phys;
    ps.energy_deposition[slot] = 0;
    ps.sec_start[slot] = 0;
    ps.sec_count[slot] = 0;
    ps.element[slot] = invalid_id8;
    state.pre_volume[slot] = state.geo.volume[slot];
    if (status == TrackStatus::errored)
        return;
    state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::alive);

    PhysicsParamsRef const& pp = params.phys;
    if (pp.num_particles == 0)
        return;
    std::uint32_t const pid = state.particle.particle_id[slot];
    ParticlePhysics const& part = pp.particle[pid];
    double mfp = ps.interaction_mfp[slot];
    if (!(mfp > 0))
    {
        XorwowEngine rng = XorwowEngine::load(state.rng, slot);
        mfp = -mth::log(draw_canonical(rng));
        rng.store(state.rng, slot);
        ps.interaction_mfp[slot] = mfp;
    }
    std::uint32_t const mat = material_of(params, state, slot);
    double const energy = state.particle.energy[slot];
    double total = 0;
    // One log and grid bracket of the energy, shared by the processes
    phys::GridPoint gp(energy);
    for (std::uint32_t i = 0; i < part.num_processes; ++i)
    {
        ProcessRecord const& proc = part.process[i];
        double const xs = proc.integral ? phys::calc_max_xs(pp, proc, mat, gp)
                                        : phys::calc_xs(pp, proc, mat, gp);
        total += xs;
        ps.per_process_xs[i * state.size + slot] = xs;
    }
    ps.macro_xs[slot] = total;
    double step;
    std::uint8_t action = pp.scalars.discrete_action;
    if (energy == 0)
    {
        step = 0;
    }
    else
    {
        step = mfp / total;
        if (part.eloss_process >= 0)
        {
            double const range
                = phys::calc_range(pp.grids[part.range_grid + mat], pp.grid_reals, energy);
            ps.dedx_range[slot] = range;
            double const eloss_step = phys::range_to_step(pp.scalars, range);
            if (eloss_step <= step)
            {
                step = eloss_step;
                action = pp.scalars.range_action;
            }
            double const fixed = pp.scalars.fixed_step_limiter;
            if (fixed > 0 && fixed < step)
            {
                step = fixed;
                action = pp.scalars.fixed_step_action;
            }
        }
        else if (part.num_processes == 0)
        {
            action = invalid_id8;
        }
    }
    state.sim.step_length[slot] = step;
    state.sim.post_step_action[slot] = action;
    state.sim.along_step_action[slot] = (pp.charge[pid] == 0)
                                            ? params.scalars.along_step_neutral_action
                                            : params.scalars.along_step_user_action;
}

//---------------------------------------------------------------------------//
/*!
 * Neutral along-step: linear propagation to the physics step or the next
 * boundary, time and interaction-MFP updates.
 */
CELER_FUNCTION void along_step_neutral_track(CoreParamsRef const& params,
                                             CoreStateRef const& state,
                                             size_type slot,
                                             TrackStatus status)
{
    double step = state.sim.step_length[slot];
    std::uint8_t post = state.sim.post_step_action[slot];
    if (step != 0)
    {
        geo::GeoTrack g = geo::GeoTrack::load(state.geo, slot);
        geo::Propagation prop = geo::find_next_step(params.geo, g, step);
        if (prop.boundary)
            geo::move_to_boundary(g);
        else
            geo::move_internal(g, step);
        g.store(state.geo, slot);
        if (CELER_UNLIKELY(!std::isfinite(g.pos[0] + g.pos[1] + g.pos[2])))
        {
            // Not reachable from a finite FP64 state
            apply_errored(params, state, slot);
            return;
        }
        if (prop.boundary)
        {
            step = prop.distance;
            post = params.scalars.boundary_action;
        }
        else if (prop.distance < step)
        {
            step = prop.distance;
            post = params.scalars.propagation_limit_action;
        }
        state.sim.step_length[slot] = step;
        state.sim.post_step_action[slot] = post;
    }
    if (status == TrackStatus::errored)
        return;
    std::uint32_t const pid = state.particle.particle_id[slot];
    double const v = phys::speed(state.particle.energy[slot], params.phys.mass[pid]);
    if (v > 0)
        state.sim.time[slot] += step / v;
    if (post != params.phys.scalars.discrete_action)
    {
        state.phys.interaction_mfp[slot] -= step * state.phys.macro_xs[slot];
    }
    state.sim.num_steps[slot] += 1;
}

//---------------------------------------------------------------------------//
/*!
 * Discrete select: choose the process by its share of the total cross
 * section, apply the integral-approach rejection, choose the model by
 * energy and sample the target element.
 */
CELER_FUNCTION void discrete_select_track(CoreParamsRef const& params,
                                          CoreStateRef const& state,
                                          size_type slot)
{
    PhysicsParamsRef const& pp = params.phys;
    state.phys.interaction_mfp[slot] = 0;
    std::uint32_t const pid = state.particle.particle_id[slot];
    ParticlePhysics const& part = pp.particle[pid];
    std::uint32_t const mat = material_of(params, state, slot);
    double const energy = state.particle.energy[slot];
    XorwowEngine rng = XorwowEngine::load(state.rng, slot);
    double accum = -state.phys.macro_xs[slot] * draw_canonical(rng);
    std::uint32_t pidx = part.num_processes - 1;
    bool found = false;
    for (std::uint32_t i = 0; i + 1 < part.num_processes; ++i)
    {
        accum += state.phys.per_process_xs[i * state.size + slot];
        if (accum > 0)
        {
            pidx = i;
            found = true;
            break;
        }
    }
    if (CELER_UNLIKELY(!found))
    {
        // Falling through to the last process is only correct when it has a
        // cross section: a draw at the top of [0, 1) with rounding could
        // otherwise select a process that cannot occur (pair production
        // below threshold); take the last process that can
        while (pidx > 0 && !(state.phys.per_process_xs[pidx * state.size + slot] > 0))
            --pidx;
    }
    ProcessRecord const& proc = part.process[pidx];
    std::uint8_t action;
    bool rejected = false;
    if (proc.integral)
    {
        double const xs = phys::calc_xs(pp, proc, mat, energy);
        if (draw_canonical(rng) * state.phys.per_process_xs[pidx * state.size + slot] > xs)
        {
            rejected = true;
        }
    }
    if (rejected)
    {
        action = pp.scalars.integral_rejection_action;
    }
    else
    {
        std::uint32_t mi = 0;
        for (std::uint32_t k = 1; k < proc.num_models; ++k)
        {
            if (energy >= proc.models[k].energy_low)
                mi = k;
        }
        state.phys.element[slot]
            = static_cast<std::uint8_t>(sample_element(pp, proc, mi, mat, energy, rng));
        action = static_cast<std::uint8_t>(proc.models[mi].action);
    }
    rng.store(state.rng, slot);
    state.sim.post_step_action[slot] = action;
}

//---------------------------------------------------------------------------//
//! Sample the selected model's interaction and apply it to the track (one
//! instantiation per model class, so a kernel carries only its sampler)
template<ModelClass M>
CELER_FUNCTION void interaction_track(CoreParamsRef const& params,
                                      CoreStateRef const& state,
                                      size_type slot)
{
    PhysicsParamsRef const& pp = params.phys;
    double const energy = state.particle.energy[slot];
    double dir[3];
    for (int k = 0; k < 3; ++k)
        dir[k] = state.geo.dir[k][slot];
    std::uint32_t const mat = material_of(params, state, slot);
    XorwowEngine rng = XorwowEngine::load(state.rng, slot);
    InteractionContext ctx{pp, state, slot, mat, state.phys.element[slot],
                           state.particle.particle_id[slot]};
    Interaction r = sample_interaction(M, ctx, energy, dir, rng);
    rng.store(state.rng, slot);
    if (CELER_UNLIKELY(rng.exhausted))
    {
        // The sample was abandoned: no secondaries, no new state
        apply_sampling_failure(params, state, slot);
        return;
    }
    apply_interaction(params, state, slot, mat, r);
}

//---------------------------------------------------------------------------//
//! Cross into the next volume; leaving the world kills the track
CELER_FUNCTION void
boundary_track(CoreParamsRef const& params, CoreStateRef const& state, size_type slot)
{
    geo::GeoTrack g = geo::GeoTrack::load(state.geo, slot);
    geo::cross_boundary(params.geo, g);
    g.store(state.geo, slot);
    if (g.failed)
    {
        apply_errored(params, state, slot);
    }
    else if (!geo::is_outside(params.geo, g))
    {
        if (params.volume_to_material[g.volume] == invalid_id)
            apply_errored(params, state, slot);
    }
    else
    {
        state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::killed);
    }
}

//! Deposit the remaining energy (plus the pair mass for e+) and kill
CELER_FUNCTION void
tracking_cut_track(CoreParamsRef const& params, CoreStateRef const& state, size_type slot)
{
    std::uint32_t const pid = state.particle.particle_id[slot];
    double dep = state.particle.energy[slot];
    if (pid == params.phys.positron_id)
        dep += 2 * params.phys.mass[pid];
    state.phys.energy_deposition[slot] += dep;
    state.particle.energy[slot] = 0;
    state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::killed);
}

//---------------------------------------------------------------------------//
}  // namespace exec
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
