//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file track/TrackInitImpl.hh
//! Per-track initialization shared by the init and secondary kernels.
//---------------------------------------------------------------------------//
#pragma once

#include "geo/OrangeLiteTracker.hh"
#include "global/CoreData.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Write the sim part of an initializer (celeritas/track/SimTrackView.hh:148)
CELER_FORCEINLINE_FUNCTION void
init_sim(SimStateRef const& sim, size_type slot, TrackInitializer const& ti)
{
    sim.track_id[slot] = ti.track_id;
    sim.parent_id[slot] = ti.parent_id;
    sim.event_id[slot] = ti.event_id;
    sim.num_steps[slot] = 0;
    sim.time[slot] = ti.time;
    sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::initializing);
    sim.step_length[slot] = 0;
    sim.post_step_action[slot] = invalid_id8;
    sim.along_step_action[slot] = invalid_id8;
    sim.num_looping_steps[slot] = 0;
}

//! Reset persistent physics state (PhysicsTrackView::operator=, :207-212)
CELER_FORCEINLINE_FUNCTION void
reset_physics(PhysicsStateRef const& phys, size_type slot)
{
    phys.interaction_mfp[slot] = 0;
    phys.msc_range_init[slot] = 0;
    phys.msc_range_factor[slot] = 0;
    phys.msc_limit_min[slot] = 0;
}

//! Mark a track as failed (celeritas/global/CoreTrackView.hh:340-347)
CELER_FORCEINLINE_FUNCTION void apply_errored(CoreParamsRef const& params,
                                              CoreStateRef const& state,
                                              size_type slot)
{
    state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::errored);
    state.sim.along_step_action[slot] = invalid_id8;
    state.sim.post_step_action[slot] = params.scalars.tracking_cut_action;
}

/*!
 * Fail a track whose sampling loop ran out of its iteration budget and
 * record it: the error flag (not fatal to the run) and a count of such
 * tracks in the device counters.
 */
CELER_FUNCTION void apply_sampling_failure(CoreParamsRef const& params,
                                           CoreStateRef const& state,
                                           size_type slot)
{
    apply_errored(params, state, slot);
#if CELER_DEVICE_COMPILE
    atomicOr(&state.counters->error_flags, static_cast<std::uint32_t>(err_sampling_budget));
    atomicAdd(&state.counters->num_sampling_failures, 1u);
#else
    state.counters->error_flags |= err_sampling_budget;
    state.counters->num_sampling_failures += 1;
#endif
}

/*!
 * Initialize a track in a vacant slot from an initializer.
 *
 * Follows InitTracksExecutor (celeritas/track/detail/InitTracksExecutor.hh:
 * 71-175). A secondary starts at its parent's position, which lies inside
 * the parent's volume when the parent was not on a surface (interactions
 * happen away from boundaries), so its geometry state is that volume with
 * the state a point location would give there, taken from the volume the
 * initializer carries instead of searched for; primaries, and secondaries
 * of parents on a surface, are located from the position. (The reference
 * copies the parent's geometry from its slot, for the secondaries of the
 * step just taken only; carrying the volume in the initializer extends the
 * same shortcut to secondaries that waited in the queue.)
 */
CELER_FUNCTION void init_track(CoreParamsRef const& params,
                               CoreStateRef const& state,
                               size_type slot,
                               TrackInitializer const& ti)
{
    init_sim(state.sim, slot, ti);
    state.particle.particle_id[slot] = static_cast<std::uint8_t>(ti.particle_id);
    state.particle.energy[slot] = ti.energy;

    geo::GeoTrack g;
    if (ti.volume != invalid_id)
    {
        geo::initialize_interior(g, ti.volume, ti.pos, ti.dir);
    }
    else
    {
        geo::initialize(params.geo, g, ti.pos, ti.dir);
    }
    g.store(state.geo, slot);
    if (CELER_UNLIKELY(g.failed || geo::is_outside(params.geo, g)))
    {
        apply_errored(params, state, slot);
        return;
    }
    std::uint32_t const mat = params.volume_to_material[g.volume];
    if (CELER_UNLIKELY(mat == invalid_id))
    {
        apply_errored(params, state, slot);
        return;
    }
    reset_physics(state.phys, slot);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
