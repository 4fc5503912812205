//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/AlongStepKernels.hip
//! Along-step action for charged particles without a field.
//!
//! One thread per slot runs the reference's appliers in order
//! (global/alongstep/AlongStepGeneralLinearAction.cc): linear propagation
//! to the physics step or the next boundary (PropagationApplier.hh:95-195),
//! time update (TimeUpdater.hh), mean continuous energy loss with the
//! lowest-electron-energy cut (ElossApplier.hh, MeanELoss.hh) and the
//! interaction-MFP update (TrackUpdater.hh). All per-slot state is read and
//! written once per launch.
//---------------------------------------------------------------------------//
#include "CoreActions.hh"
#include "corecel/WaveOps.hh"
#include "geo/OrangeLiteTracker.hh"
#include "phys/PhysicsImpl.hh"
#include "track/TrackInitImpl.hh"

namespace celeritas_amd
{
namespace
{
constexpr int block_size = 256;

__global__ void __launch_bounds__(block_size)
    along_step_charged_kernel(CoreParamsRef const params,
                              CoreStateRef const state,
                              std::uint8_t const action)
{
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    if (slot >= state.size)
        return;
    auto const status = static_cast<TrackStatus>(state.sim.status[slot]);
    if (status == TrackStatus::inactive || state.sim.along_step_action[slot] != action)
        return;
    if (state.diag)
        wave::count_active_lanes(state.diag + diag_along_charged);
    PhysicsParamsRef const& pp = params.phys;

    // Propagation
    double step = state.sim.step_length[slot];
    std::uint8_t post = state.sim.post_step_action[slot];
    if (step != 0)
    {
        geo::GeoTrack g = geo::GeoTrack::load(state.geo, slot);
        double const orig[3] = {g.pos[0], g.pos[1], g.pos[2]};
        geo::Propagation prop = geo::find_next_step(params.geo, g, step);
        if (prop.boundary)
            geo::move_to_boundary(g);
        else
            geo::move_internal(g, step);
        g.store(state.geo, slot);
        if (CELER_UNLIKELY(g.pos[0] == orig[0] && g.pos[1] == orig[1]
                           && g.pos[2] == orig[2]))
        {
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
    }
    if (status == TrackStatus::errored)
    {
        state.sim.step_length[slot] = step;
        state.sim.post_step_action[slot] = post;
        return;
    }

    // Time
    std::uint32_t const pid = state.particle.particle_id[slot];
    double energy = state.particle.energy[slot];
    double const v = phys::speed(energy, pp.mass[pid]);
    if (v > 0)
        state.sim.time[slot] += step / v;

    // Continuous energy loss
    TrackStatus new_status = status;
    ParticlePhysics const& part = pp.particle[pid];
    if (part.eloss_process >= 0 && energy > 0)
    {
        bool const apply_cut = post != params.scalars.boundary_action;
        double const lowest = pp.scalars.lowest_electron_energy;
        double deposited;
        if (apply_cut && energy < lowest)
        {
            deposited = energy;
        }
        else
        {
            std::uint32_t const mat = params.volume_to_material[state.geo.volume[slot]];
            double const eloss = phys::calc_mean_eloss(
                pp, part, mat, energy, state.phys.dedx_range[slot], step);
            deposited = (apply_cut && energy - eloss <= lowest) ? energy : eloss;
        }
        if (deposited > 0)
        {
            state.phys.energy_deposition[slot] += deposited;
            energy -= deposited;
            state.particle.energy[slot] = energy;
        }
        if (energy == 0)
        {
            if (!part.has_at_rest)
            {
                new_status = TrackStatus::killed;
                state.sim.status[slot] = static_cast<std::uint8_t>(new_status);
                post = pp.scalars.range_action;
            }
            else
            {
                post = pp.scalars.discrete_action;
            }
        }
    }

    // Interaction MFP and step count
    if (new_status == TrackStatus::alive && post != pp.scalars.discrete_action)
        state.phys.interaction_mfp[slot] -= step * state.phys.macro_xs[slot];
    state.sim.num_steps[slot] += 1;
    state.sim.step_length[slot] = step;
    state.sim.post_step_action[slot] = post;
}

//---------------------------------------------------------------------------//
}  // namespace

hipError_t launch_along_step_charged(CoreParamsRef const& p,
                                     CoreStateRef const& s,
                                     std::uint8_t action,
                                     hipStream_t st)
{
    unsigned int const nb = (s.size + block_size - 1) / block_size;
    along_step_charged_kernel<<<nb, block_size, 0, st>>>(p, s, action);
    return hipGetLastError();
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
