//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/AlongStepExecutor.hh
//! Per-track body of the charged along-step action.
//!
//! Runs the reference's appliers in order
//! (global/alongstep/AlongStepGeneralLinearAction.cc): Urban MSC step limit
//! (MscStepLimitApplier, UrbanMsc::limit_step), linear propagation to the
//! geometric step or the next boundary (PropagationApplier.hh:95-195), or
//! curved propagation in a uniform field, MSC path conversion, scattering
//! and displacement (MscApplier, UrbanMsc::apply_step), time update
//! (TimeUpdater.hh), continuous energy loss with Urban fluctuations and the
//! lowest-electron-energy cut (ElossApplier.hh, FluctELoss.hh /
//! MeanELoss.hh) and the interaction-MFP update (TrackUpdater.hh). The
//! geometry, RNG and step state of a track are read once into registers and
//! written back once. The same function compiles for the host.
//---------------------------------------------------------------------------//
#pragma once

#include <cmath>

#include "corecel/Precision.hh"
#include "field/FieldPropagator.hh"
#include "geo/OrangeLiteTracker.hh"
#include "phys/AlongStepPhysics.hh"
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
template<bool UseField>
CELER_FORCEINLINE_FUNCTION void along_step_charged_track(CoreParamsRef const& params,
                                             CoreStateRef const& state,
                                             size_type slot,
                                             TrackStatus status)
{
    PhysicsParamsRef const& pp = params.phys;
    PhysicsStateRef const& ps = state.phys;

    double step = state.sim.step_length[slot];
    std::uint8_t post = state.sim.post_step_action[slot];
    std::uint32_t const pid = state.particle.particle_id[slot];
    double energy = state.particle.energy[slot];
    std::uint32_t const mat = params.volume_to_material[state.geo.volume[slot]];
    ParticlePhysics const& part = pp.particle[pid];
    geo::GeoTrack g = geo::GeoTrack::load(state.geo, slot);
    XorwowEngine rng = XorwowEngine::load(state.rng, slot);

    // MSC step limit: true path -> geometric path
    bool const msc_applies = pp.msc.enabled && step > alongstep::msc_geom_limit
                             && status == TrackStatus::alive
                             && (pid == pp.electron_id || pid == pp.positron_id)
                             && energy > pp.msc.low_energy_limit
                             && energy < pp.msc.high_energy_limit;
    double true_path = 0, geom_path = 0, alpha = 0;
    bool displaced = false;
    alongstep::MscRange msc_range{ps.msc_range_init[slot], ps.msc_range_factor[slot],
                                  ps.msc_limit_min[slot]};
    alongstep::MscTrack mt{pp, part, mat, pid == pp.electron_id ? 0u : 1u, energy,
                           ps.dedx_range[slot], 0};
    if (msc_applies)
    {
        mt.lambda = mt.calc_msc_mfp(energy);
        if (step <= alongstep::msc_limit_min_fix)
        {
            true_path = step;
        }
        else
        {
            double safety = 0;
            bool done = false;
            if (!g.is_on_boundary())
            {
                double const max_step = mt.max_step();
                safety = geo::find_safety(params.geo, g);
                if (safety >= max_step)
                {
                    true_path = step;
                    done = true;
                }
            }
            if (!done)
            {
                displaced = true;
                true_path = alongstep::step_limit(mt, msc_range, g.is_on_boundary(), safety,
                                                  step, rng);
                ps.msc_range_init[slot] = msc_range.range_init;
                ps.msc_range_factor[slot] = msc_range.range_factor;
                ps.msc_limit_min[slot] = msc_range.limit_min;
            }
        }
        bool limited = true_path < step;
        geom_path = alongstep::to_geo(mt, true_path, &alpha);
        if (geom_path > mt.lambda)
        {
            geom_path = mt.lambda;
            limited = true;
        }
        step = geom_path;
        if (limited)
            post = pp.scalars.msc_action;
    }

    // Propagation (PropagationApplier.hh:95-195): linear, or curved in the
    // uniform field with looping-track handling
    if (step != 0)
    {
        double const orig[3] = {g.pos[0], g.pos[1], g.pos[2]};
        geo::Propagation prop;
        bool looping = false;
        if constexpr (UseField)
        {
            FieldParamsRef const& fp = params.field;
            double const mass = pp.mass[pid];
            double const momentum = std::sqrt(energy * energy + 2 * mass * energy);
            field::UniformFieldStepper const stepper{
                {fp.field[0], fp.field[1], fp.field[2]},
                pp.charge[pid] * field::lorentz_per_cm_tesla};
            field::FieldPropagation const fprop
                = field::propagate(params.geo, g, stepper, fp.options, momentum, step);
            prop.distance = fprop.distance;
            prop.boundary = fprop.boundary;
            looping = fprop.looping;
        }
        else
        {
            prop = geo::find_next_step(params.geo, g, step);
            if (prop.boundary)
                geo::move_to_boundary(g);
            else
                geo::move_internal(g, step);
        }
        // A move that leaves the position unchanged, or not finite, cannot
        // make progress: the track is failed (a finite FP64 state never
        // reaches the second test)
        if (CELER_UNLIKELY((g.pos[0] == orig[0] && g.pos[1] == orig[1]
                            && g.pos[2] == orig[2])
                           || !std::isfinite(g.pos[0] + g.pos[1] + g.pos[2])))
        {
            g.store(state.geo, slot);
            rng.store(state.rng, slot);
            apply_errored(params, state, slot);
            return;
        }
        if constexpr (UseField)
        {
            // SimTrackView::update_looping / is_looping (SimTrackView.hh:189-219)
            std::uint32_t nloop = looping ? state.sim.num_looping_steps[slot] + 1 : 0;
            state.sim.num_looping_steps[slot] = nloop;
            if (looping)
            {
                step = prop.distance;
                LoopingThreshold const& lt = params.field.looping;
                std::uint32_t const limit = energy < lt.threshold_energy
                                                ? lt.max_subthreshold_steps
                                                : lt.max_steps;
                post = nloop >= limit ? params.scalars.tracking_cut_action
                                      : params.scalars.propagation_limit_action;
            }
        }
        if (!looping)
        {
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
    }

    // MSC: geometric -> true path, scattering and displacement
    if (msc_applies && geom_path > 0 && status == TrackStatus::alive)
    {
        if (post == params.scalars.boundary_action
            || post == params.scalars.propagation_limit_action)
        {
            geom_path = step;
            true_path = alongstep::from_geo(mt, true_path, alpha, geom_path);
            displaced = false;
        }
        step = true_path;
        double safety = 0;
        if (displaced)
        {
            double displ = alongstep::calc_displacement(geom_path, true_path);
            displ *= 1 + 2 * alongstep::msc_safety_tol;
            displ = displ > alongstep::msc_geom_limit ? displ : alongstep::msc_geom_limit;
            safety = geo::find_safety(params.geo, g);
            safety = safety < displ ? safety : displ;
            if (CELER_UNLIKELY(safety == 0))
                displaced = false;
        }
        alongstep::MscResult const res = alongstep::msc_scatter(
            mt, pp.materials[mat], pid == pp.positron_id, g.dir, safety, displaced,
            geom_path, true_path, ps.msc_limit_min[slot], rng);
        if (res.action != 0)
            geo::set_dir(params.geo, g, res.dir);
        if (res.action == 2)
        {
            double const x[3] = {g.pos[0] + res.displacement[0],
                                 g.pos[1] + res.displacement[1],
                                 g.pos[2] + res.displacement[2]};
            geo::move_internal(g, x);
        }
    }
    g.store(state.geo, slot);

    // Time
    double const v = phys::speed(energy, pp.mass[pid]);
    if (v > 0)
        state.sim.time[slot] += step / v;

    // Continuous energy loss
    TrackStatus new_status = status;
    if (part.eloss_process >= 0 && energy > 0)
    {
        bool const apply_cut = post != params.scalars.boundary_action;
        double const deposited = alongstep::calc_eloss(
            pp, part, pid, mat, energy, ps.dedx_range[slot], step, apply_cut, rng);
        if (deposited > 0)
        {
            ps.energy_deposition[slot] += deposited;
            energy -= deposited;
#if CELER_PREC_IS_MIXED
            if (CELER_UNLIKELY(!(energy >= 0 && energy < HUGE_VAL)))
                energy = 0;
#endif
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
    rng.store(state.rng, slot);
    if (CELER_UNLIKELY(rng.exhausted))
    {
        // A scattering or energy-loss sample was abandoned
        apply_sampling_failure(params, state, slot);
        return;
    }

    // Interaction MFP and step count
    if (new_status == TrackStatus::alive && post != pp.scalars.discrete_action)
        ps.interaction_mfp[slot] -= step * ps.macro_xs[slot];
    state.sim.num_steps[slot] += 1;
    state.sim.step_length[slot] = step;
    state.sim.post_step_action[slot] = post;
}

//---------------------------------------------------------------------------//
}  // namespace exec
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
