//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/PhysicsActions.cc
//---------------------------------------------------------------------------//
#include "PhysicsActions.hh"

#include "global/CoreActions.hh"
#include "global/CoreParams.hh"
#include "global/CoreState.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
void PreStepAction::step(CoreParams const& params, CoreState& state) const
{
    // Pre-step visits every slot once and each slot's result depends only on
    // that slot. When the along-step sort has run before it (keyed by the
    // action the pre-step will choose, so by particle charge) it visits the
    // slots in that order: a wave then holds tracks of one kind (photons or
    // charged particles) and evaluates one set of processes; otherwise it
    // reads the slots in place
    ThreadSlots ts;
    TrackOrder const order = params.track_order();
    if (order == TrackOrder::reindex_along_step_action
        || order == TrackOrder::reindex_both_action)
    {
        ts.slots = state.ref().track_slots;
    }
    bool const mixed = params.precision() == PrecisionMode::mixed;
    CELER_HIP_CALL((mixed ? launch_pre_step_mixed : launch_pre_step)(
        params.device_ref(), state.ref(), ts, state.stream()));
}

bool DiscreteSelectAction::enqueues_work(CoreParams const& params) const
{
    return !params.device_ref().scalars.fuse_discrete_select;
}

void DiscreteSelectAction::step(CoreParams const& params, CoreState& state) const
{
    // Fused: the along-step kernels made the selection
    if (!this->enqueues_work(params))
        return;
    bool const mixed = params.precision() == PrecisionMode::mixed;
    CELER_HIP_CALL((mixed ? launch_discrete_select_mixed : launch_discrete_select)(
        params.device_ref(), state.ref(), make_thread_slots(params, state, *this),
        state.stream()));
}

bool ModelAction::enqueues_work(CoreParams const& params) const
{
    return !params.fuse_interactions()
           || this->action_id() == params.device_ref().phys.scalars.first_model_action;
}

void ModelAction::step(CoreParams const& params, CoreState& state) const
{
    bool const mixed = params.precision() == PrecisionMode::mixed;
    if (params.fuse_interactions())
    {
        // The first model action samples every model's tracks in one
        // launch over the union of their (consecutive) sorted ranges
        auto const& sc = params.device_ref().phys.scalars;
        if (this->action_id() != sc.first_model_action)
            return;
        ThreadSlots ts = make_thread_slots(params, state, *this);
        if (ts.range)
            ts.range_end = state.ref().action_offsets + sc.first_model_action
                           + sc.num_models;
        CELER_HIP_CALL((mixed ? launch_interactions_mixed : launch_interactions)(
            params.device_ref(), state.ref(), ts, state.stream()));
        return;
    }
    CELER_HIP_CALL((mixed ? launch_interaction_mixed : launch_interaction)(
        params.device_ref(), state.ref(), static_cast<std::uint8_t>(this->action_id()),
        model_class_, make_thread_slots(params, state, *this), state.stream()));
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
