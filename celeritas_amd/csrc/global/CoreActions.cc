//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/CoreActions.cc
//---------------------------------------------------------------------------//
#include "CoreActions.hh"

#include "CoreParams.hh"
#include "CoreState.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
void AlongStepNeutralAction::step(CoreParams const& params, CoreState& state) const
{
    CELER_HIP_CALL(launch_along_step_neutral(
        params.device_ref(), state.ref(),
        static_cast<std::uint8_t>(this->action_id()),
        make_thread_slots(params, state, *this), state.stream()));
}

void AlongStepChargedAction::step(CoreParams const& params, CoreState& state) const
{
    auto launch = params.precision() == PrecisionMode::mixed
                      ? launch_along_step_charged_mixed
                      : launch_along_step_charged;
    CELER_HIP_CALL(launch(
        params.device_ref(), state.ref(),
        static_cast<std::uint8_t>(this->action_id()),
        make_thread_slots(params, state, *this), state.stream()));
}

void BoundaryAction::step(CoreParams const& params, CoreState& state) const
{
    CELER_HIP_CALL(launch_boundary(params.device_ref(), state.ref(),
                                   make_thread_slots(params, state, *this),
                                   state.stream()));
}

void TrackingCutAction::step(CoreParams const& params, CoreState& state) const
{
    CELER_HIP_CALL(launch_tracking_cut(params.device_ref(), state.ref(),
                                       make_thread_slots(params, state, *this),
                                       state.stream()));
}

void CaloTallyAction::step(CoreParams const& params, CoreState& state) const
{
    // Every slot is visited once in any order, so the tally reads the slots
    // in place (coalesced) rather than through the sorted permutation
    CELER_HIP_CALL(launch_calo_tally(params.device_ref(), state.ref(),
                                     params.num_detectors(), ThreadSlots{},
                                     state.stream()));
}

//---------------------------------------------------------------------------//
bool is_action_sorted(StepActionOrder aorder, TrackOrder torder)
{
    return (aorder == StepActionOrder::post
            && torder == TrackOrder::reindex_step_limit_action)
           || (aorder == StepActionOrder::along
               && torder == TrackOrder::reindex_along_step_action)
           || (torder == TrackOrder::reindex_both_action
               && (aorder == StepActionOrder::post
                   || aorder == StepActionOrder::along));
}

ThreadSlots make_thread_slots(CoreParams const& params,
                              CoreState const& state,
                              StepActionInterface const& action)
{
    ThreadSlots ts;
    TrackOrder const torder = params.track_order();
    if (!is_reindexed(torder))
        return ts;
    ts.slots = state.ref().track_slots;
    if (is_action_sorted(action.order(), torder))
        ts.range = state.ref().action_offsets + action.action_id();
    return ts;
}

//---------------------------------------------------------------------------//
std::string make_along_step_action(std::string const& kind,
                                   PhysicsParams const&,
                                   JsonValue const&,
                                   ActionRegistry* reg)
{
    std::string label;
    if (kind == "neutral")
    {
        label = "along-step-neutral";
        reg->insert(std::make_shared<AlongStepNeutralAction>(reg->next_id()));
        return label;
    }
    if (kind == "general_linear")
        label = "along-step-general-linear";
    else if (kind == "uniform_field")
        label = "along-step-uniform-msc";
    else
        CELER_VALIDATE(false, << "unknown along-step kind '" << kind << "'");
    reg->insert(std::make_shared<AlongStepChargedAction>(reg->next_id(), label));
    return label;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
