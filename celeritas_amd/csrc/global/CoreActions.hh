//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/CoreActions.hh
//! Core physics/geometry step actions and their kernel launchers.
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>

#include "corecel/Json.hh"
#include "ActionInterface.hh"
#include "CoreData.hh"

namespace celeritas_amd
{
class PhysicsParams;

//---------------------------------------------------------------------------//
// KERNEL LAUNCHERS (implemented in the .hip translation units)
//---------------------------------------------------------------------------//
hipError_t launch_pre_step(CoreParamsRef const&,
                           CoreStateRef const&,
                           ThreadSlots const&,
                           hipStream_t);
hipError_t launch_along_step_neutral(CoreParamsRef const&,
                                     CoreStateRef const&,
                                     std::uint8_t action,
                                     ThreadSlots const&,
                                     hipStream_t);
hipError_t launch_along_step_charged(CoreParamsRef const&,
                                     CoreStateRef const&,
                                     std::uint8_t action,
                                     ThreadSlots const&,
                                     hipStream_t);
hipError_t launch_discrete_select(CoreParamsRef const&,
                                  CoreStateRef const&,
                                  ThreadSlots const&,
                                  hipStream_t);
hipError_t launch_interaction(CoreParamsRef const&,
                              CoreStateRef const&,
                              std::uint8_t action,
                              std::uint32_t model,
                              ThreadSlots const&,
                              hipStream_t);
//! Every discrete model in one launch (tracks whose post-step action is
//! one of the model actions)
hipError_t launch_interactions(CoreParamsRef const&,
                               CoreStateRef const&,
                               ThreadSlots const&,
                               hipStream_t);
hipError_t launch_interactions_mixed(CoreParamsRef const&,
                                     CoreStateRef const&,
                                     ThreadSlots const&,
                                     hipStream_t);
hipError_t launch_boundary(CoreParamsRef const&,
                           CoreStateRef const&,
                           ThreadSlots const&,
                           hipStream_t);
hipError_t launch_tracking_cut(CoreParamsRef const&,
                               CoreStateRef const&,
                               ThreadSlots const&,
                               hipStream_t);
hipError_t launch_calo_tally(CoreParamsRef const&,
                             CoreStateRef const&,
                             size_type num_detectors,
                             ThreadSlots const&,
                             hipStream_t);
hipError_t launch_kill_active(CoreParamsRef const&, CoreStateRef const&, hipStream_t);

// Mixed-precision builds of the physics kernels (FP32 transcendentals and
// draws in sampling, interpolation and MSC)
hipError_t launch_pre_step_mixed(CoreParamsRef const&,
                                 CoreStateRef const&,
                                 ThreadSlots const&,
                                 hipStream_t);
hipError_t launch_along_step_charged_mixed(CoreParamsRef const&,
                                           CoreStateRef const&,
                                           std::uint8_t action,
                                           ThreadSlots const&,
                                           hipStream_t);
hipError_t launch_discrete_select_mixed(CoreParamsRef const&,
                                        CoreStateRef const&,
                                        ThreadSlots const&,
                                        hipStream_t);
hipError_t launch_interaction_mixed(CoreParamsRef const&,
                                    CoreStateRef const&,
                                    std::uint8_t action,
                                    std::uint32_t model,
                                    ThreadSlots const&,
                                    hipStream_t);

//---------------------------------------------------------------------------//
//! Whether tracks are reindexed through track_slots (TrackOrder >= shuffle)
inline bool is_reindexed(TrackOrder t)
{
    return static_cast<int>(t) >= static_cast<int>(TrackOrder::reindex_shuffle);
}

//! Whether an action of this order launches over its sorted thread range
//! (ActionInterface.hh:63-73)
bool is_action_sorted(StepActionOrder aorder, TrackOrder torder);

//! Thread-to-slot map for launching an action on a state
ThreadSlots make_thread_slots(CoreParams const& params,
                              CoreState const& state,
                              StepActionInterface const& action);

//! Per-volume energy deposition tally (user/SimpleCalo, StepGather pre/post)
class CaloTallyAction final : public StepActionInterface, public ConcreteAction
{
  public:
    explicit CaloTallyAction(ActionId id)
        : ConcreteAction(id, "simple-calo", "tally energy deposition by volume")
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    StepActionOrder order() const final { return StepActionOrder::user_post; }
};

//---------------------------------------------------------------------------//
//! Linear propagation of neutral particles (AlongStepNeutralAction)
class AlongStepNeutralAction final : public StepActionInterface,
                                     public ConcreteAction
{
  public:
    explicit AlongStepNeutralAction(ActionId id)
        : ConcreteAction(id,
                         "along-step-neutral",
                         "along-step for neutral particles")
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    bool track_disjoint() const final { return true; }
    StepActionOrder order() const final { return StepActionOrder::along; }
};

//! Charged along-step: MSC, propagation, energy loss (linear or field)
class AlongStepChargedAction final : public StepActionInterface,
                                     public ConcreteAction
{
  public:
    AlongStepChargedAction(ActionId id, std::string label)
        : ConcreteAction(id, std::move(label), "along-step for charged particles")
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    bool track_disjoint() const final { return true; }
    StepActionOrder order() const final { return StepActionOrder::along; }
};

//! Cross a geometry boundary (geo/detail/BoundaryExecutor.hh)
class BoundaryAction final : public StepActionInterface, public ConcreteAction
{
  public:
    explicit BoundaryAction(ActionId id)
        : ConcreteAction(id, "geo-boundary", "cross a geometry boundary")
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    bool track_disjoint() const final { return true; }
    StepActionOrder order() const final { return StepActionOrder::post; }
};

//! Kill a track and deposit its energy (phys/detail/TrackingCutExecutor.hh)
class TrackingCutAction final : public StepActionInterface, public ConcreteAction
{
  public:
    explicit TrackingCutAction(ActionId id)
        : ConcreteAction(id, "tracking-cut", "kill a track and deposit its energy")
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    bool track_disjoint() const final { return true; }
    StepActionOrder order() const final { return StepActionOrder::post; }
};

//! Create and register the charged along-step action; returns its label
std::string make_along_step_action(std::string const& kind,
                                   PhysicsParams const& physics,
                                   JsonValue const& doc,
                                   ActionRegistry* reg);

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
