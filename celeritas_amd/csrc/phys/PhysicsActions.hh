//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/PhysicsActions.hh
//! Step actions registered by the physics (celeritas/phys/detail/*Action).
//---------------------------------------------------------------------------//
#pragma once

#include "global/ActionInterface.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Pre-step: MFP sampling and physics step limits (PreStepAction)
class PreStepAction final : public StepActionInterface, public ConcreteAction
{
  public:
    explicit PreStepAction(ActionId id)
        : ConcreteAction(id, "pre-step", "update beginning-of-step state")
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    StepActionOrder order() const final { return StepActionOrder::pre; }
};

//! Select the discrete interaction (DiscreteSelectAction)
class DiscreteSelectAction final : public StepActionInterface, public ConcreteAction
{
  public:
    explicit DiscreteSelectAction(ActionId id)
        : ConcreteAction(id, "physics-discrete-select", "select a discrete interaction")
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    bool enqueues_work(CoreParams const&) const final;
    StepActionOrder order() const final { return StepActionOrder::pre_post; }
};

//! Sample the interaction of one model for the tracks that selected it
class ModelAction final : public StepActionInterface, public ConcreteAction
{
  public:
    ModelAction(ActionId id, std::string label, std::uint32_t model_class)
        : ConcreteAction(id, std::move(label), "interact by a discrete model")
        , model_class_(model_class)
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    bool track_disjoint() const final { return true; }
    bool enqueues_work(CoreParams const&) const final;
    StepActionOrder order() const final { return StepActionOrder::post; }
    std::uint32_t model_class() const { return model_class_; }

  private:
    std::uint32_t model_class_;
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
