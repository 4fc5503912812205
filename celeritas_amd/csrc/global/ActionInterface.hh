//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/ActionInterface.hh
//! Step-action plugin interface and registry.
//!
//! Same contract as the reference (corecel/sys/ActionInterface.hh:58-186,
//! corecel/sys/ActionRegistry.hh:42-105): actions have a unique ID equal to
//! the registry's next_id() at insertion, a unique label, a description and
//! (for step actions) an order; the sequence runs step actions sorted by
//! (order, ID). Only the device memory space exists here: there is no host
//! execution of the transport loop, so step(params, CoreState&) is the one
//! entry point and it enqueues kernels on the state's HIP stream.
//---------------------------------------------------------------------------//
#pragma once

#include <map>
#include <memory>
#include <string>
#include <string_view>
#include <vector>

#include "corecel/DeviceRuntime.hh"
#include "corecel/Types.hh"

namespace celeritas_amd
{
class CoreParams;
class CoreState;

using ActionId = std::uint32_t;

//---------------------------------------------------------------------------//
//! Base for all actions: ID, label, description
class ActionInterface
{
  public:
    virtual ~ActionInterface() = default;
    virtual ActionId action_id() const = 0;
    virtual std::string_view label() const = 0;
    virtual std::string_view description() const = 0;
};

//! Action that sets up per-state data before the first step
class BeginRunActionInterface : virtual public ActionInterface
{
  public:
    virtual void begin_run(CoreParams const&, CoreState&) = 0;
};

//! Action executed every step iteration
class StepActionInterface : virtual public ActionInterface
{
  public:
    virtual StepActionOrder order() const = 0;
    virtual void step(CoreParams const&, CoreState&) const = 0;
    //! Whether the action reads and writes only the tracks whose along- or
    //! post-step action is its own (the reference's ConditionalTrackExecutor
    //! filter), so that such actions of one order touch disjoint tracks and
    //! may run concurrently
    virtual bool track_disjoint() const { return false; }
    //! Whether a step of this action enqueues any work (an action whose
    //! work another action launches on its behalf returns false)
    virtual bool enqueues_work(CoreParams const&) const { return true; }
};

using CoreStepActionInterface = StepActionInterface;

//---------------------------------------------------------------------------//
//! Mixin with ID, label and description storage
class ConcreteAction : virtual public ActionInterface
{
  public:
    ConcreteAction(ActionId id, std::string label, std::string description = {})
        : id_(id), label_(std::move(label)), description_(std::move(description))
    {
    }
    ActionId action_id() const final { return id_; }
    std::string_view label() const final { return label_; }
    std::string_view description() const final { return description_; }

  private:
    ActionId id_;
    std::string label_;
    std::string description_;
};

//! Implicit action: an ID that tracks can select but that launches nothing
class ImplicitAction final : public ConcreteAction
{
  public:
    using ConcreteAction::ConcreteAction;
};

//---------------------------------------------------------------------------//
/*!
 * Registry of actions; IDs are assigned in insertion order.
 */
class ActionRegistry
{
  public:
    using SPConstAction = std::shared_ptr<ActionInterface const>;

    ActionId next_id() const { return static_cast<ActionId>(actions_.size()); }

    //! Insert an immutable action (ID must equal next_id())
    void insert(SPConstAction action);
    //! Insert a mutable (begin-run) action
    void insert_mutable(std::shared_ptr<ActionInterface> action);

    size_type num_actions() const { return static_cast<size_type>(actions_.size()); }
    bool empty() const { return actions_.empty(); }
    SPConstAction const& action(ActionId id) const;
    std::string const& id_to_label(ActionId id) const { return labels_.at(id); }
    //! Find by label: returns invalid_id if absent
    ActionId find_action(std::string const& label) const;

    std::vector<std::shared_ptr<ActionInterface>> const& mutable_actions() const
    {
        return mutable_;
    }

  private:
    std::vector<SPConstAction> actions_;
    std::vector<std::string> labels_;
    std::map<std::string, ActionId> label_to_id_;
    std::vector<std::shared_ptr<ActionInterface>> mutable_;

    void insert_impl(SPConstAction action);
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
