//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file track/TrackActions.hh
//! Track-management step actions (celeritas/track/*Action.hh).
//---------------------------------------------------------------------------//
#pragma once

#include <vector>

#include "global/ActionInterface.hh"
#include "global/CoreData.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
/*!
 * Primaries -> initializers (ExtendFromPrimariesAction.cc:103-200).
 *
 * insert() copies host primaries to the state's device buffer; step()
 * appends them to the initializer queue. Inserting twice before a step is
 * rejected like the reference (CELER_NOT_IMPLEMENTED there).
 */
class ExtendFromPrimariesAction final : public StepActionInterface,
                                        public ConcreteAction
{
  public:
    explicit ExtendFromPrimariesAction(ActionId id);
    static char const* label_str() { return "extend-from-primaries"; }

    void insert(CoreParams const& params,
                CoreState& state,
                Primary const* host_primaries,
                size_type count) const;
    //! Insert primaries already resident in device memory
    void insert_device(CoreParams const& params,
                       CoreState& state,
                       Primary const* device_primaries,
                       size_type count) const;
    void step(CoreParams const&, CoreState&) const final;
    StepActionOrder order() const final { return StepActionOrder::generate; }

  private:
    struct Pending
    {
        CoreState const* state;
        Primary const* data;
        std::uint32_t const* ranks;
        size_type count;
    };
    mutable std::vector<Pending> pending_;
};

//---------------------------------------------------------------------------//
//! Fill vacancies from initializers (InitializeTracksAction.cc:56-97)
class InitializeTracksAction final : public StepActionInterface,
                                     public ConcreteAction
{
  public:
    explicit InitializeTracksAction(ActionId id)
        : ConcreteAction(id, "initialize-tracks", "initialize track states")
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    StepActionOrder order() const final { return StepActionOrder::start; }
};

//---------------------------------------------------------------------------//
//! Compaction and secondaries (ExtendFromSecondariesAction.cc:56-100)
class ExtendFromSecondariesAction final : public StepActionInterface,
                                          public ConcreteAction
{
  public:
    explicit ExtendFromSecondariesAction(ActionId id)
        : ConcreteAction(id,
                         "extend-from-secondaries",
                         "create track initializers from secondaries")
    {
    }
    void step(CoreParams const&, CoreState&) const final;
    StepActionOrder order() const final { return StepActionOrder::end; }
};

//---------------------------------------------------------------------------//
/*!
 * Reindex track slots (SortTracksAction.cc:44-135).
 *
 * Labels and step orders follow the reference: reindex_status and
 * reindex_particle_type sort at sort_start, reindex_along_step_action at
 * sort_pre (after pre-step), reindex_step_limit_action at sort_pre_post
 * (after discrete selection). The per-key offsets stay in device memory
 * for the sorted launches of the next actions.
 */
class SortTracksAction final : public StepActionInterface, public ConcreteAction
{
  public:
    SortTracksAction(ActionId id, TrackOrder order, size_type num_keys);
    void step(CoreParams const&, CoreState&) const final;
    StepActionOrder order() const final { return action_order_; }
    TrackOrder track_order() const { return track_order_; }

  private:
    TrackOrder track_order_;
    StepActionOrder action_order_;
    SortKey key_;
    size_type num_keys_;
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
