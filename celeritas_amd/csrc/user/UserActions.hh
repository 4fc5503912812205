//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file user/UserActions.hh
//! User step actions: post-step action and step-count diagnostics and the
//! step collector for detector step points.
//!
//! Counterparts of user/ActionDiagnostic.hh, user/StepDiagnostic.hh and
//! user/StepCollector.hh with detail/StepGatherAction.hh: each is created
//! with make_and_insert() on the core params before a stepper is built,
//! allocates its per-stream data in begin_run(), and accumulates on the
//! device across steps (and across states of the same stream ID) until the
//! host reads it.
//---------------------------------------------------------------------------//
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "corecel/DeviceRuntime.hh"
#include "global/ActionInterface.hh"
#include "UserData.hh"

namespace celeritas_amd
{
class CoreParams;

//---------------------------------------------------------------------------//
/*!
 * Histogram of (particle, bin) counts per stream, summed over streams on
 * read (ParticleTallyData.hh with the StreamStore of the reference).
 */
class ParticleTallyStore
{
  public:
    ParticleTallyStore(size_type num_particles, size_type num_bins);

    //! Device view of a stream's counts (allocated and zeroed on first use)
    ParticleTallyRef state(std::uint32_t stream_id, hipStream_t stream);
    //! Counts [particle][bin] summed over streams (synchronizes the device)
    std::vector<std::uint64_t> sum() const;
    //! Zero every stream's counts
    void clear();

    size_type num_particles() const { return num_particles_; }
    size_type num_bins() const { return num_bins_; }

  private:
    size_type num_particles_;
    size_type num_bins_;
    std::map<std::uint32_t, DeviceVector<unsigned long long>> counts_;
};

//---------------------------------------------------------------------------//
/*!
 * Count post-step actions by particle type (ActionDiagnostic.hh:38-120).
 *
 * Runs at the post-step order over every slot whose track is valid (alive,
 * errored or killed this step). The action dimension is the registry size
 * at the first begin_run, so every action created before the stepper is
 * counted.
 */
class ActionDiagnostic final : public StepActionInterface,
                               public BeginRunActionInterface,
                               public ConcreteAction
{
  public:
    static std::shared_ptr<ActionDiagnostic> make_and_insert(CoreParams const& core);

    ActionDiagnostic(ActionId id, size_type num_particles);

    void begin_run(CoreParams const&, CoreState&) final;
    void step(CoreParams const&, CoreState&) const final;
    StepActionOrder order() const final { return StepActionOrder::post; }

    //! Counts [particle][action] summed over streams
    std::vector<std::uint64_t> calc_actions() const;
    size_type num_particles() const { return num_particles_; }
    size_type num_actions() const { return store_ ? store_->num_bins() : 0; }
    void clear();

  private:
    size_type num_particles_;
    std::unique_ptr<ParticleTallyStore> store_;
};

//---------------------------------------------------------------------------//
/*!
 * Distribution of steps per track by particle type (StepDiagnostic.hh).
 *
 * Runs at the user post-step order over the tracks killed this step; bins
 * are 0..max_step_bin plus one overflow bin (num_bins = max_step_bin + 2).
 */
class StepDiagnostic final : public StepActionInterface,
                             public BeginRunActionInterface,
                             public ConcreteAction
{
  public:
    static std::shared_ptr<StepDiagnostic>
    make_and_insert(CoreParams const& core, size_type max_step_bin);

    StepDiagnostic(ActionId id, size_type num_particles, size_type max_step_bin);

    void begin_run(CoreParams const&, CoreState&) final;
    void step(CoreParams const&, CoreState&) const final;
    StepActionOrder order() const final { return StepActionOrder::user_post; }

    //! Counts [particle][num_steps] summed over streams
    std::vector<std::uint64_t> calc_steps() const;
    size_type num_particles() const { return store_.num_particles(); }
    size_type num_bins() const { return store_.num_bins(); }
    void clear() { store_.clear(); }

  private:
    mutable ParticleTallyStore store_;
};

//---------------------------------------------------------------------------//
/*!
 * Host copy of the hits of detector steps (DetectorStepOutput,
 * DetectorSteps.hh:47-80): one entry per step of a track in a detector, in
 * step order and, within a step, in slot order. Columns of fields that were
 * not selected are empty.
 */
struct DetectorStepOutput
{
    struct Point
    {
        std::vector<double> time;
        std::vector<double> pos;  //!< [hit][3]
        std::vector<double> dir;  //!< [hit][3]
        std::vector<std::uint32_t> volume_id;
        std::vector<double> energy;
    };
    std::vector<std::uint32_t> detector;
    std::vector<std::uint32_t> track_id;
    Point points[2];
    std::vector<std::uint32_t> event_id;
    std::vector<std::uint32_t> parent_id;
    std::vector<std::uint32_t> track_step_count;
    std::vector<std::uint32_t> action_id;
    std::vector<double> step_length;
    std::vector<std::uint32_t> particle;
    std::vector<double> energy_deposition;
    //! Steps whose hits did not fit in the buffer, and the hits lost
    std::uint32_t overflow_steps{0};
    std::uint32_t dropped_hits{0};

    size_type size() const { return static_cast<size_type>(detector.size()); }
};

/*!
 * Gather step points of tracks in detectors (StepCollector.hh:47-126 with
 * StepGatherAction<pre/post> and DetectorSteps' copy_steps).
 *
 * make_and_insert registers the "step-gather-pre" action (user pre-step
 * order; it also allocates each stream's data at begin_run) and the
 * "step-gather-post" action (user post-step order: compaction into the hit
 * buffer). Each stream's hits accumulate in a device buffer of
 * hit_capacity entries until read(); a step whose hits do not all fit is
 * dropped and counted.
 */
class StepCollector
{
  public:
    struct Input
    {
        StepSelection selection;
        //! Volume -> detector ID (invalid_id: not a detector); empty:
        //! every active track is collected with detector 0
        std::vector<std::uint32_t> volume_to_detector;
        bool nonzero_energy_deposition{false};
        size_type hit_capacity{1u << 20};
    };

    static std::shared_ptr<StepCollector>
    make_and_insert(CoreParams const& core, Input inp);

    StepCollector(CoreParams const& core, Input inp);
    ~StepCollector();

    //! Allocate a stream's data for a state (reused while the size matches)
    void begin_run(CoreParams const&, CoreState&);
    //! Copy and clear the hits of a stream (synchronizes its stream)
    DetectorStepOutput read(std::uint32_t stream_id);

    StepSelection const& selection() const { return inp_.selection; }
    //! Device view of a stream's data (null before its begin_run)
    StepCollectorRef const* stream_ref(std::uint32_t stream_id) const;

  private:
    struct StreamData;

    Input inp_;
    DeviceVector<std::uint32_t> d_detector_map_;
    std::map<std::uint32_t, std::unique_ptr<StreamData>> streams_;
};

//! Step-point gather action (detail/StepGatherAction.hh)
class StepGatherAction final : public StepActionInterface,
                               public BeginRunActionInterface,
                               public ConcreteAction
{
  public:
    StepGatherAction(ActionId id, StepPoint point, std::shared_ptr<StepCollector> owner);
    void begin_run(CoreParams const&, CoreState&) final;
    void step(CoreParams const&, CoreState&) const final;
    StepActionOrder order() const final
    {
        return point_ == step_pre ? StepActionOrder::user_pre
                                  : StepActionOrder::user_post;
    }

  private:
    StepPoint point_;
    std::shared_ptr<StepCollector> owner_;
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
