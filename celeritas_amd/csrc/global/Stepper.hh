//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/Stepper.hh
//---------------------------------------------------------------------------//
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "ActionInterface.hh"
#include "CoreData.hh"

namespace celeritas_amd
{
class CoreParams;
class CoreState;
class ExtendFromPrimariesAction;

//---------------------------------------------------------------------------//
//! Stepper construction (celeritas/global/Stepper.hh:43-55)
struct StepperInput
{
    std::shared_ptr<CoreParams const> params;
    std::uint32_t stream_id{0};
    size_type num_track_slots{0};
    bool action_times{false};
    //! Replay steps without new primaries from a captured hipGraph
    bool use_graph{false};
};

//! Counters after one step iteration (celeritas/global/Stepper.hh:61-70)
struct StepperResult
{
    size_type generated{0};
    size_type queued{0};
    size_type active{0};
    size_type alive{0};

    explicit operator bool() const { return queued > 0 || alive > 0; }
};

//---------------------------------------------------------------------------//
/*!
 * Runs step iterations over one state (celeritas/global/Stepper.cc:66-201).
 *
 * The sequence is the registered step actions sorted by (order, ID)
 * (corecel/sys/ActionGroups.t.hh:23-54). Each action enqueues its kernels on
 * the state's stream; the only host synchronization per step is the counter
 * readback that produces the StepperResult.
 *
 * Consecutive actions of one order that touch disjoint tracks (the along-
 * step actions; the interactions, boundary crossing and tracking cut) are
 * forked onto side streams and joined back with events, so a latency-bound
 * kernel with few waves shares the GPU with its siblings instead of running
 * alone; a captured step graph holds the same fork and join.
 */
class Stepper
{
  public:
    using SpanConstPrimary = std::pair<Primary const*, size_type>;

    explicit Stepper(StepperInput input);
    ~Stepper();

    //! Run all actions with no active tracks (allocations, first launches)
    void warm_up();
    //! Transport existing tracks one step
    StepperResult operator()();
    //! Transport existing tracks several steps with one counter readback
    //! at the end (the steps need no host input: launch sizes and ranges
    //! are device-resident)
    StepperResult run_steps(size_type count);
    //! Insert primaries then take a step
    StepperResult operator()(Primary const* primaries, size_type count);
    //! Kill all active tracks (next step applies the tracking cut)
    void kill_active();
    //! Reseed RNGs and track counters at the start of an event
    void reseed(std::uint64_t event_id);

    //! Enqueue one step without reading counters back
    void enqueue_step();

    //! Capture the step sequence (no new primaries) as a hipGraph and
    //! replay it for later steps without primaries
    void capture_step();
    //! Whether steps without primaries replay the captured graph
    bool uses_graph() const { return graph_exec_ != nullptr; }
    //! Drop the captured graph (after the state's device view changed); the
    //! next step without primaries captures it again
    void reset_graph();

    std::vector<std::shared_ptr<StepActionInterface const>> const& actions() const
    {
        return actions_;
    }
    CoreState& state() { return *state_; }
    std::shared_ptr<CoreState> const& sp_state() const { return state_; }
    CoreParams const& params() const { return *params_; }
    //! Accumulated per-action device times [s] when action_times is on
    std::vector<double> const& action_times() const { return times_; }
    //! Number of timed steps per action
    std::vector<std::uint64_t> const& action_counts() const { return counts_; }
    void reset_action_times();

  private:
    std::shared_ptr<CoreParams const> params_;
    std::shared_ptr<CoreState> state_;
    std::vector<std::shared_ptr<StepActionInterface const>> actions_;
    std::vector<std::string> labels_;  //!< profiler range names of actions_
    std::shared_ptr<ExtendFromPrimariesAction const> primaries_action_;
    bool action_times_{false};
    std::vector<double> times_;
    std::vector<std::uint64_t> counts_;
    //! Timed steps whose events are recorded before a readback
    static constexpr std::size_t max_timed_steps = 16;
    std::vector<hipEvent_t> events_;  //!< [step][action][begin, end]
    std::size_t timed_steps_pending_{0};
    bool use_graph_{false};
    hipGraphExec_t graph_exec_{nullptr};
    //! Side streams of concurrent groups (with the state's own stream, one
    //! per hardware queue a process gets by default)
    static constexpr std::size_t max_side_streams = 3;
    //! [begin, end) of each group of actions_ launched together
    std::vector<std::pair<std::size_t, std::size_t>> groups_;
    std::vector<hipStream_t> side_streams_;
    std::vector<hipEvent_t> fork_events_;  //!< one per concurrent group
    std::vector<hipEvent_t> join_events_;  //!< one per side launch

    StepperResult make_result();
    void launch_actions();
    void read_action_times();
    StepperResult step_impl(bool new_primaries);
    void launch_step(bool new_primaries);
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
