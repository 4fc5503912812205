//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/Stepper.cc
//---------------------------------------------------------------------------//
#include "Stepper.hh"

#include <algorithm>
#include <chrono>

#include "corecel/ScopedProfiling.hh"
#include "track/TrackActions.hh"
#include "track/TrackKernels.hh"
#include "CoreParams.hh"
#include "CoreState.hh"

namespace celeritas_amd
{
hipError_t launch_kill_active(CoreParamsRef const&, CoreStateRef const&, hipStream_t);

//---------------------------------------------------------------------------//
Stepper::Stepper(StepperInput input)
    : params_(std::move(input.params))
    , action_times_(input.action_times)
    , use_graph_(input.use_graph)
{
    CELER_VALIDATE(params_, << "stepper requires params");
    CELER_VALIDATE(input.num_track_slots > 0, << "stepper requires track slots");
    auto const& reg = *params_->action_reg();
    for (ActionId id = 0; id < reg.num_actions(); ++id)
    {
        auto sa = std::dynamic_pointer_cast<StepActionInterface const>(reg.action(id));
        if (sa)
            actions_.push_back(std::move(sa));
    }
    std::stable_sort(actions_.begin(), actions_.end(), [](auto const& a, auto const& b) {
        return std::make_pair(a->order(), a->action_id())
               < std::make_pair(b->order(), b->action_id());
    });
    times_.assign(actions_.size(), 0.0);
    counts_.assign(actions_.size(), 0);
    for (auto const& a : actions_)
        labels_.emplace_back(a->label());

    // Groups of consecutive track-disjoint actions of one order run on
    // concurrent streams; every other action is a group of its own
    std::size_t num_concurrent = 0;
    for (std::size_t i = 0; i < actions_.size();)
    {
        std::size_t j = i + 1;
        if (params_->concurrent_actions() && actions_[i]->track_disjoint())
        {
            while (j < actions_.size() && actions_[j]->track_disjoint()
                   && actions_[j]->order() == actions_[i]->order())
                ++j;
        }
        std::size_t working = 0;
        for (std::size_t k = i; k < j; ++k)
            working += actions_[k]->enqueues_work(*params_) ? 1 : 0;
        if (working > 1)
            ++num_concurrent;
        else
            j = i + 1;
        groups_.emplace_back(i, j);
        i = j;
    }
    if (num_concurrent > 0)
    {
        side_streams_.resize(max_side_streams);
        for (auto& s : side_streams_)
            CELER_HIP_CALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        fork_events_.resize(num_concurrent);
        join_events_.resize(num_concurrent * max_side_streams);
        for (auto* ev : {&fork_events_, &join_events_})
            for (auto& e : *ev)
                CELER_HIP_CALL(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }

    ActionId pid = reg.find_action(ExtendFromPrimariesAction::label_str());
    CELER_VALIDATE(pid != invalid_id,
                   << "primary generator was not added to the stepping loop");
    primaries_action_
        = std::dynamic_pointer_cast<ExtendFromPrimariesAction const>(reg.action(pid));

    state_ = std::make_shared<CoreState>(*params_, input.stream_id, input.num_track_slots);
    for (auto const& m : reg.mutable_actions())
    {
        if (auto br = std::dynamic_pointer_cast<BeginRunActionInterface>(m))
            br->begin_run(*params_, *state_);
    }
}

Stepper::~Stepper()
{
    for (auto e : events_)
        (void)hipEventDestroy(e);
    for (auto* ev : {&fork_events_, &join_events_})
        for (auto e : *ev)
            (void)hipEventDestroy(e);
    for (auto s : side_streams_)
        (void)hipStreamDestroy(s);
    if (graph_exec_)
        (void)hipGraphExecDestroy(graph_exec_);
}

//---------------------------------------------------------------------------//
/*!
 * Capture one step of the action sequence into a graph.
 *
 * Every action enqueues only kernels whose arguments are fixed for the
 * state (device pointers, action IDs, device-resident launch ranges and
 * counters), and the primaries action launches nothing when none are
 * pending, so the same graph is valid for every later step without new
 * primaries. The per-step counter reset and readback stay outside it.
 */
void Stepper::capture_step()
{
    CELER_VALIDATE(!action_times_, << "per-action timing cannot be captured in a graph");
    if (graph_exec_)
        return;
    hipStream_t const stream = state_->stream();
    hipGraph_t graph = nullptr;
    CELER_HIP_CALL(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    this->launch_actions();
    CELER_HIP_CALL(hipStreamEndCapture(stream, &graph));
    hipError_t const err = hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    CELER_HIP_CALL(err);
    use_graph_ = true;
}

void Stepper::reset_graph()
{
    if (graph_exec_)
    {
        CELER_HIP_CALL(hipStreamSynchronize(state_->own_stream()));
        CELER_HIP_CALL(hipGraphExecDestroy(graph_exec_));
        graph_exec_ = nullptr;
    }
}

void Stepper::reset_action_times()
{
    timed_steps_pending_ = 0;
    std::fill(times_.begin(), times_.end(), 0.0);
    std::fill(counts_.begin(), counts_.end(), 0);
}

//---------------------------------------------------------------------------//
void Stepper::enqueue_step()
{
    if (action_times_ && !state_->warming_up())
    {
        // Device time per action from events bracketing its launches on the
        // state's stream; read back at the next counter sync (or when the
        // event ring is full), so the host can run ahead of the device and
        // the brackets hold no host gaps
        if (events_.empty())
        {
            events_.resize(2 * actions_.size() * max_timed_steps);
            for (auto& e : events_)
                CELER_HIP_CALL(hipEventCreate(&e));
        }
        if (timed_steps_pending_ == max_timed_steps)
            this->read_action_times();
        std::size_t const base = 2 * actions_.size() * timed_steps_pending_;
        for (std::size_t i = 0; i < actions_.size(); ++i)
        {
            ScopedProfiling prof(labels_[i].c_str());
            CELER_HIP_CALL(hipEventRecord(events_[base + 2 * i], state_->stream()));
            actions_[i]->step(*params_, *state_);
            CELER_HIP_CALL(hipEventRecord(events_[base + 2 * i + 1], state_->stream()));
        }
        ++timed_steps_pending_;
    }
    else
    {
        this->launch_actions();
    }
}

/*!
 * Enqueue every action once, forking each concurrent group onto the side
 * streams: the group's first working action stays on the state's stream,
 * the others go round-robin to the side streams, which wait for the work
 * before the group and are waited for by the state's stream after it.
 */
void Stepper::launch_actions()
{
    // Launches go back to the state's own stream however the group ends
    struct Restore
    {
        CoreState& s;
        ~Restore() { s.launch_stream(nullptr); }
    } restore{*state_};

    hipStream_t const own = state_->own_stream();
    std::size_t fork = 0;
    for (auto const& [begin, end] : groups_)
    {
        if (end - begin == 1)
        {
            ScopedProfiling prof(labels_[begin].c_str());
            actions_[begin]->step(*params_, *state_);
            continue;
        }
        CELER_HIP_CALL(hipEventRecord(fork_events_[fork], own));
        std::vector<bool> used(side_streams_.size(), false);
        std::size_t working = 0;
        for (std::size_t i = begin; i < end; ++i)
        {
            hipStream_t launch = nullptr;
            if (actions_[i]->enqueues_work(*params_) && working++ > 0)
            {
                std::size_t const s = (working - 2) % side_streams_.size();
                if (!used[s])
                {
                    CELER_HIP_CALL(
                        hipStreamWaitEvent(side_streams_[s], fork_events_[fork], 0));
                    used[s] = true;
                }
                launch = side_streams_[s];
            }
            state_->launch_stream(launch);
            ScopedProfiling prof(labels_[i].c_str());
            actions_[i]->step(*params_, *state_);
        }
        state_->launch_stream(nullptr);
        for (std::size_t s = 0; s < side_streams_.size(); ++s)
        {
            if (!used[s])
                continue;
            hipEvent_t const join = join_events_[fork * side_streams_.size() + s];
            CELER_HIP_CALL(hipEventRecord(join, side_streams_[s]));
            CELER_HIP_CALL(hipStreamWaitEvent(own, join, 0));
        }
        ++fork;
    }
}

void Stepper::read_action_times()
{
    if (timed_steps_pending_ == 0)
        return;
    std::size_t const per_step = 2 * actions_.size();
    CELER_HIP_CALL(hipEventSynchronize(events_[per_step * timed_steps_pending_ - 1]));
    for (std::size_t p = 0; p < timed_steps_pending_; ++p)
    {
        for (std::size_t i = 0; i < actions_.size(); ++i)
        {
            float ms = 0;
            CELER_HIP_CALL(hipEventElapsedTime(
                &ms, events_[p * per_step + 2 * i], events_[p * per_step + 2 * i + 1]));
            times_[i] += 1e-3 * ms;
            counts_[i] += 1;
        }
    }
    timed_steps_pending_ = 0;
}

StepperResult Stepper::make_result()
{
    CoreCounters const& c = state_->sync_counters();
    this->read_action_times();
    CELER_VALIDATE(!(c.error_flags & err_initializer_overflow),
                   << "insufficient capacity (" << params_->init_capacity()
                   << ") for track initializers (created " << c.num_secondaries
                   << " new secondaries for a total capacity requirement of "
                   << c.num_initializers << ")");
    CELER_VALIDATE(!(c.error_flags & err_secondary_overflow),
                   << "secondary stack capacity exceeded");
    CELER_VALIDATE(!(c.error_flags & err_event_overflow),
                   << "event number exceeds max_events=" << params_->max_events());
    StepperResult r;
    r.generated = c.num_generated;
    r.active = c.num_active;
    r.alive = c.num_alive;
    r.queued = c.num_initializers;
    return r;
}

StepperResult Stepper::operator()()
{
    return this->step_impl(false);
}

void Stepper::launch_step(bool new_primaries)
{
    if (use_graph_ && !new_primaries && !action_times_ && !state_->warming_up())
    {
        if (!graph_exec_)
            this->capture_step();
        ScopedProfiling prof("step-graph");
        CELER_HIP_CALL(hipGraphLaunch(graph_exec_, state_->stream()));
    }
    else
    {
        this->enqueue_step();
    }
}

StepperResult Stepper::step_impl(bool new_primaries)
{
    // Generated count is per step (Stepper.cc:125-139)
    CELER_HIP_CALL(hipMemsetAsync(&state_->ref().counters->num_generated, 0,
                                  sizeof(std::uint32_t), state_->stream()));
    this->launch_step(new_primaries);
    return this->make_result();
}

StepperResult Stepper::run_steps(size_type count)
{
    CELER_VALIDATE(count > 0, << "no steps requested");
    CELER_HIP_CALL(hipMemsetAsync(&state_->ref().counters->num_generated, 0,
                                  sizeof(std::uint32_t), state_->stream()));
    for (size_type i = 0; i < count; ++i)
        this->launch_step(false);
    return this->make_result();
}

StepperResult Stepper::operator()(Primary const* primaries, size_type count)
{
    CELER_VALIDATE(count > 0, << "no primaries given");
    std::uint32_t max_event = 0;
    for (size_type i = 0; i < count; ++i)
        max_event = std::max(max_event, primaries[i].event_id);
    CELER_VALIDATE(max_event < params_->max_events(),
                   << "event number " << max_event
                   << " exceeds max_events=" << params_->max_events());
    primaries_action_->insert(*params_, *state_, primaries, count);
    return this->step_impl(true);
}

//---------------------------------------------------------------------------//
void Stepper::warm_up()
{
    CELER_VALIDATE(state_->counters().num_active == 0,
                   << "cannot warm up when state has active tracks");
    state_->warming_up(true);
    this->enqueue_step();
    state_->warming_up(false);
    state_->sync_counters();
}

void Stepper::kill_active()
{
    CELER_HIP_CALL(launch_kill_active(params_->device_ref(), state_->ref(),
                                      state_->stream()));
}

void Stepper::reseed(std::uint64_t event_id)
{
    CELER_HIP_CALL(launch_reseed_rng(state_->ref().rng,
                                     params_->device_ref().rng_jump,
                                     params_->rng().seed(),
                                     event_id,
                                     state_->stream()));
    CELER_HIP_CALL(launch_fill_u32(state_->ref().init.track_counters,
                                   params_->max_events(), 0, state_->stream()));
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
