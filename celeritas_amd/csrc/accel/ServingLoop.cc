//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file accel/ServingLoop.cc
//---------------------------------------------------------------------------//
#include "ServingLoop.hh"

#include <algorithm>
#include <chrono>
#include <thread>

#include "global/CoreParams.hh"
#include "global/CoreState.hh"
#include "track/TrackKernels.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
ServingLoop::ServingLoop(std::shared_ptr<CoreParams const> params,
                         size_type num_track_slots,
                         std::uint32_t stream_id)
    : params_(std::move(params))
{
    CELER_VALIDATE(params_, << "serving loop requires params");
    StepperInput inp;
    inp.params = params_;
    inp.stream_id = stream_id;
    inp.num_track_slots = num_track_slots;
    // Steps without newly admitted events replay one captured graph
    inp.use_graph = true;
    stepper_ = std::make_unique<Stepper>(std::move(inp));
    // Completion is detected from the per-event tallies
    stepper_->state().event_tally(true);
}

//---------------------------------------------------------------------------//
/*!
 * Size the device queue for a run and attach it to the state.
 *
 * The captured step graph holds the state's device view, so a queue that
 * moves (grows) or a new admission limit drops the graph; the next step
 * captures it again.
 */
void ServingLoop::attach_queue(size_type num_primaries,
                               size_type num_events,
                               size_type max_queued)
{
    bool changed = false;
    if (primaries_.size() < num_primaries)
    {
        primaries_.resize(num_primaries);
        changed = true;
    }
    if (event_first_.size() < num_events)
    {
        event_first_.resize(num_events);
        event_count_.resize(num_events);
        admit_step_.resize(num_events);
        changed = true;
    }
    if (counters_.size() == 0)
    {
        counters_.resize(1);
        changed = true;
    }
    if (max_queued != max_queued_)
        changed = true;
    if (!changed)
        return;
    ServeQueueRef q;
    q.primaries = primaries_.data();
    q.event_first = event_first_.data();
    q.event_count = event_count_.data();
    q.admit_step = admit_step_.data();
    q.counters = counters_.data();
    q.num_events = static_cast<size_type>(event_first_.size());
    q.max_queued = max_queued;
    max_queued_ = max_queued;
    stepper_->reset_graph();
    stepper_->state().serve_queue(q);
}

//---------------------------------------------------------------------------//
ServeResult ServingLoop::run(Primary const* primaries,
                             ServeEvent const* events,
                             size_type num_events,
                             ServeOptions const& options)
{
    CELER_VALIDATE(num_events <= params_->max_events(),
                   << "serving " << num_events << " events needs max_events >= "
                   << num_events << " (have " << params_->max_events() << ")");
    for (size_type e = 1; e < num_events; ++e)
    {
        CELER_VALIDATE(events[e].arrival >= events[e - 1].arrival,
                       << "events must be given in arrival order");
    }
    CoreState& state = stepper_->state();
    CELER_VALIDATE(state.counters().num_alive == 0 && state.counters().num_initializers == 0,
                   << "serving loop needs an empty state");
    size_type const max_queued = options.max_queued > 0 ? options.max_queued
                                                        : state.size();
    CELER_VALIDATE(max_queued <= params_->init_capacity(),
                   << "max_queued exceeds the initializer capacity");
    size_type const steps_per_sync = std::max<size_type>(options.steps_per_sync, 1);

    // Events lie in arrival order in the queue, each contiguous, with the
    // event index as event ID (IDs restart at zero every run)
    std::vector<std::uint32_t> first(num_events);
    std::vector<std::uint32_t> count(num_events);
    std::vector<Primary> staged;
    for (size_type e = 0; e < num_events; ++e)
    {
        ServeEvent const& ev = events[e];
        CELER_VALIDATE(ev.num_primaries > 0, << "event " << e << " has no primaries");
        CELER_VALIDATE(ev.num_primaries <= params_->init_capacity(),
                       << "event " << e << " has more primaries (" << ev.num_primaries
                       << ") than the initializer capacity");
        first[e] = static_cast<std::uint32_t>(staged.size());
        count[e] = static_cast<std::uint32_t>(ev.num_primaries);
        for (size_type i = 0; i < ev.num_primaries; ++i)
        {
            Primary p = primaries[ev.first_primary + i];
            p.event_id = static_cast<std::uint32_t>(e);
            staged.push_back(p);
        }
    }
    this->attach_queue(std::max<size_type>(staged.size(), 1),
                       std::max<size_type>(num_events, 1), max_queued);

    // Clear the per-event tallies, counters and the queue's counters
    hipStream_t const stream = state.stream();
    auto const& init = state.ref().init;
    CELER_HIP_CALL(hipMemsetAsync(init.event_edep, 0, init.max_events * sizeof(double), stream));
    CELER_HIP_CALL(
        hipMemsetAsync(init.event_live, 0, init.max_events * sizeof(std::int32_t), stream));
    CELER_HIP_CALL(hipMemsetAsync(init.track_counters, 0,
                                  init.max_events * sizeof(std::uint32_t), stream));
    CELER_HIP_CALL(hipMemsetAsync(&state.ref().counters->num_completed, 0,
                                  sizeof(std::uint32_t), stream));
    CELER_HIP_CALL(hipMemsetAsync(counters_.data(), 0, sizeof(ServeCounters), stream));
    if (num_events > 0)
    {
        CELER_HIP_CALL(hipMemcpyAsync(event_first_.data(), first.data(),
                                      num_events * sizeof(std::uint32_t),
                                      hipMemcpyHostToDevice, stream));
        CELER_HIP_CALL(hipMemcpyAsync(event_count_.data(), count.data(),
                                      num_events * sizeof(std::uint32_t),
                                      hipMemcpyHostToDevice, stream));
    }
    CELER_HIP_CALL(hipStreamSynchronize(stream));

    ServeResult result;
    result.admit_time.assign(num_events, -1);
    result.done_time.assign(num_events, -1);
    result.edep.assign(num_events, 0);

    using clock = std::chrono::steady_clock;
    auto const t0 = clock::now();
    auto elapsed = [&t0] {
        return std::chrono::duration<double>(clock::now() - t0).count();
    };

    std::vector<std::uint32_t> done_ids;
    size_type arrived = 0;  // events staged on the device
    size_type admitted = 0;  // events the device has admitted (last readback)
    size_type completed = 0;
    std::uint32_t seen = 0;  // completions already read
    std::uint32_t const steps0 = state.counters().steps_taken;
    unsigned long long const track_steps0 = state.counters().track_steps;
    StepperResult last;
    ServeCounters sc{};
    while (completed < num_events)
    {
        // Stage the events that have arrived and publish their count
        size_type const before = arrived;
        double const now = elapsed();
        while (arrived < num_events && events[arrived].arrival <= now)
        {
            CELER_HIP_CALL(hipMemcpyAsync(primaries_.data() + first[arrived],
                                          staged.data() + first[arrived],
                                          count[arrived] * sizeof(Primary),
                                          hipMemcpyHostToDevice, stream));
            ++arrived;
        }
        if (arrived != before)
        {
            CELER_HIP_CALL(launch_fill_u32(&counters_.data()->arrived, 1,
                                           static_cast<std::uint32_t>(arrived), stream));
        }
        if (admitted == arrived && last.queued == 0 && last.alive == 0)
        {
            // Nothing in flight: wait for the next arrival
            CELER_VALIDATE(arrived < num_events,
                           << "serving loop stalled with " << num_events - completed
                           << " events unfinished");
            double const wait = events[arrived].arrival - elapsed();
            if (wait > 0)
            {
                std::this_thread::sleep_for(
                    std::chrono::duration<double>(std::min(wait, 1e-3)));
            }
            continue;
        }
        // Run ahead unless an arrival is due to be staged
        bool const due = arrived < num_events && events[arrived].arrival <= elapsed();
        last = stepper_->run_steps(due ? 1 : steps_per_sync);
        CoreCounters const& c = state.counters();
        result.steps = c.steps_taken - steps0;
        CELER_VALIDATE(options.max_steps == 0 || result.steps < options.max_steps,
                       << "serving loop exceeded " << options.max_steps << " steps");

        // Admissions and completions recorded by the device
        CELER_HIP_CALL(hipMemcpyAsync(&sc, counters_.data(), sizeof(ServeCounters),
                                      hipMemcpyDeviceToHost, stream));
        std::uint32_t const total = std::min<std::uint32_t>(c.num_completed, num_events);
        if (total > seen)
        {
            done_ids.resize(total - seen);
            CELER_HIP_CALL(hipMemcpyAsync(done_ids.data(), state.ref().init.completed + seen,
                                          done_ids.size() * sizeof(std::uint32_t),
                                          hipMemcpyDeviceToHost, stream));
        }
        CELER_HIP_CALL(hipStreamSynchronize(stream));
        double const t = elapsed();
        for (; admitted < sc.admitted; ++admitted)
            result.admit_time[admitted] = t;
        if (total > seen)
        {
            for (std::uint32_t e : done_ids)
            {
                CELER_VALIDATE(e < num_events && result.done_time[e] < 0,
                               << "event " << e << " completed twice");
                result.done_time[e] = t;
                ++completed;
            }
            seen = total;
        }
    }
    result.track_steps = state.counters().track_steps - track_steps0;
    result.wall_time = elapsed();
    if (num_events > 0)
    {
        CELER_HIP_CALL(hipMemcpyAsync(result.edep.data(), state.ref().init.event_edep,
                                      num_events * sizeof(double), hipMemcpyDeviceToHost,
                                      stream));
        CELER_HIP_CALL(hipStreamSynchronize(stream));
    }
    return result;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
