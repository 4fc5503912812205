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
    for (size_type e = 0; e < num_events; ++e)
    {
        CELER_VALIDATE(events[e].num_primaries > 0, << "event " << e << " has no primaries");
    }
    CoreState& state = stepper_->state();
    CELER_VALIDATE(state.counters().num_alive == 0 && state.counters().num_initializers == 0,
                   << "serving loop needs an empty state");
    // Event IDs restart at zero: clear the per-event tallies and counters
    auto const& init = state.ref().init;
    CELER_HIP_CALL(hipMemsetAsync(init.event_edep, 0, init.max_events * sizeof(double),
                                  state.stream()));
    CELER_HIP_CALL(hipMemsetAsync(init.event_live, 0,
                                  init.max_events * sizeof(std::int32_t), state.stream()));
    CELER_HIP_CALL(hipMemsetAsync(init.track_counters, 0,
                                  init.max_events * sizeof(std::uint32_t), state.stream()));
    CELER_HIP_CALL(hipMemsetAsync(&state.ref().counters->num_completed, 0,
                                  sizeof(std::uint32_t), state.stream()));
    size_type const max_queued = options.max_queued > 0 ? options.max_queued
                                                        : state.size();
    CELER_VALIDATE(max_queued <= params_->init_capacity(),
                   << "max_queued exceeds the initializer capacity");
    size_type const steps_per_sync = std::max<size_type>(options.steps_per_sync, 1);

    ServeResult result;
    result.admit_time.assign(num_events, -1);
    result.done_time.assign(num_events, -1);
    result.edep.assign(num_events, 0);

    using clock = std::chrono::steady_clock;
    auto const t0 = clock::now();
    auto elapsed = [&t0] {
        return std::chrono::duration<double>(clock::now() - t0).count();
    };

    std::vector<Primary> batch;
    std::vector<std::uint32_t> done_ids;
    size_type next = 0;  // next event to admit
    size_type completed = 0;
    std::uint32_t seen = 0;  // completions already read
    std::uint32_t const steps0 = state.counters().steps_taken;
    unsigned long long const track_steps0 = state.counters().track_steps;
    StepperResult last;
    while (completed < num_events)
    {
        // Admit events that have arrived while the queue has room
        double const now = elapsed();
        batch.clear();
        while (next < num_events && events[next].arrival <= now)
        {
            ServeEvent const& ev = events[next];
            bool const idle = batch.empty() && last.queued == 0 && last.alive == 0;
            if (!idle && last.queued + batch.size() + ev.num_primaries > max_queued)
                break;
            for (size_type i = 0; i < ev.num_primaries; ++i)
            {
                Primary p = primaries[ev.first_primary + i];
                p.event_id = next;
                batch.push_back(p);
            }
            result.admit_time[next] = now;
            ++next;
        }
        if (batch.empty() && last.queued == 0 && last.alive == 0)
        {
            // Nothing in flight: wait for the next arrival
            if (next < num_events)
            {
                double const wait = events[next].arrival - elapsed();
                if (wait > 0)
                {
                    std::this_thread::sleep_for(
                        std::chrono::duration<double>(std::min(wait, 1e-3)));
                }
                continue;
            }
            CELER_VALIDATE(completed == num_events,
                           << "serving loop stalled with " << num_events - completed
                           << " events unfinished");
            break;
        }
        if (!batch.empty())
        {
            last = (*stepper_)(batch.data(), static_cast<size_type>(batch.size()));
        }
        else
        {
            // Run ahead while no arrived event waits for admission
            bool const waiting = next < num_events && events[next].arrival <= elapsed();
            last = stepper_->run_steps(waiting ? 1 : steps_per_sync);
        }
        CoreCounters const& c = state.counters();
        result.steps = c.steps_taken - steps0;
        CELER_VALIDATE(options.max_steps == 0 || result.steps < options.max_steps,
                       << "serving loop exceeded " << options.max_steps << " steps");

        // New completions recorded by the device
        std::uint32_t const total = std::min<std::uint32_t>(c.num_completed, num_events);
        if (total > seen)
        {
            done_ids.resize(total - seen);
            CELER_HIP_CALL(hipMemcpyAsync(done_ids.data(), state.ref().init.completed + seen,
                                          done_ids.size() * sizeof(std::uint32_t),
                                          hipMemcpyDeviceToHost, state.stream()));
            CELER_HIP_CALL(hipStreamSynchronize(state.stream()));
            double const t = elapsed();
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
                                      state.stream()));
        CELER_HIP_CALL(hipStreamSynchronize(state.stream()));
    }
    return result;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
