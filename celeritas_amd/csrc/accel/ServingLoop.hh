//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file accel/ServingLoop.hh
//! Continuous-batching transport of events that arrive over time.
//---------------------------------------------------------------------------//
#pragma once

#include <memory>
#include <vector>

#include "global/CoreData.hh"
#include "global/Stepper.hh"

namespace celeritas_amd
{
class CoreParams;

//---------------------------------------------------------------------------//
//! One request: a contiguous range of primaries and its arrival time
struct ServeEvent
{
    size_type first_primary{0};
    size_type num_primaries{0};
    double arrival{0};  //!< [s] after the start of the run
};

struct ServeOptions
{
    //! Admit new events only while queued initializers stay below this
    //! (0: the number of track slots)
    size_type max_queued{0};
    //! Abort after this many step iterations (0: no limit)
    size_type max_steps{0};
    //! Step iterations launched per counter readback while no event is
    //! waiting to be admitted (the device runs them back to back)
    size_type steps_per_sync{4};
};

//! Per-event outcome and run totals
struct ServeResult
{
    std::vector<double> admit_time;  //!< [s]
    std::vector<double> done_time;  //!< [s]
    std::vector<double> edep;  //!< [MeV] deposited by the event's tracks
    double wall_time{0};  //!< [s]
    size_type steps{0};  //!< step iterations
    unsigned long long track_steps{0};  //!< sum of active tracks per step
};

/*!
 * Transport a stream of events through one track-slot state.
 *
 * Events are admitted in arrival order once they have arrived (wall clock)
 * and while the initializer queue has room, so requests of uneven size
 * share the device with those already in flight: their primaries join the
 * next step, and their tracks fill slots as others free them. Completion
 * is detected on the device: the track-management kernels keep a live
 * count per event (initializers plus tracks) and append an event to a
 * completion list when its count reaches zero. While no arrived event is
 * waiting, the loop launches several steps (graph replays) back to back
 * and reads only the counters and the new completions afterwards; an
 * event's completion time is that of the readback that first shows it.
 * Latency is completion minus arrival.
 */
class ServingLoop
{
  public:
    ServingLoop(std::shared_ptr<CoreParams const> params,
                size_type num_track_slots,
                std::uint32_t stream_id = 0);

    ServeResult run(Primary const* primaries,
                    ServeEvent const* events,
                    size_type num_events,
                    ServeOptions const& options);

    Stepper& stepper() { return *stepper_; }

  private:
    std::shared_ptr<CoreParams const> params_;
    std::unique_ptr<Stepper> stepper_;
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
