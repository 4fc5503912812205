//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file accel/ServingLoop.hh
//! Continuous-batching transport of events that arrive over time.
//---------------------------------------------------------------------------//
#pragma once

#include <memory>
#include <vector>

#include "corecel/DeviceRuntime.hh"
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
    //! Step iterations launched per counter readback while no arrival is
    //! due to be staged (the device runs them back to back, admitting
    //! staged events at the start of each)
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
 * Admission and completion happen on the device. When an event arrives
 * (wall clock) the host copies its primaries into a device queue and
 * raises the queue's arrival count; nothing else about admission is
 * decided on the host. At the start of every step the device admits
 * arrived events in order while the queued initializers stay within
 * \c max_queued (the first waiting event regardless when nothing is queued
 * or alive), so requests of uneven size share the device with those in
 * flight: their tracks fill slots as others free them. The track-
 * management kernels keep a live count per event (initializers plus
 * tracks) and append an event to a completion list when its count reaches
 * zero. Every step, admission included, is one replay of the captured step
 * graph; the host launches several back to back (fewer while arrivals are
 * pending, so that they are staged promptly) and reads the counters, the
 * queue's admission count and the new completions afterwards. An event's
 * admission and completion times are those of the readback that first
 * shows them; latency is completion minus arrival.
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
    // Device queue (attached to the state; grown and re-attached on demand)
    DeviceVector<Primary> primaries_;
    DeviceVector<std::uint32_t> event_first_;
    DeviceVector<std::uint32_t> event_count_;
    DeviceVector<std::uint32_t> admit_step_;
    DeviceVector<ServeCounters> counters_;
    size_type max_queued_{0};

    void attach_queue(size_type num_primaries, size_type num_events, size_type max_queued);
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
