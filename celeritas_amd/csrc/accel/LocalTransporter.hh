//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file accel/LocalTransporter.hh
//! Offload front end: buffer primaries and transport them on flush.
//---------------------------------------------------------------------------//
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "global/CoreData.hh"

namespace celeritas_amd
{
class CoreParams;
class Stepper;

//---------------------------------------------------------------------------//
//! Options used by the offload front end (accel/SetupOptions.hh subset)
struct SetupOptions
{
    size_type max_num_tracks{0};  //!< track slots of the state
    size_type auto_flush{0};  //!< flush at this many buffered primaries (0: max_num_tracks)
    size_type max_steps{1000000};  //!< step iterations before giving up
    bool nonfatal_flush{false};  //!< kill active tracks instead of throwing
    std::uint32_t stream_id{0};
    bool action_times{false};
};

//---------------------------------------------------------------------------//
/*!
 * Buffer tracks offloaded by a host application and transport them.
 *
 * Follows accel/LocalTransporter.{hh,cc}:52-300 without Geant4: Push takes a
 * Primary (the G4Track conversion is the host application's job), the
 * buffer flushes automatically at auto_flush tracks, Flush runs the stepper
 * until no track is queued or alive, and more than max_steps iterations
 * either throw or (nonfatal_flush) kill the active tracks. InitializeEvent
 * reseeds the track RNG streams for the event. As in the reference, pushed
 * tracks are given event 0 of the current event's RNG reseed.
 */
class LocalTransporter
{
  public:
    LocalTransporter(SetupOptions const& options, std::shared_ptr<CoreParams const> params);
    ~LocalTransporter();

    void InitializeEvent(int id);
    void Push(Primary const& track);
    void Flush();
    void Finalize();

    //! Accumulated device time per step action [s] (needs action_times)
    std::map<std::string, double> GetActionTime() const;
    size_type GetBufferSize() const { return static_cast<size_type>(buffer_.size()); }
    //! Step iterations of the last flush
    size_type last_flush_iterations() const { return last_iters_; }

    Stepper& stepper() { return *step_; }

  private:
    std::shared_ptr<CoreParams const> params_;
    std::shared_ptr<Stepper> step_;
    std::vector<Primary> buffer_;
    size_type auto_flush_{0};
    size_type max_steps_{0};
    bool nonfatal_{false};
    std::int64_t event_id_{-1};
    size_type last_iters_{0};
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
