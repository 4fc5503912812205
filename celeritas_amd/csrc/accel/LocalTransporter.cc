//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file accel/LocalTransporter.cc
//---------------------------------------------------------------------------//
#include "LocalTransporter.hh"

#include <iostream>

#include "corecel/DeviceRuntime.hh"
#include "global/CoreParams.hh"
#include "global/Stepper.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
LocalTransporter::LocalTransporter(SetupOptions const& options,
                                   std::shared_ptr<CoreParams const> params)
    : params_(std::move(params))
    , auto_flush_(options.auto_flush ? options.auto_flush : options.max_num_tracks)
    , max_steps_(options.max_steps)
    , nonfatal_(options.nonfatal_flush)
{
    CELER_VALIDATE(params_, << "params were not initialized before constructing "
                               "the local transporter");
    CELER_VALIDATE(options.max_num_tracks > 0, << "max_num_tracks must be positive");
    CELER_VALIDATE(max_steps_ > 0, << "max_steps must be positive");
    StepperInput inp;
    inp.params = params_;
    inp.stream_id = options.stream_id;
    inp.num_track_slots = options.max_num_tracks;
    inp.action_times = options.action_times;
    step_ = std::make_shared<Stepper>(std::move(inp));
}

LocalTransporter::~LocalTransporter() = default;

//---------------------------------------------------------------------------//
void LocalTransporter::InitializeEvent(int id)
{
    CELER_VALIDATE(id >= 0, << "event ID must be nonnegative");
    event_id_ = id;
    step_->reseed(static_cast<std::uint64_t>(id));
}

//---------------------------------------------------------------------------//
void LocalTransporter::Push(Primary const& track)
{
    CELER_VALIDATE(track.particle_id < params_->physics().num_particles(),
                   << "cannot offload particle " << track.particle_id);
    Primary p = track;
    p.event_id = 0;
    buffer_.push_back(p);
    if (buffer_.size() >= auto_flush_)
        this->Flush();
}

//---------------------------------------------------------------------------//
void LocalTransporter::Flush()
{
    if (buffer_.empty())
        return;
    auto counts = (*step_)(buffer_.data(), static_cast<size_type>(buffer_.size()));
    buffer_.clear();
    size_type iters = 1;
    while (counts)
    {
        if (iters >= max_steps_)
        {
            if (!nonfatal_)
            {
                throw RuntimeError("number of step iterations exceeded the allowed maximum ("
                                   + std::to_string(max_steps_) + ")");
            }
            std::cerr << "celeritas_amd: number of step iterations exceeded the "
                         "allowed maximum ("
                      << max_steps_ << "); killing active tracks\n";
            step_->kill_active();
        }
        counts = (*step_)();
        ++iters;
        CELER_VALIDATE(iters < 2 * max_steps_ + 2,
                       << "tracks remained after killing active tracks");
    }
    last_iters_ = iters;
}

//---------------------------------------------------------------------------//
void LocalTransporter::Finalize()
{
    CELER_VALIDATE(buffer_.empty(), << "offloaded tracks (" << buffer_.size()
                                    << " in buffer) were not flushed");
}

//---------------------------------------------------------------------------//
std::map<std::string, double> LocalTransporter::GetActionTime() const
{
    std::map<std::string, double> result;
    auto const& actions = step_->actions();
    auto const& times = step_->action_times();
    for (std::size_t i = 0; i < actions.size(); ++i)
        result[std::string(actions[i]->label())] = times[i];
    return result;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
