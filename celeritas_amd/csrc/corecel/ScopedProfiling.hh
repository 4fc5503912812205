//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/ScopedProfiling.hh
//! Named profiler ranges around host-side work (roctx), so a trace from
//! rocprofv3 --marker-trace shows each step action's launches under its
//! label.
//!
//! Counterpart of the reference's ScopedProfiling
//! (corecel/sys/ScopedProfiling.hip.cc:29-56, .hh): ranges are recorded
//! only when profiling is enabled, here by CELER_ENABLE_PROFILING set to a
//! value other than "" or "0", read once.
//---------------------------------------------------------------------------//
#pragma once

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Whether profiling ranges are recorded
bool use_profiling();

//---------------------------------------------------------------------------//
//! RAII profiler range: pushed at construction, popped at destruction
class ScopedProfiling
{
  public:
    explicit ScopedProfiling(char const* label);
    ~ScopedProfiling();

    ScopedProfiling(ScopedProfiling const&) = delete;
    ScopedProfiling& operator=(ScopedProfiling const&) = delete;

  private:
    bool active_;
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
