//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/ScopedProfiling.cc
//---------------------------------------------------------------------------//
#include "ScopedProfiling.hh"

#include <cstdlib>
#include <cstring>

#include <rocprofiler-sdk-roctx/roctx.h>

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
bool use_profiling()
{
    static bool const enabled = [] {
        char const* v = std::getenv("CELER_ENABLE_PROFILING");
        return v && *v && std::strcmp(v, "0") != 0;
    }();
    return enabled;
}

//---------------------------------------------------------------------------//
ScopedProfiling::ScopedProfiling(char const* label) : active_(use_profiling())
{
    if (active_)
        roctxRangePushA(label);
}

ScopedProfiling::~ScopedProfiling()
{
    if (active_)
        roctxRangePop();
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
