//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file track/SortKernels.hh
//! Host-callable launcher of the track-slot reindexing.
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime_api.h>

#include "global/CoreData.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
/*!
 * Sort track_slots stably by a per-slot key and write the per-key thread
 * offsets to action_offsets[0..num_keys] (the last entry is the state
 * size). Inactive slots take key num_keys - 1.
 */
hipError_t launch_sort_tracks(CoreStateRef const& state,
                              SortKey kind,
                              unsigned int num_keys,
                              hipStream_t stream);

//! Words of sort_counts scratch needed for a state size
size_type sort_counts_size(size_type num_slots);

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
