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
                              hipStream_t stream,
                              std::uint8_t const* along_of_particle = nullptr);

/*!
 * Stable multi-split of the indices 0..n-1 by a byte key per index (keys
 * at or above num_keys - 1 take the last key): sorted lists, for each key,
 * its indices in increasing order, and offsets[0..num_keys] its thread
 * ranges. counts is scratch of sort_counts_size(n) words.
 */
hipError_t launch_partition_keys(std::uint8_t const* keys,
                                 size_type n,
                                 unsigned int num_keys,
                                 std::uint32_t* counts,
                                 std::uint32_t* sorted,
                                 std::uint32_t* offsets,
                                 hipStream_t stream);

//! Words of histogram scratch needed to split n indices
size_type sort_counts_size(size_type num_slots);

//! Largest n of the allocation-free partition (its scratch is static)
inline constexpr size_type max_partition_noalloc = size_type(1) << 22;

/*!
 * Device address, on the current device, of the static scratch of the
 * allocation-free partition (looked up once per device). Calls that
 * overlap on one device share it and race.
 */
hipError_t partition_noalloc_scratch(std::uint32_t** ptr);

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
