//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file track/TrackKernels.hh
//! Host-callable launchers for track management and RNG kernels.
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime_api.h>

#include "global/CoreData.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! \param ranks null, or 2*count words: event-local rank, then last flag
hipError_t launch_extend_from_primaries(CoreStateRef const& state,
                                        Primary const* primaries,
                                        std::uint32_t const* ranks,
                                        size_type count,
                                        hipStream_t stream);
hipError_t launch_initialize_tracks(CoreParamsRef const& params,
                                    CoreStateRef const& state,
                                    hipStream_t stream);
hipError_t launch_extend_from_secondaries(CoreParamsRef const& params,
                                          CoreStateRef const& state,
                                          hipStream_t stream);
hipError_t launch_fill_u32(std::uint32_t* data,
                           size_type n,
                           std::uint32_t value,
                           hipStream_t stream);
hipError_t launch_iota_u32(std::uint32_t* data, size_type n, hipStream_t stream);
size_type extend_block_scratch_size(size_type num_slots);

//---------------------------------------------------------------------------//
// RNG
hipError_t launch_reseed_rng(RngStateRef const& state,
                             XorwowJumpTables const* tables,
                             std::uint32_t seed,
                             ull_int event_id,
                             hipStream_t stream);
hipError_t launch_rng_draw(RngStateRef const& state,
                           size_type draws_per_slot,
                           int kind,
                           void* out,
                           hipStream_t stream);
hipError_t launch_rng_initialize(RngStateRef const& state,
                                 XorwowJumpTables const* tables,
                                 std::uint32_t seed,
                                 ull_int const* subsequence,
                                 ull_int const* offset,
                                 hipStream_t stream);

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
