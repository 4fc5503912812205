//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/Types.hh
//! Integer and enum types shared by host code, kernels and the C ABI.
//!
//! Enumerations keep the numeric values of the reference so that state dumps
//! compare directly: TrackStatus (celeritas/Types.hh:113-122), TrackOrder
//! (celeritas/Types.hh:151-171) and StepActionOrder
//! (corecel/sys/ActionInterface.hh:29-46).
//---------------------------------------------------------------------------//
#pragma once

#include <cstdint>

#include "Macros.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
using size_type = std::uint32_t;
using ull_int = unsigned long long;
using real_type = double;

//! Sentinel for every 32-bit identifier ("false" OpaqueId)
inline constexpr std::uint32_t invalid_id = 0xffffffffu;
//! Sentinel for 8-bit action/particle identifiers stored per slot
inline constexpr std::uint8_t invalid_id8 = 0xffu;

//---------------------------------------------------------------------------//
//! Status of a track slot
enum class TrackStatus : std::uint8_t
{
    inactive = 0,
    initializing = 1,
    alive = 2,
    errored = 3,
    killed = 4,
};

//! Ordering/remapping of track slots
enum class TrackOrder : std::uint8_t
{
    none = 0,
    init_charge = 1,
    reindex_shuffle = 2,
    reindex_status = 3,
    reindex_particle_type = 4,
    reindex_along_step_action = 5,
    reindex_step_limit_action = 6,
    reindex_both_action = 7,
};

//! Dependency ordering of actions inside a step
enum class StepActionOrder : std::uint8_t
{
    generate = 0,
    start,
    user_start,
    sort_start,
    pre,
    user_pre,
    sort_pre,
    along,
    sort_along,
    pre_post,
    sort_pre_post,
    post,
    user_post,
    end,
    size_
};

//! Floating point policy of the physics kernels
enum class PrecisionMode : std::uint8_t
{
    fp64 = 0,  //!< Everything in double precision (reference behavior)
    mixed = 1,  //!< FP32 sampling/XS/MSC; FP64 position, safety, edep
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
