//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file user/UserData.hh
//! Device views of the user actions: particle tallies (ActionDiagnostic,
//! StepDiagnostic) and step-point gathering for detectors (StepCollector,
//! DetectorSteps).
//!
//! Counterparts of user/ParticleTallyData.hh and user/StepData.hh. The step
//! data keep only what a step's end cannot recover: the pre-step point and
//! the detector of each slot. At the step's end the hits of the slots in a
//! detector are compacted in slot order (the order of the reference's
//! copy_if in DetectorSteps.cu:154-160) and appended to a device-resident
//! hit buffer, with the post-step fields read straight from the core state,
//! so no per-step copy to the host is needed; the host drains the buffer
//! when it wants the hits.
//---------------------------------------------------------------------------//
#pragma once

#include <cstdint>

#include "corecel/Types.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Histogram over (particle, bin) in one stream (ParticleTallyData.hh)
struct ParticleTallyRef
{
    std::uint32_t num_particles{0};
    std::uint32_t num_bins{0};
    unsigned long long* counts{nullptr};  //!< [particle][bin]
};

//! Bins of the LDS histograms of the tally kernels
inline constexpr unsigned int max_tally_bins = 4096;
//! Workgroups of a tally pass (each strides over the slots)
inline constexpr unsigned int tally_blocks = 512;

//---------------------------------------------------------------------------//
//! Track properties gathered at a step point (StepPointSelection)
struct StepPointSelection
{
    std::uint8_t time{0};
    std::uint8_t pos{0};
    std::uint8_t dir{0};
    std::uint8_t volume_id{0};
    std::uint8_t energy{0};
};

//! Track properties gathered every step (StepSelection, StepData.hh:64-120)
struct StepSelection
{
    StepPointSelection points[2];  //!< pre, post
    std::uint8_t event_id{0};
    std::uint8_t parent_id{0};
    std::uint8_t track_step_count{0};
    std::uint8_t action_id{0};
    std::uint8_t step_length{0};
    std::uint8_t particle{0};
    std::uint8_t energy_deposition{0};
};

enum StepPoint : int
{
    step_pre = 0,
    step_post = 1
};

//! Per-slot data of the pre-step point
struct StepPointSlots
{
    double* time{nullptr};
    double* pos[3]{nullptr, nullptr, nullptr};
    double* dir[3]{nullptr, nullptr, nullptr};
    std::uint32_t* volume_id{nullptr};
    double* energy{nullptr};
};

//! Compacted step-point columns of the hit buffer
struct StepPointHits
{
    double* time{nullptr};
    double* pos{nullptr};  //!< [hit][3]
    double* dir{nullptr};  //!< [hit][3]
    std::uint32_t* volume_id{nullptr};
    double* energy{nullptr};
};

/*!
 * Hit buffer: one entry per step of a track in a detector (DetectorStepOutput,
 * DetectorSteps.hh:47-80), columns allocated only for selected fields.
 */
struct StepHitsRef
{
    std::uint32_t capacity{0};
    std::uint32_t* detector{nullptr};
    std::uint32_t* track_id{nullptr};
    StepPointHits points[2];
    std::uint32_t* event_id{nullptr};
    std::uint32_t* parent_id{nullptr};
    std::uint32_t* track_step_count{nullptr};
    std::uint32_t* action_id{nullptr};
    double* step_length{nullptr};
    std::uint32_t* particle{nullptr};
    double* energy_deposition{nullptr};
};

//! Device counters of the collector
struct StepHitCounters
{
    std::uint32_t num_hits;  //!< entries in the hit buffer
    std::uint32_t step_hits;  //!< hits of the current step
    std::uint32_t overflow;  //!< steps whose hits did not fit
    std::uint32_t dropped;  //!< hits lost to overflow
};

/*!
 * Shared and per-stream data of the collector (StepParamsData and
 * StepStateData, StepData.hh:124-340).
 */
struct StepCollectorRef
{
    StepSelection selection;
    //! Volume -> detector (invalid_id: none); null collects every active
    //! track
    std::uint32_t const* detector_map{nullptr};
    //! Skip post-step points without energy deposition
    std::uint32_t nonzero_energy_deposition{0};
    //! Per-slot detector of the current step (invalid_id: not collected)
    std::uint32_t* slot_detector{nullptr};
    StepPointSlots pre;
    //! Per-workgroup counts and exclusive offsets of the compaction
    std::uint32_t* block_counts{nullptr};
    StepHitCounters* counters{nullptr};
    StepHitsRef hits;
};

//! Workgroup size of the compaction kernels
inline constexpr unsigned int step_gather_block = 256;

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
