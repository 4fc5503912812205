//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/user.hh
//! User observers of the step loop: ActionDiagnostic
//! (user/detail/ActionDiagnosticExecutor.hh), StepDiagnostic
//! (user/detail/StepDiagnosticExecutor.hh) and the detector step gather
//! (user/detail/StepGatherExecutor.hh:41-153 with DetectorSteps.cc's host
//! copy_steps: hits of the valid slots in slot order).
//!
//! Pinned by: Diagnostic.test.cc:110-131 (SimpleCompton action keys, counts
//! and step histogram) and StepCollector.test.cc:179-226 (KN MC-truth
//! pre-step points after one and two steps); see tests/test_oracle_user.py.
//---------------------------------------------------------------------------//
#pragma once

#include <cstdint>
#include <vector>

#include "core.hh"

namespace orc
{
//---------------------------------------------------------------------------//
//! Every field of one detector step (the selection picks columns later)
struct StepHit
{
    std::uint32_t detector = none, track_id = none, event_id = none, parent_id = none;
    std::uint32_t track_step_count = 0, action_id = none, particle = none;
    std::uint32_t volume[2] = {none, none};  // pre, post
    double step_length = 0, edep = 0;
    double time[2] = {0, 0};
    double pos[2][3] = {{0, 0, 0}, {0, 0, 0}};
    double dir[2][3] = {{0, 0, 0}, {0, 0, 0}};
    double energy[2] = {0, 0};
};

struct UserObservers
{
    // ActionDiagnostic: [particle][action]
    bool action_diag = false;
    std::uint32_t num_actions = 0;
    std::vector<std::uint64_t> action_counts;
    // StepDiagnostic: [particle][min(steps, bins - 1)]
    bool step_diag = false;
    std::uint32_t step_bins = 0;
    std::vector<std::uint64_t> step_counts;
    // Step collector
    bool collect = false;
    std::vector<std::uint32_t> detector_map;  // empty: every active track
    bool nonzero_edep = false;
    std::vector<std::uint32_t> slot_detector;
    std::vector<StepHit> pre;  // per slot: the pre-step point
    std::vector<StepHit> hits;
};

//! Planted mistakes for the pin tests (0: none). 1: the action tally skips
//! killed tracks; 2: the step tally bins one step low; 3: the pre-step
//! position is taken at the end of the step
extern int g_user_mutation;

//! Pre-step gather (user_pre order)
void user_gather_pre(State& s);
//! Post-step action tally (post order, after the post-step actions)
void user_action_diag(State& s);
//! Step-count tally and post-step gather (user_post order)
void user_post(State& s);

//---------------------------------------------------------------------------//
}  // namespace orc
