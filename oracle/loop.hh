//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/loop.hh
//! Action dispatch and the step loop (celeritas/global/Stepper.cc:105-168,
//! celeritas/global/ActionSequence.cc:78-138): actions run in (order, ID)
//! order, each over all slots in slot order.
//---------------------------------------------------------------------------//
#pragma once

#include <string>
#include <vector>

#include "core.hh"

namespace orc
{
//! Run one action by label on all slots
void run_action(State& s, std::string const& label);

//! One step iteration; returns {generated, queued, active, alive}
struct StepResult
{
    std::uint32_t generated, queued, active, alive;
};
StepResult step_once(State& s);
//---------------------------------------------------------------------------//
}  // namespace orc
