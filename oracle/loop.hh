//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/loop.hh
//! Action dispatch and the step loop (celeritas/global/Stepper.cc:105-168,
//! celeritas/global/ActionSequence.cc:78-138).
//---------------------------------------------------------------------------//
#pragma once

#include <string>

#include "core.hh"

namespace orc
{
//! Run one action by label on all slots
void run_action(State& s, std::string const& label);
//---------------------------------------------------------------------------//
}  // namespace orc
