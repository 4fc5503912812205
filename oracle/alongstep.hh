//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/alongstep.hh
//! Charged along-step (Urban MSC, linear propagation, Urban fluctuations).
//---------------------------------------------------------------------------//
#pragma once

#include "core.hh"
#include "physics.hh"

namespace orc
{
//! Run the general-linear along-step action on all matching tracks
void along_step_charged_impl(State& s);

//! Urban fluctuation sample (EnergyLossUrbanDistribution) on an engine
template<class U>
double urban_fluct_sample(FluctO const& par, MaterialDef const& mat, double unscaled_mean,
                          double max_energy, double two_mebsgs, double beta_sq, U& u);
}  // namespace orc
