//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file user/UserKernels.hh
//! Launchers of the user-action kernels (UserKernels.hip).
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime_api.h>

#include "global/CoreData.hh"
#include "UserData.hh"

namespace celeritas_amd
{
//! Post-step action of every valid track by particle (ActionDiagnostic)
hipError_t launch_action_tally(CoreStateRef const&, ParticleTallyRef const&, hipStream_t);
//! Step count of tracks killed this step by particle (StepDiagnostic)
hipError_t launch_step_tally(CoreStateRef const&, ParticleTallyRef const&, hipStream_t);
//! Pre-step point of the slots in a detector (StepGatherAction<pre>)
hipError_t launch_step_gather_pre(CoreParamsRef const&,
                                  CoreStateRef const&,
                                  StepCollectorRef const&,
                                  hipStream_t);
//! Compact this step's hits into the hit buffer (StepGatherAction<post> and
//! DetectorSteps' copy_steps)
hipError_t launch_step_gather_post(CoreParamsRef const&,
                                   CoreStateRef const&,
                                   StepCollectorRef const&,
                                   hipStream_t);
//! Empty the hit buffer and its overflow counters
hipError_t launch_step_hits_clear(StepHitCounters*, hipStream_t);
}  // namespace celeritas_amd
