//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/StepKernelsMixed.hip
//! Mixed-precision build of the pre-step, discrete-selection and
//! interaction kernels (the _mixed launchers).
//---------------------------------------------------------------------------//
#define CELER_MIXED_PRECISION 1
#include "StepKernels.hip"
