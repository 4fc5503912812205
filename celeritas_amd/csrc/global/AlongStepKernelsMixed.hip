//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/AlongStepKernelsMixed.hip
//! Mixed-precision build of the charged along-step kernel (MSC, energy-loss
//! fluctuations; propagation and positions stay FP64).
//---------------------------------------------------------------------------//
#define CELER_MIXED_PRECISION 1
#include "AlongStepKernels.hip"
