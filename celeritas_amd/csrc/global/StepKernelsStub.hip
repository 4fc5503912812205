//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: temporary launchers for actions not yet on the device
//---------------------------------------------------------------------------//
#include "CoreActions.hh"

namespace celeritas_amd
{
hipError_t launch_pre_step(CoreParamsRef const&, CoreStateRef const&, hipStream_t)
{ return hipErrorNotSupported; }
hipError_t launch_along_step_neutral(CoreParamsRef const&, CoreStateRef const&, std::uint8_t, hipStream_t)
{ return hipErrorNotSupported; }
hipError_t launch_along_step_charged(CoreParamsRef const&, CoreStateRef const&, std::uint8_t, hipStream_t)
{ return hipErrorNotSupported; }
hipError_t launch_discrete_select(CoreParamsRef const&, CoreStateRef const&, hipStream_t)
{ return hipErrorNotSupported; }
hipError_t launch_interactions(CoreParamsRef const&, CoreStateRef const&, hipStream_t)
{ return hipErrorNotSupported; }
hipError_t launch_boundary(CoreParamsRef const&, CoreStateRef const&, hipStream_t)
{ return hipErrorNotSupported; }
hipError_t launch_tracking_cut(CoreParamsRef const&, CoreStateRef const&, hipStream_t)
{ return hipErrorNotSupported; }
hipError_t launch_calo_tally(CoreParamsRef const&, CoreStateRef const&, hipStream_t)
{ return hipErrorNotSupported; }
hipError_t launch_fused_step(CoreParamsRef const&, CoreStateRef const&, hipStream_t)
{ return hipErrorNotSupported; }
}
namespace celeritas_amd
{
hipError_t launch_kill_active(CoreParamsRef const&, CoreStateRef const&, hipStream_t)
{ return hipErrorNotSupported; }
}
