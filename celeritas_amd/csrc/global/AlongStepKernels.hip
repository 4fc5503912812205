//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/AlongStepKernels.hip
//! Along-step action for charged particles (linear or in a uniform field).
//!
//! One thread per slot runs the charged along-step body
//! (AlongStepExecutor.hh) for the tracks of this action.
//---------------------------------------------------------------------------//
#include "corecel/Precision.hh"
#include "CoreActions.hh"
#include "corecel/KernelLaunch.hh"
#include "corecel/WaveOps.hh"
#include "AlongStepExecutor.hh"
#include "StagedTables.hh"

namespace celeritas_amd
{
namespace
{
constexpr int block_size = 256;

template<bool UseField, Staging S>
__global__ void __launch_bounds__(block_size)
    along_step_charged_kernel(CoreParamsRef const params_in,
                              CoreStateRef const state,
                              std::uint8_t const action,
                              ThreadSlots const ts)
{
    extern __shared__ uint4 staged_tables[];
    if (!workgroup_has_work(ts, state.size))
        return;
    CoreParamsRef const params = stage<S>(params_in, staged_tables);
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    auto const status = static_cast<TrackStatus>(state.sim.status[slot]);
    if (status == TrackStatus::inactive || state.sim.along_step_action[slot] != action)
        return;
    if (state.diag)
        wave::count_active_lanes_spread<diag_spread>(state.diag + diag_along_charged * diag_spread);
    exec::along_step_charged_track<UseField>(params, state, slot, status);
    fused_discrete_select<true>(params, state, slot);
}

//---------------------------------------------------------------------------//
}  // namespace

hipError_t CELER_PREC_FN(launch_along_step_charged)(CoreParamsRef const& p,
                                     CoreStateRef const& s,
                                     std::uint8_t action,
                                     ThreadSlots const& ts,
                                     hipStream_t st)
{
    // Physics and geometry tables in LDS where both fit, else the physics
    unsigned int lds = staged_both_bytes(p);
    Staging const staging = lds ? Staging::both
                                : (staged_table_bytes(p) ? Staging::physics : Staging::none);
    if (staging == Staging::physics)
        lds = staged_table_bytes(p);
    auto pick = [staging](auto both, auto physics, auto none) {
        return staging == Staging::both ? both : staging == Staging::physics ? physics : none;
    };
    auto* kernel = p.field.enabled ? pick(along_step_charged_kernel<true, Staging::both>,
                                          along_step_charged_kernel<true, Staging::physics>,
                                          along_step_charged_kernel<true, Staging::none>)
                                   : pick(along_step_charged_kernel<false, Staging::both>,
                                          along_step_charged_kernel<false, Staging::physics>,
                                          along_step_charged_kernel<false, Staging::none>);
#if CELER_PREC_IS_MIXED
    char const* name = p.field.enabled ? "along-step-field-mixed" : "along-step-linear-mixed";
#else
    char const* name = p.field.enabled ? "along-step-field" : "along-step-linear";
#endif
    launch_kernel(name, kernel, grid_of(s), s.block_size, lds, st, p, s, action, ts);
    return hipGetLastError();
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
