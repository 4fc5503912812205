//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/StepKernels.hip
//! Per-action kernels of one step iteration (one thread per track slot).
//!
//! Each kernel mirrors one reference action (the per-track bodies are in
//! StepExecutors.hh) plus the StepGather + SimpleCalo tally. Kernels
//! early-exit for slots whose action does not match, exactly like the
//! reference's ConditionalTrackExecutor.
//---------------------------------------------------------------------------//
#include <algorithm>

#include "corecel/Precision.hh"
#include "CoreActions.hh"

#include "corecel/KernelLaunch.hh"
#include "corecel/WaveOps.hh"
#include "StagedTables.hh"
#include "StepExecutors.hh"

namespace celeritas_amd
{
namespace
{
//---------------------------------------------------------------------------//
constexpr int block_size = 256;
constexpr int num_waves = block_size / wave::size;

inline unsigned int num_blocks(size_type n)
{
    return (n + block_size - 1) / block_size;
}

using exec::material_of;
using exec::status_of;

//---------------------------------------------------------------------------//
// PRE-STEP
//---------------------------------------------------------------------------//
template<bool Staged>
__global__ void __launch_bounds__(block_size)
    pre_step_kernel(CoreParamsRef const params_in,
                    CoreStateRef const state,
                    ThreadSlots const ts)
{
    extern __shared__ uint4 staged_tables[];
    size_type const tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0)
        state.counters->secondary_stack_size = 0;
    if (!workgroup_has_work(ts, state.size))
        return;
    CoreParamsRef const params = stage_tables<Staged>(params_in, staged_tables);
    size_type slot;
    if (!thread_to_slot(ts, tid, state.size, &slot))
        return;
    auto const status = status_of(state, slot);
    if (status == TrackStatus::inactive)
    {
        exec::pre_step_inactive(state, slot);
        return;
    }
    if (state.diag)
        wave::count_active_lanes_spread<diag_spread>(state.diag + diag_pre_step * diag_spread);
    exec::pre_step_track(params, state, slot, status);
}

//---------------------------------------------------------------------------//
// ALONG-STEP (NEUTRAL)
//---------------------------------------------------------------------------//
template<Staging S>
__global__ void __launch_bounds__(block_size)
    along_step_neutral_kernel(CoreParamsRef const params_in,
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
    auto const status = status_of(state, slot);
    if (status == TrackStatus::inactive || state.sim.along_step_action[slot] != action)
        return;
    if (state.diag)
        wave::count_active_lanes_spread<diag_spread>(state.diag + diag_along_neutral * diag_spread);
    exec::along_step_neutral_track(params, state, slot, status);
    fused_discrete_select<false>(params, state, slot);
}

//---------------------------------------------------------------------------//
// DISCRETE SELECT
//---------------------------------------------------------------------------//
template<bool Staged>
__global__ void __launch_bounds__(block_size)
    discrete_select_kernel(CoreParamsRef const params_in,
                           CoreStateRef const state,
                           ThreadSlots const ts)
{
    extern __shared__ uint4 staged_tables[];
    if (!workgroup_has_work(ts, state.size))
        return;
    CoreParamsRef const params = stage_tables<Staged>(params_in, staged_tables);
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    if (state.sim.post_step_action[slot] != params.phys.scalars.discrete_action
        || status_of(state, slot) == TrackStatus::inactive)
        return;
    if (state.diag)
        wave::count_active_lanes_spread<diag_spread>(state.diag + diag_discrete * diag_spread);
    exec::discrete_select_track(params, state, slot);
}

//---------------------------------------------------------------------------//
// INTERACTIONS
//---------------------------------------------------------------------------//
//! One instantiation per model class: each launch carries only the
//! registers of its own sampler
template<ModelClass M>
__global__ void __launch_bounds__(block_size)
    interaction_kernel(CoreParamsRef const params,
                       CoreStateRef const state,
                       std::uint8_t const action,
                       ThreadSlots const ts)
{
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    if (state.sim.post_step_action[slot] != action
        || status_of(state, slot) == TrackStatus::inactive)
        return;
    if (state.diag)
        wave::count_active_lanes_spread<diag_spread>(state.diag + diag_interaction * diag_spread);
    exec::interaction_track<M>(params, state, slot);
}

/*!
 * All discrete models in one launch: each thread samples the model its
 * track selected. Over the union of the models' sorted ranges consecutive
 * threads share a model, so waves stay uniform, and the short latency-
 * bound runs of the separate per-model launches overlap instead of
 * following one another.
 */
template<bool Staged>
__global__ void __launch_bounds__(block_size)
    interactions_kernel(CoreParamsRef const params_in,
                        CoreStateRef const state,
                        ThreadSlots const ts)
{
    extern __shared__ uint4 staged_tables[];
    if (!workgroup_has_work(ts, state.size))
        return;
    CoreParamsRef const params = stage_tables<Staged>(params_in, staged_tables);
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    if (status_of(state, slot) == TrackStatus::inactive)
        return;
    std::uint32_t const model
        = std::uint32_t(state.sim.post_step_action[slot])
          - params.phys.scalars.first_model_action;
    if (model >= params.phys.scalars.num_models)
        return;
    if (state.diag)
        wave::count_active_lanes_spread<diag_spread>(state.diag + diag_interaction * diag_spread);
    switch (static_cast<ModelClass>(params.phys.model_class[model]))
    {
#define CELER_MODEL_CASE(M)                                  \
    case ModelClass::M:                                      \
        exec::interaction_track<ModelClass::M>(params, state, slot); \
        break;
        CELER_MODEL_CASE(klein_nishina)
        CELER_MODEL_CASE(livermore_pe)
        CELER_MODEL_CASE(bethe_heitler)
        CELER_MODEL_CASE(moller_bhabha)
        CELER_MODEL_CASE(seltzer_berger)
        CELER_MODEL_CASE(relativistic_brem)
        CELER_MODEL_CASE(eplusgg)
        CELER_MODEL_CASE(combined_brem)
#undef CELER_MODEL_CASE
        default:
            break;
    }
}

//---------------------------------------------------------------------------//
// BOUNDARY AND TRACKING CUT
//---------------------------------------------------------------------------//
template<Staging S>
__global__ void __launch_bounds__(block_size)
    boundary_kernel(CoreParamsRef const params_in,
                    CoreStateRef const state,
                    ThreadSlots const ts)
{
    extern __shared__ uint4 staged_tables[];
    if (!workgroup_has_work(ts, state.size))
        return;
    CoreParamsRef const params = stage<S>(params_in, staged_tables);
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    if (state.sim.post_step_action[slot] != params.scalars.boundary_action
        || status_of(state, slot) != TrackStatus::alive)
        return;
    if (state.diag)
        wave::count_active_lanes_spread<diag_spread>(state.diag + diag_boundary * diag_spread);
    exec::boundary_track(params, state, slot);
}

__global__ void __launch_bounds__(block_size)
    tracking_cut_kernel(CoreParamsRef const params,
                        CoreStateRef const state,
                        ThreadSlots const ts)
{
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    if (state.sim.post_step_action[slot] != params.scalars.tracking_cut_action
        || status_of(state, slot) == TrackStatus::inactive)
        return;
    exec::tracking_cut_track(params, state, slot);
}

//! Kill all active tracks (global/detail/KillActive.cu:25-32)
__global__ void __launch_bounds__(block_size)
    kill_active_kernel(CoreParamsRef const params, CoreStateRef const state)
{
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    if (slot >= state.size)
        return;
    auto const st = status_of(state, slot);
    if (st == TrackStatus::inactive || st == TrackStatus::killed)
        return;
    apply_errored(params, state, slot);
}

//---------------------------------------------------------------------------//
// CALORIMETER TALLY (the SimpleCalo user action's step callback)
//
// First pass: at most calo_tally_blocks workgroups stride over the slots
// and sum their deposits into LDS bins, then write the bins to their
// column of calo_partial ([bin][workgroup]). Second pass: one workgroup per
// bin adds its row of partials to the tally. Compared with one global atomic per bin per workgroup of 256
// slots, the same-address atomics that serialize in L2 are gone.
//---------------------------------------------------------------------------//
//! Detector bin and deposit of a slot (false if nothing to tally)
__device__ __forceinline__ bool calo_deposit(CoreParamsRef const& params,
                                             CoreStateRef const& state,
                                             size_type slot,
                                             std::uint32_t* det,
                                             double* edep)
{
    if (status_of(state, slot) == TrackStatus::inactive)
        return false;
    *edep = state.phys.energy_deposition[slot];
    if (*edep == 0)
        return false;
    *det = params.volume_to_detector[state.pre_volume[slot]];
    return *det != invalid_id;
}

__global__ void __launch_bounds__(block_size)
    calo_partial_kernel(CoreParamsRef const params,
                        CoreStateRef const state,
                        size_type const num_det,
                        ThreadSlots const ts)
{
    __shared__ double bins[max_lds_detectors];
    for (size_type i = threadIdx.x; i < num_det; i += block_size)
        bins[i] = 0;
    __syncthreads();
    size_type const stride = gridDim.x * block_size;
    size_type slot;
    for (size_type tid = blockIdx.x * block_size + threadIdx.x;
         thread_to_slot(ts, tid, state.size, &slot);
         tid += stride)
    {
        std::uint32_t det;
        double edep;
        if (calo_deposit(params, state, slot, &det, &edep))
            atomicAdd(&bins[det], edep);
    }
    __syncthreads();
    // Bin-major: the second pass reads each bin's partials contiguously
    for (size_type i = threadIdx.x; i < num_det; i += block_size)
        state.calo_partial[std::size_t(i) * gridDim.x + blockIdx.x] = bins[i];
}

//! Sum of the partial rows of one bin (workgroup blockIdx.x = bin)
__global__ void __launch_bounds__(block_size)
    calo_reduce_kernel(CoreStateRef const state,
                       size_type const num_det,
                       unsigned int const num_rows)
{
    __shared__ double lds[num_waves];
    size_type const det = blockIdx.x;
    double sum = 0;
    double const* column = state.calo_partial + std::size_t(det) * num_rows;
    for (unsigned int r = threadIdx.x; r < num_rows; r += block_size)
        sum += column[r];
    sum = wave::reduce_sum(sum);
    if (wave::lane_id() == 0)
        lds[threadIdx.x / wave::size] = sum;
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double total = 0;
        for (int w = 0; w < num_waves; ++w)
            total += lds[w];
        state.calo_edep[det] += total;
    }
}

//! Direct global atomics, for more bins than LDS holds
__global__ void __launch_bounds__(block_size)
    calo_atomic_kernel(CoreParamsRef const params,
                       CoreStateRef const state,
                       ThreadSlots const ts)
{
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    std::uint32_t det;
    double edep;
    if (calo_deposit(params, state, slot, &det, &edep))
        atomicAdd(&state.calo_edep[det], edep);
}

//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
// LAUNCHERS
//---------------------------------------------------------------------------//
#if CELER_PREC_IS_MIXED
#    define CELER_KERNEL_NAME(S) S "-mixed"
#else
#    define CELER_KERNEL_NAME(S) S
#endif

hipError_t CELER_PREC_FN(launch_pre_step)(CoreParamsRef const& p,
                           CoreStateRef const& s,
                           ThreadSlots const& ts,
                           hipStream_t st)
{
    unsigned int const lds = staged_table_bytes(p);
    launch_kernel(CELER_KERNEL_NAME("pre-step"), lds ? pre_step_kernel<true> : pre_step_kernel<false>,
           grid_of(s), s.block_size, lds, st, p, s, ts);
    return hipGetLastError();
}

#if !CELER_PREC_IS_MIXED
hipError_t launch_along_step_neutral(CoreParamsRef const& p,
                                     CoreStateRef const& s,
                                     std::uint8_t action,
                                     ThreadSlots const& ts,
                                     hipStream_t st)
{
    unsigned int const lds = staged_geo_bytes(p);
    launch_kernel("along-step-neutral",
                  lds ? along_step_neutral_kernel<Staging::geometry>
                      : along_step_neutral_kernel<Staging::none>,
                  grid_of(s), s.block_size, lds, st, p, s, action, ts);
    return hipGetLastError();
}

#endif

hipError_t CELER_PREC_FN(launch_discrete_select)(CoreParamsRef const& p,
                                  CoreStateRef const& s,
                                  ThreadSlots const& ts,
                                  hipStream_t st)
{
    unsigned int const lds = staged_table_bytes(p);
    launch_kernel(CELER_KERNEL_NAME("discrete-select"),
           lds ? discrete_select_kernel<true> : discrete_select_kernel<false>, grid_of(s),
           s.block_size, lds, st, p, s, ts);
    return hipGetLastError();
}

hipError_t CELER_PREC_FN(launch_interaction)(CoreParamsRef const& p,
                              CoreStateRef const& s,
                              std::uint8_t action,
                              std::uint32_t model,
                              ThreadSlots const& ts,
                              hipStream_t st)
{
    switch (static_cast<ModelClass>(model))
    {
#define CELER_LAUNCH_MODEL(M)                                                               \
    case ModelClass::M:                                                                     \
        launch_kernel(CELER_KERNEL_NAME("interaction-" #M), interaction_kernel<ModelClass::M>,    \
               grid_of(s), s.block_size, 0, st, p, s, action, ts);                          \
        break;
        CELER_LAUNCH_MODEL(klein_nishina)
        CELER_LAUNCH_MODEL(livermore_pe)
        CELER_LAUNCH_MODEL(bethe_heitler)
        CELER_LAUNCH_MODEL(moller_bhabha)
        CELER_LAUNCH_MODEL(seltzer_berger)
        CELER_LAUNCH_MODEL(relativistic_brem)
        CELER_LAUNCH_MODEL(eplusgg)
        CELER_LAUNCH_MODEL(combined_brem)
#undef CELER_LAUNCH_MODEL
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t CELER_PREC_FN(launch_interactions)(CoreParamsRef const& p,
                               CoreStateRef const& s,
                               ThreadSlots const& ts,
                               hipStream_t st)
{
    unsigned int const lds = staged_table_bytes(p);
    launch_kernel(CELER_KERNEL_NAME("interactions"),
           lds ? interactions_kernel<true> : interactions_kernel<false>, grid_of(s), s.block_size,
           lds, st, p, s, ts);
    return hipGetLastError();
}

#if !CELER_PREC_IS_MIXED
hipError_t launch_boundary(CoreParamsRef const& p,
                           CoreStateRef const& s,
                           ThreadSlots const& ts,
                           hipStream_t st)
{
    unsigned int const lds = staged_geo_bytes(p);
    launch_kernel("boundary",
                  lds ? boundary_kernel<Staging::geometry> : boundary_kernel<Staging::none>,
                  grid_of(s), s.block_size, lds, st, p, s, ts);
    return hipGetLastError();
}

hipError_t launch_tracking_cut(CoreParamsRef const& p,
                               CoreStateRef const& s,
                               ThreadSlots const& ts,
                               hipStream_t st)
{
    launch_kernel("tracking-cut", tracking_cut_kernel, grid_of(s), s.block_size, 0, st, p, s, ts);
    return hipGetLastError();
}

hipError_t launch_kill_active(CoreParamsRef const& p, CoreStateRef const& s, hipStream_t st)
{
    kill_active_kernel<<<num_blocks(s.size), block_size, 0, st>>>(p, s);
    return hipGetLastError();
}

hipError_t launch_calo_tally(CoreParamsRef const& p,
                             CoreStateRef const& s,
                             size_type num_det,
                             ThreadSlots const& ts,
                             hipStream_t st)
{
    if (num_det == 0)
        return hipSuccess;
    if (num_det > max_lds_detectors)
    {
        launch_kernel("calo-atomic", calo_atomic_kernel, num_blocks(s.size), block_size, 0, st, p, s, ts);
        return hipGetLastError();
    }
    unsigned int const rows = std::min<unsigned int>(num_blocks(s.size), calo_tally_blocks);
    launch_kernel("calo-partial", calo_partial_kernel, rows, block_size, 0, st, p, s, num_det, ts);
    launch_kernel("calo-reduce", calo_reduce_kernel, num_det, block_size, 0, st, s, num_det, rows);
    return hipGetLastError();
}

#endif

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
