//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/StepKernels.hip
//! Per-action kernels of one step iteration (one thread per track slot).
//!
//! Each kernel mirrors one reference action: PreStepExecutor
//! (phys/detail/PreStepExecutor.hh:46-115), the neutral along-step
//! (global/alongstep/AlongStep.hh:38-48 with PropagationApplier,
//! TimeUpdater, TrackUpdater), DiscreteSelectExecutor, InteractionApplier
//! (phys/InteractionApplier.hh:106-171), BoundaryExecutor,
//! TrackingCutExecutor and the StepGather + SimpleCalo tally. Kernels
//! early-exit for slots whose action does not match, exactly like the
//! reference's ConditionalTrackExecutor.
//---------------------------------------------------------------------------//
#include <algorithm>

#include "corecel/Precision.hh"
#include "CoreActions.hh"

#include "corecel/WaveOps.hh"
#include "em/Interactors.hh"
#include "geo/OrangeLiteTracker.hh"
#include "phys/PhysicsImpl.hh"
#include "track/TrackInitImpl.hh"

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

__device__ __forceinline__ TrackStatus status_of(CoreStateRef const& s, size_type i)
{
    return static_cast<TrackStatus>(s.sim.status[i]);
}

__device__ __forceinline__ std::uint32_t material_of(CoreParamsRef const& p,
                                                     CoreStateRef const& s,
                                                     size_type i)
{
    return p.volume_to_material[s.geo.volume[i]];
}

//---------------------------------------------------------------------------//
// PRE-STEP
//---------------------------------------------------------------------------//
__global__ void __launch_bounds__(block_size)
    pre_step_kernel(CoreParamsRef const params,
                    CoreStateRef const state,
                    ThreadSlots const ts)
{
    size_type const tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid == 0)
        state.counters->secondary_stack_size = 0;
    size_type slot;
    if (!thread_to_slot(ts, tid, state.size, &slot))
        return;
    auto const status = status_of(state, slot);
    if (status == TrackStatus::inactive)
    {
        state.sim.step_length[slot] = HUGE_VAL;
        state.sim.post_step_action[slot] = invalid_id8;
        state.sim.along_step_action[slot] = invalid_id8;
        return;
    }
    if (state.diag)
        wave::count_active_lanes(state.diag + diag_pre_step);
    PhysicsStateRef const& ps = state.phys;
    ps.energy_deposition[slot] = 0;
    ps.sec_start[slot] = 0;
    ps.sec_count[slot] = 0;
    ps.element[slot] = invalid_id8;
    state.pre_volume[slot] = state.geo.volume[slot];
    if (status == TrackStatus::errored)
        return;
    state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::alive);

    PhysicsParamsRef const& pp = params.phys;
    if (pp.num_particles == 0)
        return;
    std::uint32_t const pid = state.particle.particle_id[slot];
    ParticlePhysics const& part = pp.particle[pid];
    double mfp = ps.interaction_mfp[slot];
    if (!(mfp > 0))
    {
        XorwowEngine rng = XorwowEngine::load(state.rng, slot);
        mfp = -mth::log(draw_canonical(rng));
        rng.store(state.rng, slot);
        ps.interaction_mfp[slot] = mfp;
    }
    std::uint32_t const mat = material_of(params, state, slot);
    double const energy = state.particle.energy[slot];
    double total = 0;
    // One log and grid bracket of the energy, shared by the processes
    phys::GridPoint gp(energy);
    for (std::uint32_t i = 0; i < part.num_processes; ++i)
    {
        ProcessRecord const& proc = part.process[i];
        double const xs = proc.integral ? phys::calc_max_xs(pp, proc, mat, gp)
                                        : phys::calc_xs(pp, proc, mat, gp);
        total += xs;
        ps.per_process_xs[i * state.size + slot] = xs;
    }
    ps.macro_xs[slot] = total;
    double step;
    std::uint8_t action = pp.scalars.discrete_action;
    if (energy == 0)
    {
        step = 0;
    }
    else
    {
        step = mfp / total;
        if (part.eloss_process >= 0)
        {
            double const range
                = phys::calc_range(pp.grids[part.range_grid + mat], pp.reals, energy);
            ps.dedx_range[slot] = range;
            double const eloss_step = phys::range_to_step(pp.scalars, range);
            if (eloss_step <= step)
            {
                step = eloss_step;
                action = pp.scalars.range_action;
            }
            double const fixed = pp.scalars.fixed_step_limiter;
            if (fixed > 0 && fixed < step)
            {
                step = fixed;
                action = pp.scalars.fixed_step_action;
            }
        }
        else if (part.num_processes == 0)
        {
            action = invalid_id8;
        }
    }
    state.sim.step_length[slot] = step;
    state.sim.post_step_action[slot] = action;
    state.sim.along_step_action[slot] = (pp.charge[pid] == 0)
                                            ? params.scalars.along_step_neutral_action
                                            : params.scalars.along_step_user_action;
}

//---------------------------------------------------------------------------//
// ALONG-STEP (NEUTRAL)
//---------------------------------------------------------------------------//
__global__ void __launch_bounds__(block_size)
    along_step_neutral_kernel(CoreParamsRef const params,
                              CoreStateRef const state,
                              std::uint8_t const action,
                              ThreadSlots const ts)
{
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    auto const status = status_of(state, slot);
    if (status == TrackStatus::inactive || state.sim.along_step_action[slot] != action)
        return;
    if (state.diag)
        wave::count_active_lanes(state.diag + diag_along_neutral);

    double step = state.sim.step_length[slot];
    std::uint8_t post = state.sim.post_step_action[slot];
    if (step != 0)
    {
        geo::GeoTrack g = geo::GeoTrack::load(state.geo, slot);
        geo::Propagation prop = geo::find_next_step(params.geo, g, step);
        if (prop.boundary)
            geo::move_to_boundary(g);
        else
            geo::move_internal(g, step);
        g.store(state.geo, slot);
        if (CELER_UNLIKELY(!std::isfinite(g.pos[0] + g.pos[1] + g.pos[2])))
        {
            // Not reachable from a finite FP64 state
            apply_errored(params, state, slot);
            return;
        }
        if (prop.boundary)
        {
            step = prop.distance;
            post = params.scalars.boundary_action;
        }
        else if (prop.distance < step)
        {
            step = prop.distance;
            post = params.scalars.propagation_limit_action;
        }
        state.sim.step_length[slot] = step;
        state.sim.post_step_action[slot] = post;
    }
    if (status == TrackStatus::errored)
        return;
    std::uint32_t const pid = state.particle.particle_id[slot];
    double const v = phys::speed(state.particle.energy[slot], params.phys.mass[pid]);
    if (v > 0)
        state.sim.time[slot] += step / v;
    if (post != params.phys.scalars.discrete_action)
    {
        state.phys.interaction_mfp[slot] -= step * state.phys.macro_xs[slot];
    }
    state.sim.num_steps[slot] += 1;
}

//---------------------------------------------------------------------------//
// DISCRETE SELECT
//---------------------------------------------------------------------------//
__global__ void __launch_bounds__(block_size)
    discrete_select_kernel(CoreParamsRef const params,
                           CoreStateRef const state,
                           ThreadSlots const ts)
{
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    PhysicsParamsRef const& pp = params.phys;
    if (state.sim.post_step_action[slot] != pp.scalars.discrete_action
        || status_of(state, slot) == TrackStatus::inactive)
        return;
    state.phys.interaction_mfp[slot] = 0;
    std::uint32_t const pid = state.particle.particle_id[slot];
    ParticlePhysics const& part = pp.particle[pid];
    std::uint32_t const mat = material_of(params, state, slot);
    double const energy = state.particle.energy[slot];
    XorwowEngine rng = XorwowEngine::load(state.rng, slot);
    double accum = -state.phys.macro_xs[slot] * draw_canonical(rng);
    std::uint32_t pidx = part.num_processes - 1;
    for (std::uint32_t i = 0; i + 1 < part.num_processes; ++i)
    {
        accum += state.phys.per_process_xs[i * state.size + slot];
        if (accum > 0)
        {
            pidx = i;
            break;
        }
    }
    ProcessRecord const& proc = part.process[pidx];
    std::uint8_t action;
    bool rejected = false;
    if (proc.integral)
    {
        double const xs = phys::calc_xs(pp, proc, mat, energy);
        if (draw_canonical(rng)
                * state.phys.per_process_xs[pidx * state.size + slot]
            > xs)
        {
            rejected = true;
        }
    }
    if (rejected)
    {
        action = pp.scalars.integral_rejection_action;
    }
    else
    {
        std::uint32_t mi = 0;
        for (std::uint32_t k = 1; k < proc.num_models; ++k)
        {
            if (energy >= proc.models[k].energy_low)
                mi = k;
        }
        state.phys.element[slot] = static_cast<std::uint8_t>(
            sample_element(pp, proc, mi, mat, energy, rng));
        action = static_cast<std::uint8_t>(proc.models[mi].action);
    }
    rng.store(state.rng, slot);
    state.sim.post_step_action[slot] = action;
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
    PhysicsParamsRef const& pp = params.phys;
    double const energy = state.particle.energy[slot];
    double dir[3];
    for (int k = 0; k < 3; ++k)
        dir[k] = state.geo.dir[k][slot];
    std::uint32_t const mat = material_of(params, state, slot);
    XorwowEngine rng = XorwowEngine::load(state.rng, slot);
    InteractionContext ctx{pp, state, slot, mat, state.phys.element[slot],
                           state.particle.particle_id[slot]};
    Interaction r = sample_interaction(M, ctx, energy, dir, rng);
    rng.store(state.rng, slot);
    apply_interaction(params, state, slot, mat, r);
}

//---------------------------------------------------------------------------//
// BOUNDARY AND TRACKING CUT
//---------------------------------------------------------------------------//
__global__ void __launch_bounds__(block_size)
    boundary_kernel(CoreParamsRef const params,
                    CoreStateRef const state,
                    ThreadSlots const ts)
{
    size_type slot;
    if (!thread_to_slot(ts, blockIdx.x * blockDim.x + threadIdx.x, state.size, &slot))
        return;
    if (state.sim.post_step_action[slot] != params.scalars.boundary_action
        || status_of(state, slot) != TrackStatus::alive)
        return;
    if (state.diag)
        wave::count_active_lanes(state.diag + diag_boundary);
    geo::GeoTrack g = geo::GeoTrack::load(state.geo, slot);
    geo::cross_boundary(params.geo, g);
    g.store(state.geo, slot);
    if (g.failed)
    {
        apply_errored(params, state, slot);
    }
    else if (!geo::is_outside(params.geo, g))
    {
        if (params.volume_to_material[g.volume] == invalid_id)
            apply_errored(params, state, slot);
    }
    else
    {
        state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::killed);
    }
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
    std::uint32_t const pid = state.particle.particle_id[slot];
    double dep = state.particle.energy[slot];
    if (pid == params.phys.positron_id)
        dep += 2 * params.phys.mass[pid];
    state.phys.energy_deposition[slot] += dep;
    state.particle.energy[slot] = 0;
    state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::killed);
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
hipError_t CELER_PREC_FN(launch_pre_step)(CoreParamsRef const& p,
                           CoreStateRef const& s,
                           ThreadSlots const& ts,
                           hipStream_t st)
{
    pre_step_kernel<<<grid_of(s), s.block_size, 0, st>>>(p, s, ts);
    return hipGetLastError();
}

#if !CELER_PREC_IS_MIXED
hipError_t launch_along_step_neutral(CoreParamsRef const& p,
                                     CoreStateRef const& s,
                                     std::uint8_t action,
                                     ThreadSlots const& ts,
                                     hipStream_t st)
{
    along_step_neutral_kernel<<<grid_of(s), s.block_size, 0, st>>>(p, s, action, ts);
    return hipGetLastError();
}

#endif

hipError_t CELER_PREC_FN(launch_discrete_select)(CoreParamsRef const& p,
                                  CoreStateRef const& s,
                                  ThreadSlots const& ts,
                                  hipStream_t st)
{
    discrete_select_kernel<<<grid_of(s), s.block_size, 0, st>>>(p, s, ts);
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
#define CELER_LAUNCH_MODEL(M)                                              \
    case ModelClass::M:                                                    \
        interaction_kernel<ModelClass::M><<<grid_of(s), s.block_size, 0, st>>>(p, s, action, ts); \
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

#if !CELER_PREC_IS_MIXED
hipError_t launch_boundary(CoreParamsRef const& p,
                           CoreStateRef const& s,
                           ThreadSlots const& ts,
                           hipStream_t st)
{
    boundary_kernel<<<grid_of(s), s.block_size, 0, st>>>(p, s, ts);
    return hipGetLastError();
}

hipError_t launch_tracking_cut(CoreParamsRef const& p,
                               CoreStateRef const& s,
                               ThreadSlots const& ts,
                               hipStream_t st)
{
    tracking_cut_kernel<<<grid_of(s), s.block_size, 0, st>>>(p, s, ts);
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
        calo_atomic_kernel<<<num_blocks(s.size), block_size, 0, st>>>(p, s, ts);
        return hipGetLastError();
    }
    unsigned int const rows = std::min<unsigned int>(num_blocks(s.size), calo_tally_blocks);
    calo_partial_kernel<<<rows, block_size, 0, st>>>(p, s, num_det, ts);
    calo_reduce_kernel<<<num_det, block_size, 0, st>>>(s, num_det, rows);
    return hipGetLastError();
}

#endif

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
