//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file user/UserKernels.hip
//! Kernels of the user actions: particle tallies and detector step points.
//!
//! The tallies (ActionDiagnosticExecutor.hh, StepDiagnosticExecutor.hh)
//! histogram into LDS: at most tally_blocks workgroups stride over the slots,
//! count in 32-bit LDS bins and flush each nonzero bin with one 64-bit
//! global atomic, instead of one global atomic per track on a few hot bins.
//!
//! The step-point gather (StepGatherExecutor.hh:41-153) stores the pre-step
//! point of each slot in a detector. At the step's end three launches
//! compact the slots with hits in slot order (ballot + mbcnt within a wave,
//! an LDS scan within a workgroup, one workgroup scanning the per-workgroup
//! counts) and write the hits to the device hit buffer, the post-step fields
//! read from the core state (DetectorSteps.cu:31-90).
//---------------------------------------------------------------------------//
#include "UserKernels.hh"

#include "corecel/WaveOps.hh"

namespace celeritas_amd
{
namespace
{
//---------------------------------------------------------------------------//
constexpr int block_size = 256;

__device__ __forceinline__ TrackStatus status_at(CoreStateRef const& s, size_type i)
{
    return static_cast<TrackStatus>(s.sim.status[i]);
}

//---------------------------------------------------------------------------//
// PARTICLE TALLIES
//---------------------------------------------------------------------------//
enum class TallyKind
{
    action,  //!< post-step action of every valid track
    steps,  //!< step count of each track killed this step
};

template<TallyKind K>
__global__ void __launch_bounds__(block_size)
    particle_tally_kernel(CoreStateRef const state, ParticleTallyRef const tally)
{
    __shared__ unsigned int bins[max_tally_bins];
    unsigned int const nbins = tally.num_particles * tally.num_bins;
    for (unsigned int i = threadIdx.x; i < nbins; i += block_size)
        bins[i] = 0;
    __syncthreads();
    for (size_type slot = blockIdx.x * block_size + threadIdx.x; slot < state.size;
         slot += gridDim.x * block_size)
    {
        TrackStatus const st = status_at(state, slot);
        unsigned int bin;
        if (K == TallyKind::action)
        {
            if (st == TrackStatus::inactive)
                continue;
            bin = state.sim.post_step_action[slot];
            if (bin >= tally.num_bins)
                continue;
        }
        else
        {
            if (st != TrackStatus::killed)
                continue;
            unsigned int const n = state.sim.num_steps[slot];
            bin = n < tally.num_bins - 1 ? n : tally.num_bins - 1;
        }
        unsigned int const pid = state.particle.particle_id[slot];
        if (pid >= tally.num_particles)
            continue;
        atomicAdd(&bins[pid * tally.num_bins + bin], 1u);
    }
    __syncthreads();
    for (unsigned int i = threadIdx.x; i < nbins; i += block_size)
    {
        if (bins[i] != 0)
            atomicAdd(tally.counts + i, static_cast<unsigned long long>(bins[i]));
    }
}

//---------------------------------------------------------------------------//
// STEP-POINT GATHER
//---------------------------------------------------------------------------//
//! Store the pre-step point of an active slot in a detector
__global__ void __launch_bounds__(block_size)
    step_gather_pre_kernel(CoreParamsRef const params,
                           CoreStateRef const state,
                           StepCollectorRef const coll)
{
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    if (slot >= state.size)
        return;
    if (status_at(state, slot) == TrackStatus::inactive)
    {
        coll.slot_detector[slot] = invalid_id;
        return;
    }
    std::uint32_t const vol = state.geo.volume[slot];
    std::uint32_t const det = coll.detector_map ? coll.detector_map[vol] : 0u;
    coll.slot_detector[slot] = det;
    if (det == invalid_id)
        return;
    StepPointSelection const& sel = coll.selection.points[step_pre];
    if (sel.time)
        coll.pre.time[slot] = state.sim.time[slot];
    for (int j = 0; j < 3; ++j)
    {
        if (sel.pos)
            coll.pre.pos[j][slot] = state.geo.pos[j][slot];
        if (sel.dir)
            coll.pre.dir[j][slot] = state.geo.dir[j][slot];
    }
    if (sel.volume_id)
        coll.pre.volume_id[slot] = vol;
    if (sel.energy)
        coll.pre.energy[slot] = state.particle.energy[slot];
}

//! Whether the slot has a hit at the end of this step
__device__ __forceinline__ bool has_hit(CoreStateRef const& state,
                                        StepCollectorRef const& coll,
                                        size_type slot)
{
    if (slot >= state.size || status_at(state, slot) == TrackStatus::inactive)
        return false;
    if (coll.slot_detector[slot] == invalid_id)
        return false;
    return !coll.nonzero_energy_deposition || state.phys.energy_deposition[slot] != 0;
}

//! Number of hits per workgroup
__global__ void __launch_bounds__(block_size)
    step_gather_count_kernel(CoreStateRef const state, StepCollectorRef const coll)
{
    __shared__ unsigned int wave_counts[block_size / wave::size];
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    std::uint64_t const mask = wave::ballot(has_hit(state, coll, slot));
    if (wave::lane_id() == 0)
        wave_counts[threadIdx.x / wave::size] = __popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned int total = 0;
        for (int w = 0; w < block_size / wave::size; ++w)
            total += wave_counts[w];
        coll.block_counts[blockIdx.x] = total;
    }
}

//! One workgroup: exclusive scan of the workgroup counts into hit-buffer
//! offsets; the step's hits are kept only if they all fit
__global__ void __launch_bounds__(1024)
    step_gather_scan_kernel(StepCollectorRef const coll, unsigned int num_blocks)
{
    constexpr int nt = 1024;
    constexpr int nw = nt / wave::size;
    __shared__ unsigned int wave_tot[nw];
    unsigned int const per = (num_blocks + nt - 1) / nt;
    unsigned int const first = threadIdx.x * per;
    unsigned int local = 0;
    for (unsigned int i = 0; i < per && first + i < num_blocks; ++i)
        local += coll.block_counts[first + i];
    unsigned int const incl = wave::inclusive_scan(local);
    int const w = threadIdx.x / wave::size;
    if (wave::lane_id() == wave::size - 1)
        wave_tot[w] = incl;
    __syncthreads();
    unsigned int before = 0, total = 0;
    for (int i = 0; i < nw; ++i)
    {
        if (i < w)
            before += wave_tot[i];
        total += wave_tot[i];
    }
    StepHitCounters* c = coll.counters;
    unsigned int const base = c->num_hits;
    bool const fits = static_cast<unsigned long long>(base) + total <= coll.hits.capacity;
    unsigned int offset = base + before + incl - local;
    for (unsigned int i = 0; i < per && first + i < num_blocks; ++i)
    {
        unsigned int const n = coll.block_counts[first + i];
        coll.block_counts[first + i] = offset;
        offset += n;
    }
    // Every thread read num_hits before thread 0 updates it
    __syncthreads();
    if (threadIdx.x == 0)
    {
        if (fits)
        {
            c->num_hits = base + total;
            c->step_hits = total;
        }
        else
        {
            c->step_hits = 0;
            c->overflow += total > 0 ? 1u : 0u;
            c->dropped += total;
        }
    }
}

//! Write the hits of a workgroup at its offset, in slot order
__global__ void __launch_bounds__(block_size)
    step_gather_scatter_kernel(CoreParamsRef const params,
                               CoreStateRef const state,
                               StepCollectorRef const coll)
{
    __shared__ unsigned int lds[2 * (block_size / wave::size)];
    if (coll.counters->step_hits == 0)
        return;
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    bool const hit = has_hit(state, coll, slot);
    wave::BlockScan2<block_size / wave::size> scan(hit ? 1u : 0u, 0u, lds);
    if (!hit)
        return;
    std::uint32_t const h = coll.block_counts[blockIdx.x] + scan.excl_a;
    StepHitsRef const& out = coll.hits;
    StepSelection const& sel = coll.selection;
    out.detector[h] = coll.slot_detector[slot];
    out.track_id[h] = state.sim.track_id[slot];
    // Pre-step point from the slot arrays
    {
        StepPointSelection const& ps = sel.points[step_pre];
        StepPointHits const& o = out.points[step_pre];
        if (ps.time)
            o.time[h] = coll.pre.time[slot];
        for (int j = 0; j < 3; ++j)
        {
            if (ps.pos)
                o.pos[3 * h + j] = coll.pre.pos[j][slot];
            if (ps.dir)
                o.dir[3 * h + j] = coll.pre.dir[j][slot];
        }
        if (ps.volume_id)
            o.volume_id[h] = coll.pre.volume_id[slot];
        if (ps.energy)
            o.energy[h] = coll.pre.energy[slot];
    }
    // Post-step point from the core state
    {
        StepPointSelection const& ps = sel.points[step_post];
        StepPointHits const& o = out.points[step_post];
        if (ps.time)
            o.time[h] = state.sim.time[slot];
        for (int j = 0; j < 3; ++j)
        {
            if (ps.pos)
                o.pos[3 * h + j] = state.geo.pos[j][slot];
            if (ps.dir)
                o.dir[3 * h + j] = state.geo.dir[j][slot];
        }
        if (ps.volume_id)
        {
            std::uint32_t const vol = state.geo.volume[slot];
            o.volume_id[h] = vol == params.geo.exterior ? invalid_id : vol;
        }
        if (ps.energy)
            o.energy[h] = state.particle.energy[slot];
    }
    if (sel.event_id)
        out.event_id[h] = state.sim.event_id[slot];
    if (sel.parent_id)
        out.parent_id[h] = state.sim.parent_id[slot];
    if (sel.track_step_count)
        out.track_step_count[h] = state.sim.num_steps[slot];
    if (sel.action_id)
    {
        std::uint8_t const a = state.sim.post_step_action[slot];
        out.action_id[h] = a == invalid_id8 ? invalid_id : a;
    }
    if (sel.step_length)
        out.step_length[h] = state.sim.step_length[slot];
    if (sel.particle)
        out.particle[h] = state.particle.particle_id[slot];
    if (sel.energy_deposition)
        out.energy_deposition[h] = state.phys.energy_deposition[slot];
}

//! Empty the hit buffer (never captured in a graph)
__global__ void step_hits_clear_kernel(StepHitCounters* c)
{
    c->num_hits = 0;
    c->step_hits = 0;
    c->overflow = 0;
    c->dropped = 0;
}

unsigned int blocks_of(size_type n)
{
    return (n + block_size - 1) / block_size;
}

//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
hipError_t launch_action_tally(CoreStateRef const& state,
                               ParticleTallyRef const& tally,
                               hipStream_t stream)
{
    if (state.size == 0)
        return hipSuccess;
    unsigned int const grid = std::min(tally_blocks, blocks_of(state.size));
    particle_tally_kernel<TallyKind::action><<<grid, block_size, 0, stream>>>(state, tally);
    return hipGetLastError();
}

hipError_t launch_step_tally(CoreStateRef const& state,
                             ParticleTallyRef const& tally,
                             hipStream_t stream)
{
    if (state.size == 0)
        return hipSuccess;
    unsigned int const grid = std::min(tally_blocks, blocks_of(state.size));
    particle_tally_kernel<TallyKind::steps><<<grid, block_size, 0, stream>>>(state, tally);
    return hipGetLastError();
}

hipError_t launch_step_gather_pre(CoreParamsRef const& params,
                                  CoreStateRef const& state,
                                  StepCollectorRef const& coll,
                                  hipStream_t stream)
{
    if (state.size == 0)
        return hipSuccess;
    step_gather_pre_kernel<<<blocks_of(state.size), block_size, 0, stream>>>(
        params, state, coll);
    return hipGetLastError();
}

hipError_t launch_step_gather_post(CoreParamsRef const& params,
                                   CoreStateRef const& state,
                                   StepCollectorRef const& coll,
                                   hipStream_t stream)
{
    if (state.size == 0)
        return hipSuccess;
    unsigned int const nb = blocks_of(state.size);
    step_gather_count_kernel<<<nb, block_size, 0, stream>>>(state, coll);
    step_gather_scan_kernel<<<1, 1024, 0, stream>>>(coll, nb);
    step_gather_scatter_kernel<<<nb, block_size, 0, stream>>>(params, state, coll);
    return hipGetLastError();
}

hipError_t launch_step_hits_clear(StepHitCounters* counters, hipStream_t stream)
{
    step_hits_clear_kernel<<<1, 1, 0, stream>>>(counters);
    return hipGetLastError();
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
