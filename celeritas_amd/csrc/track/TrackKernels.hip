//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file track/TrackKernels.hip
//! Primaries -> initializers, track initialization and the fused
//! locate-alive / vacancy compaction / secondary scan / initializer creation.
//!
//! All counters live in device memory (CoreCounters). Launch extents that
//! depend on counters are fixed upper bounds; a one-lane "prepare" kernel
//! computes the real extent and the base offsets before the worker kernel,
//! so a whole step can be enqueued (or captured in a graph) with no host
//! round trip. The reference instead copies counts back to the host after
//! thrust::remove_if and exclusive_scan (celeritas/track/detail/
//! TrackInitAlgorithms.cu:34-105).
//---------------------------------------------------------------------------//
#include "TrackKernels.hh"

#include "corecel/KernelLaunch.hh"
#include "corecel/WaveOps.hh"
#include "TrackInitImpl.hh"

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

//---------------------------------------------------------------------------//
// PRIMARIES
//---------------------------------------------------------------------------//
__global__ void primaries_prepare_kernel(CoreCounters* c, size_type count)
{
    c->init_base = c->num_initializers;
    c->num_initializers += count;
    c->num_generated += count;
    // New primaries go on top of the previous step's secondaries, whose
    // parent slots are then no longer at the front of the stack
    if (count > 0)
        c->prev_secondaries = 0;
}

//! Initializer of a primary (ProcessPrimariesExecutor.hh:56-76)
__device__ __forceinline__ TrackInitializer primary_initializer(Primary const& p,
                                                                std::uint32_t track_id)
{
    TrackInitializer ti;
    ti.track_id = track_id;
    ti.parent_id = invalid_id;
    ti.event_id = p.event_id;
    ti.particle_id = p.particle_id;
    ti.time = p.time;
    for (int k = 0; k < 3; ++k)
    {
        ti.pos[k] = p.pos[k];
        ti.dir[k] = p.dir[k];
    }
    ti.energy = p.energy;
    ti.volume = invalid_id;
    ti.reserved = 0;
    return ti;
}

/*!
 * Create initializers from primaries (ProcessPrimariesExecutor.hh:56-76).
 *
 * With event-local ranks (computed by the host when it copies the batch)
 * the track ID is the event's counter plus the rank, which reproduces the
 * serial order of the reference's host loop; the counters are advanced by
 * a second kernel. Without ranks the ID comes from an atomic per-event
 * counter, as in the reference's device path.
 */
__global__ void __launch_bounds__(block_size)
    process_primaries_kernel(CoreStateRef const state,
                             Primary const* __restrict__ primaries,
                             std::uint32_t const* __restrict__ ranks,
                             size_type count)
{
    size_type const tid = blockIdx.x * block_size + threadIdx.x;
    if (tid >= count)
        return;
    Primary const p = primaries[tid];
    size_type const idx = state.counters->init_base + tid;
    std::uint32_t track_id = invalid_id;
    if (p.event_id < state.init.max_events)
    {
        track_id = ranks ? state.init.track_counters[p.event_id] + ranks[tid]
                         : atomicAdd(&state.init.track_counters[p.event_id], 1u);
        if (state.init.event_tally)
            atomicAdd(&state.init.event_live[p.event_id], 1);
    }
    else
    {
        atomicOr(&state.counters->error_flags, err_event_overflow);
    }
    if (idx < state.init.capacity)
        state.init.initializers[idx] = primary_initializer(p, track_id);
}

//---------------------------------------------------------------------------//
// DEVICE-SIDE ADMISSION (serving queue)
//---------------------------------------------------------------------------//
/*!
 * Admit arrived events in order while the queued initializers stay within
 * the limit; the first waiting event is admitted regardless when nothing
 * is queued or alive. Takes the place of primaries_prepare_kernel: the
 * admitted primaries go on top of the initializer stack.
 */
__global__ void serve_admit_kernel(CoreStateRef const state)
{
    if (threadIdx.x != 0)
        return;
    ServeQueueRef const& q = state.serve;
    ServeCounters& sc = *q.counters;
    CoreCounters& c = *state.counters;
    std::uint32_t const begin = sc.admitted;
    std::uint32_t const end = min(sc.arrived, static_cast<std::uint32_t>(q.num_events));
    std::uint32_t const queued = c.num_initializers;
    bool const idle = queued == 0 && c.num_alive == 0;
    std::uint32_t total = 0;
    std::uint32_t e = begin;
    for (; e < end; ++e)
    {
        std::uint32_t const n = q.event_count[e];
        if (!(idle && e == begin) && queued + total + n > q.max_queued)
            break;
        total += n;
        q.admit_step[e] = c.steps_taken;
    }
    sc.first = begin < end ? q.event_first[begin] : 0;
    sc.count = total;
    sc.admitted = e;
    c.init_base = c.num_initializers;
    c.num_initializers += total;
    c.num_generated += total;
    if (total > 0)
        c.prev_secondaries = 0;
}

/*!
 * Initializers of the primaries admitted this step. A new event's track
 * counter is zero, so a primary's track ID is its rank within its event;
 * the event's last primary sets the counter past them.
 */
__global__ void __launch_bounds__(block_size) serve_primaries_kernel(CoreStateRef const state)
{
    ServeQueueRef const& q = state.serve;
    std::uint32_t const count = q.counters->count;
    std::uint32_t const first = q.counters->first;
    size_type const base = state.counters->init_base;
    for (std::uint32_t i = blockIdx.x * block_size + threadIdx.x; i < count;
         i += gridDim.x * block_size)
    {
        Primary const p = q.primaries[first + i];
        std::uint32_t const ev = p.event_id;
        std::uint32_t const rank = first + i - q.event_first[ev];
        if (state.init.event_tally)
            atomicAdd(&state.init.event_live[ev], 1);
        if (rank + 1 == q.event_count[ev])
            state.init.track_counters[ev] = q.event_count[ev];
        if (base + i < state.init.capacity)
            state.init.initializers[base + i] = primary_initializer(p, rank);
    }
}

//! Workgroups of the admission copy (a grid-stride loop over the primaries)
constexpr unsigned int serve_blocks = 512;

//! Advance per-event counters after ranked ID assignment (last of each event)
__global__ void __launch_bounds__(block_size)
    advance_track_counters_kernel(CoreStateRef const state,
                                  Primary const* __restrict__ primaries,
                                  std::uint32_t const* __restrict__ ranks,
                                  std::uint32_t const* __restrict__ is_last,
                                  size_type count)
{
    size_type const tid = blockIdx.x * block_size + threadIdx.x;
    if (tid >= count || !is_last[tid])
        return;
    std::uint32_t const ev = primaries[tid].event_id;
    if (ev < state.init.max_events)
        state.init.track_counters[ev] += ranks[tid] + 1;
}

//---------------------------------------------------------------------------//
// TRACK INITIALIZATION
//---------------------------------------------------------------------------//
__global__ void init_prepare_kernel(CoreCounters* c, size_type num_slots)
{
    size_type const n = min(c->num_vacancies, c->num_initializers);
    c->num_new_tracks = n;
    c->init_base = c->num_initializers;
    c->vacancy_base = c->num_vacancies;
    c->num_initializers -= n;
    c->num_vacancies -= n;
    c->num_active = num_slots - c->num_vacancies;
    c->steps_taken += 1;
    c->track_steps += c->num_active;
}

/*!
 * Pop initializers from the back into vacancies from the back
 * (InitTracksExecutor.hh:71-175 with TrackOrder::none).
 */
__global__ void __launch_bounds__(block_size)
    init_tracks_kernel(CoreParamsRef const params, CoreStateRef const state)
{
    size_type const tid = blockIdx.x * block_size + threadIdx.x;
    CoreCounters const& c = *state.counters;
    if (tid >= c.num_new_tracks)
        return;
    TrackInitializer const ti = state.init.initializers[c.init_base - tid - 1];
    size_type const slot = state.init.vacancies[c.vacancy_base - tid - 1];
    init_track(params, state, slot, ti);
}

//---------------------------------------------------------------------------//
// CHARGE PARTITION (TrackOrder::init_charge)
//---------------------------------------------------------------------------//
__device__ __forceinline__ unsigned int
is_neutral(CoreParamsRef const& params, TrackInitializer const& ti)
{
    return params.phys.charge[ti.particle_id] == 0 ? 1u : 0u;
}

//! Neutral flags of the new initializers, reduced per workgroup
__global__ void __launch_bounds__(block_size)
    partition_count_kernel(CoreParamsRef const params, CoreStateRef const state)
{
    __shared__ unsigned int lds[2 * num_waves];
    size_type const tid = blockIdx.x * block_size + threadIdx.x;
    CoreCounters const& c = *state.counters;
    size_type const n = c.num_new_tracks;
    unsigned int neutral = 0;
    if (tid < n)
        neutral = is_neutral(params, state.init.initializers[c.init_base - n + tid]);
    wave::BlockScan2<num_waves> scan(neutral, 0u, lds);
    if (threadIdx.x == 0)
        state.init.block_scratch[2 * blockIdx.x] = scan.total_a;
}

//! Exclusive scan of the workgroup neutral counts (one workgroup)
__global__ void __launch_bounds__(1024)
    partition_scan_kernel(CoreStateRef const state, unsigned int nblocks)
{
    constexpr int nw = 1024 / wave::size;
    __shared__ unsigned int lds[2 * nw];
    __shared__ unsigned int carry;
    if (threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    unsigned int* part = state.init.block_scratch;
    for (unsigned int base = 0; base < nblocks; base += 1024)
    {
        unsigned int const i = base + threadIdx.x;
        unsigned int const a = i < nblocks ? part[2 * i] : 0u;
        wave::BlockScan2<nw> scan(a, 0u, lds);
        unsigned int const ca = carry;
        if (i < nblocks)
            part[2 * i] = ca + scan.excl_a;
        __syncthreads();
        if (threadIdx.x == 0)
            carry = ca + scan.total_a;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        state.counters->num_new_neutral = carry;
}

/*!
 * Stable partition of the indices 0..n-1 of the new initializers, neutral
 * first (std::stable_partition in TrackInitAlgorithms.cc:80-94).
 */
__global__ void __launch_bounds__(block_size)
    partition_scatter_kernel(CoreParamsRef const params, CoreStateRef const state)
{
    __shared__ unsigned int lds[2 * num_waves];
    size_type const tid = blockIdx.x * block_size + threadIdx.x;
    CoreCounters const& c = *state.counters;
    size_type const n = c.num_new_tracks;
    unsigned int neutral = 0;
    if (tid < n)
        neutral = is_neutral(params, state.init.initializers[c.init_base - n + tid]);
    wave::BlockScan2<num_waves> scan(neutral, 0u, lds);
    if (tid >= n)
        return;
    unsigned int const excl = state.init.block_scratch[2 * blockIdx.x] + scan.excl_a;
    unsigned int const pos = neutral ? excl : c.num_new_neutral + (tid - excl);
    state.init.indices[pos] = tid;
}

/*!
 * Initialize with the charge partition (InitTracksExecutor.hh:78-104):
 * thread tid takes initializer indices[n-1-tid]; neutral tracks fill the
 * vacancies from the front, charged ones from the back.
 */
__global__ void __launch_bounds__(block_size)
    init_tracks_charge_kernel(CoreParamsRef const params, CoreStateRef const state)
{
    size_type const tid = blockIdx.x * block_size + threadIdx.x;
    CoreCounters const& c = *state.counters;
    size_type const n = c.num_new_tracks;
    if (tid >= n)
        return;
    size_type const idx = state.init.indices[n - 1 - tid] + c.init_base - n;
    TrackInitializer const ti = state.init.initializers[idx];
    size_type const vac = is_neutral(params, ti) ? n - 1 - tid
                                                 : c.vacancy_base - 1 - tid;
    init_track(params, state, state.init.vacancies[vac], ti);
}

//---------------------------------------------------------------------------//
// EXTEND FROM SECONDARIES
//---------------------------------------------------------------------------//
//! Number of live secondaries of a slot, counting empty (cut) entries out
__device__ __forceinline__ size_type count_secondaries(CoreStateRef const& s,
                                                       size_type slot)
{
    size_type const start = s.phys.sec_start[slot];
    size_type const n = s.phys.sec_count[slot];
    size_type result = 0;
    for (size_type i = 0; i < n; ++i)
    {
        if (s.secondaries.storage[start + i].particle_id != invalid_id)
            ++result;
    }
    return result;
}

constexpr std::uint32_t vacancy_bit = 0x80000000u;

/*!
 * Locate alive tracks and count surviving secondaries
 * (LocateAliveExecutor.hh:60-106), then reduce per workgroup.
 *
 * secondary_counts[slot] holds the count with the vacancy flag in bit 31.
 */
__global__ void __launch_bounds__(block_size)
    locate_alive_kernel(CoreStateRef const state, std::uint8_t in_place)
{
    __shared__ unsigned int lds[2 * num_waves];
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    unsigned int vac = 0, nsec = 0;
    if (slot < state.size)
    {
        auto const status = static_cast<TrackStatus>(state.sim.status[slot]);
        if (status != TrackStatus::inactive)
            nsec = count_secondaries(state, slot);
        if (status == TrackStatus::alive)
        {
            vac = 0;
        }
        else if (nsec > 0 && in_place)
        {
            --nsec;
            vac = 0;
        }
        else
        {
            vac = 1;
        }
        state.init.secondary_counts[slot] = nsec | (vac ? vacancy_bit : 0u);
    }
    wave::BlockScan2<num_waves> scan(vac, nsec, lds);
    if (threadIdx.x == 0)
    {
        state.init.block_scratch[2 * blockIdx.x] = scan.total_a;
        state.init.block_scratch[2 * blockIdx.x + 1] = scan.total_b;
    }
}

/*!
 * Exclusive scan over workgroup partials (one workgroup) and update of the
 * counters (ExtendFromSecondariesAction.cc:56-100).
 */
__global__ void __launch_bounds__(1024)
    scan_partials_kernel(CoreStateRef const state, unsigned int nblocks)
{
    constexpr int nw = 1024 / wave::size;
    __shared__ unsigned int lds[2 * nw];
    __shared__ unsigned int carry[2];
    if (threadIdx.x == 0)
    {
        carry[0] = carry[1] = 0;
    }
    __syncthreads();
    unsigned int* part = state.init.block_scratch;
    // Each pass gives every thread a run of up to held consecutive pairs,
    // read with independent loads, so a pass costs one memory latency
    constexpr unsigned int held = 8;
    for (unsigned int base = 0; base < nblocks; base += 1024 * held)
    {
        unsigned int const first = base + threadIdx.x * held;
        unsigned int va[held], vb[held];
        unsigned int a = 0, b = 0;
#pragma unroll
        for (unsigned int u = 0; u < held; ++u)
        {
            unsigned int const i = first + u;
            va[u] = i < nblocks ? part[2 * i] : 0u;
            vb[u] = i < nblocks ? part[2 * i + 1] : 0u;
        }
#pragma unroll
        for (unsigned int u = 0; u < held; ++u)
        {
            a += va[u];
            b += vb[u];
        }
        wave::BlockScan2<nw> scan(a, b, lds);
        unsigned int ra = carry[0] + scan.excl_a;
        unsigned int rb = carry[1] + scan.excl_b;
#pragma unroll
        for (unsigned int u = 0; u < held; ++u)
        {
            unsigned int const i = first + u;
            if (i < nblocks)
            {
                part[2 * i] = ra;
                part[2 * i + 1] = rb;
            }
            ra += va[u];
            rb += vb[u];
        }
        __syncthreads();
        if (threadIdx.x == 0)
        {
            carry[0] += scan.total_a;
            carry[1] += scan.total_b;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
    {
        CoreCounters* c = state.counters;
        unsigned int const nvac = carry[0];
        unsigned int const nsec = carry[1];
        c->num_vacancies = nvac;
        c->num_secondaries = nsec;
        c->init_base = c->num_initializers;
        c->num_initializers += nsec;
        if (c->num_initializers > state.init.capacity)
            c->error_flags |= err_initializer_overflow;
        c->num_alive = state.size - nvac;
        c->prev_secondaries = nsec;
    }
}

/*!
 * Compact vacancies (stable, like std::remove_if) and create initializers
 * from secondaries in slot order (ProcessSecondariesExecutor.hh:70-185).
 */
__global__ void __launch_bounds__(block_size)
    process_secondaries_kernel(CoreStateRef const state, std::uint8_t in_place)
{
    __shared__ unsigned int lds[2 * num_waves];
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    unsigned int vac = 0, nsec = 0;
    if (slot < state.size)
    {
        unsigned int const packed = state.init.secondary_counts[slot];
        vac = (packed & vacancy_bit) ? 1 : 0;
        nsec = packed & ~vacancy_bit;
    }
    wave::BlockScan2<num_waves> scan(vac, nsec, lds);
    if (slot >= state.size)
        return;

    unsigned int const vac_idx = state.init.block_scratch[2 * blockIdx.x]
                                 + scan.excl_a;
    unsigned int const sec_prefix
        = state.init.block_scratch[2 * blockIdx.x + 1] + scan.excl_b;
    if (vac)
        state.init.vacancies[vac_idx] = slot;
    // Expose the scanned offsets like the reference's exclusive_scan output
    state.init.secondary_counts[slot] = sec_prefix;

    auto status = static_cast<TrackStatus>(state.sim.status[slot]);
    if (status == TrackStatus::inactive)
        return;

    CoreCounters const& c = *state.counters;
    unsigned int next_idx = c.init_base + sec_prefix;
    bool initialized = false;
    size_type const start = state.phys.sec_start[slot];
    size_type const count = state.phys.sec_count[slot];
    // Event bookkeeping: +1 per secondary made, -1 for a dead parent
    std::uint32_t const event = state.sim.event_id[slot];
    int live_delta = status == TrackStatus::alive ? 0 : -1;
    if (count > 0)
    {
        std::uint32_t const parent_id = state.sim.track_id[slot];
        std::uint32_t const event_id = state.sim.event_id[slot];
        double const time = state.sim.time[slot];
        double pos[3];
        for (int k = 0; k < 3; ++k)
            pos[k] = state.geo.pos[k][slot];
        // A parent in a volume's interior (not on a surface) hands its volume
        // to the secondaries made at its position
        std::uint32_t const volume
            = (state.geo.surface[slot] == invalid_id && !state.geo.failed[slot])
                  ? state.geo.volume[slot]
                  : invalid_id;
        for (size_type i = 0; i < count; ++i)
        {
            Secondary const sec = state.secondaries.storage[start + i];
            if (sec.particle_id == invalid_id)
                continue;
            ++live_delta;
            TrackInitializer ti;
            ti.track_id = atomicAdd(&state.init.track_counters[event_id], 1u);
            ti.parent_id = parent_id;
            ti.event_id = event_id;
            ti.particle_id = sec.particle_id;
            ti.time = time;
            ti.energy = sec.energy;
            for (int k = 0; k < 3; ++k)
            {
                ti.pos[k] = pos[k];
                ti.dir[k] = sec.dir[k];
            }
            ti.volume = volume;
            ti.reserved = 0;
            if (!initialized && status != TrackStatus::alive && in_place)
            {
                // Refill the killed parent's slot in place: keep geometry,
                // take the secondary's direction, reset physics
                init_sim(state.sim, slot, ti);
                state.particle.particle_id[slot]
                    = static_cast<std::uint8_t>(ti.particle_id);
                state.particle.energy[slot] = ti.energy;
                for (int k = 0; k < 3; ++k)
                    state.geo.dir[k][slot] = ti.dir[k];
                state.geo.failed[slot] = 0;
                state.geo.next_step[slot] = 0;
                state.geo.next_surface[slot] = invalid_id;
                reset_physics(state.phys, slot);
                initialized = true;
            }
            else
            {
                if (next_idx < state.init.capacity)
                    state.init.initializers[next_idx] = ti;
                ++next_idx;
            }
        }
    }
    if (state.init.event_tally && event < state.init.max_events)
    {
        if (live_delta != 0)
        {
            // A track only changes its own event's count by -1 (death
            // without secondaries) or more, and every other live track or
            // initializer of the event is still counted, so the update that
            // takes the count to zero is the event's completion
            int const prev = atomicAdd(&state.init.event_live[event], live_delta);
            if (prev + live_delta == 0)
            {
                std::uint32_t const idx = atomicAdd(&state.counters->num_completed, 1u);
                if (idx < state.init.max_events)
                    state.init.completed[idx] = event;
            }
        }
        double const edep = state.phys.energy_deposition[slot];
        if (edep != 0)
            atomicAdd(&state.init.event_edep[event], edep);
    }
    if (!initialized && status == TrackStatus::killed)
    {
        state.sim.status[slot] = static_cast<std::uint8_t>(TrackStatus::inactive);
    }
}

//---------------------------------------------------------------------------//
// UTILITIES
//---------------------------------------------------------------------------//
__global__ void fill_u32_kernel(std::uint32_t* data, size_type n, std::uint32_t v)
{
    size_type const i = blockIdx.x * block_size + threadIdx.x;
    if (i < n)
        data[i] = v;
}

__global__ void iota_u32_kernel(std::uint32_t* data, size_type n)
{
    size_type const i = blockIdx.x * block_size + threadIdx.x;
    if (i < n)
        data[i] = i;
}

//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
// LAUNCHERS
//---------------------------------------------------------------------------//
hipError_t launch_extend_from_primaries(CoreStateRef const& state,
                                        Primary const* primaries,
                                        std::uint32_t const* ranks,
                                        size_type count,
                                        hipStream_t stream)
{
    if (state.serve.counters && count == 0)
    {
        // Admission from the serving queue, decided on the device
        launch_kernel("serve-admit", serve_admit_kernel, 1, wave::size, 0, stream, state);
        launch_kernel("serve-primaries", serve_primaries_kernel, serve_blocks, block_size, 0,
                      stream, state);
        return hipGetLastError();
    }
    primaries_prepare_kernel<<<1, 1, 0, stream>>>(state.counters, count);
    if (count > 0)
    {
        process_primaries_kernel<<<num_blocks(count), block_size, 0, stream>>>(
            state, primaries, ranks, count);
        if (ranks)
        {
            advance_track_counters_kernel<<<num_blocks(count), block_size, 0, stream>>>(
                state, primaries, ranks, ranks + count, count);
        }
    }
    return hipGetLastError();
}

hipError_t launch_initialize_tracks(CoreParamsRef const& params,
                                    CoreStateRef const& state,
                                    hipStream_t stream)
{
    unsigned int const nb = num_blocks(state.size);
    init_prepare_kernel<<<1, 1, 0, stream>>>(state.counters, state.size);
    if (params.scalars.track_order == TrackOrder::init_charge)
    {
        partition_count_kernel<<<nb, block_size, 0, stream>>>(params, state);
        partition_scan_kernel<<<1, 1024, 0, stream>>>(state, nb);
        partition_scatter_kernel<<<nb, block_size, 0, stream>>>(params, state);
        launch_kernel("init-tracks-charge", init_tracks_charge_kernel, nb, block_size, 0, stream,
                      params, state);
    }
    else
    {
        launch_kernel("init-tracks", init_tracks_kernel, nb, block_size, 0, stream, params, state);
    }
    return hipGetLastError();
}

hipError_t launch_extend_from_secondaries(CoreParamsRef const& params,
                                          CoreStateRef const& state,
                                          hipStream_t stream)
{
    unsigned int const nb = num_blocks(state.size);
    std::uint8_t const in_place = params.scalars.track_order
                                  != TrackOrder::init_charge;
    launch_kernel("locate-alive", locate_alive_kernel, nb, block_size, 0, stream, state, in_place);
    launch_kernel("scan-partials", scan_partials_kernel, 1, 1024, 0, stream, state, nb);
    launch_kernel("process-secondaries", process_secondaries_kernel, nb, block_size, 0, stream,
                  state, in_place);
    return hipGetLastError();
}

hipError_t launch_fill_u32(std::uint32_t* data,
                           size_type n,
                           std::uint32_t value,
                           hipStream_t stream)
{
    if (n > 0)
        fill_u32_kernel<<<num_blocks(n), block_size, 0, stream>>>(data, n, value);
    return hipGetLastError();
}

hipError_t launch_iota_u32(std::uint32_t* data, size_type n, hipStream_t stream)
{
    if (n > 0)
        iota_u32_kernel<<<num_blocks(n), block_size, 0, stream>>>(data, n);
    return hipGetLastError();
}

size_type extend_block_scratch_size(size_type num_slots)
{
    return 2 * num_blocks(num_slots);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
