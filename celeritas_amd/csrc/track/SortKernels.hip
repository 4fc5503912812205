//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file track/SortKernels.hip
//! Reindexing of track slots by a small key (status, particle type, along-
//! step or post-step action) and the per-key thread offsets.
//!
//! Counterpart of sort_tracks / count_tracks_per_action
//! (celeritas/track/detail/TrackSortUtils.cu:145-206, host
//! TrackSortUtils.cc:71-157). The reference sorts the track_slots array
//! with thrust::sort_by_key / std::sort (unstable) and then copies the
//! per-action offsets to the host. Here the keys are few (at most
//! max_sort_keys) so the sort is a three-pass stable multi-split over the
//! slot index: per-tile key histograms in LDS, an exclusive scan of each
//! key's row of the key-major histogram table (one workgroup per key), and
//! a scatter that places the key ranges from the key totals and ranks equal
//! keys inside each wave with ballots. Tiles of 1,024 indices give a 2^20
//! slot sort 1,024 workgroups, four waves per SIMD. The result
//! lists, for each key, its slots in increasing order, so the gathers of a
//! sorted launch stay monotonic; the offsets stay on the device and sorted
//! kernels read their range from there, so no host round trip is needed.
//---------------------------------------------------------------------------//
#include "SortKernels.hh"

#include <mutex>

#include "corecel/WaveOps.hh"

namespace celeritas_amd
{
namespace
{
//---------------------------------------------------------------------------//
constexpr int block_size = 256;
constexpr int num_waves = block_size / wave::size;
constexpr int items_per_thread = 4;
constexpr int tile_size = block_size * items_per_thread;

inline unsigned int num_tiles(size_type n)
{
    return (n + tile_size - 1) / tile_size;
}

//! Sort key of a slot from the track state; inactive slots take the last key
struct StateKey
{
    CoreStateRef state;
    SortKey kind;
    //! Along-step action per particle type (along_action_of_particle)
    std::uint8_t const* along_of_particle{nullptr};

    __device__ __forceinline__ unsigned int operator()(size_type slot, unsigned int last) const
    {
        auto const status = static_cast<TrackStatus>(state.sim.status[slot]);
        if (status == TrackStatus::inactive)
            return last;
        unsigned int k;
        switch (kind)
        {
            case SortKey::status:
                return 0;
            case SortKey::particle:
                k = state.particle.particle_id[slot];
                break;
            case SortKey::along_action:
                k = state.sim.along_step_action[slot];
                break;
            case SortKey::along_action_of_particle:
                // The pre-step leaves an errored track's action unset
                // (invalid) and gives others their particle's
                if (status == TrackStatus::errored || !along_of_particle)
                    return last;
                k = along_of_particle[state.particle.particle_id[slot]];
                break;
            default:
                k = state.sim.post_step_action[slot];
                break;
        }
        return k < last ? k : last;
    }
};

//! Key of an index from a caller's byte array, clamped to the last key
struct ArrayKey
{
    std::uint8_t const* keys;

    __device__ __forceinline__ unsigned int operator()(size_type i, unsigned int last) const
    {
        unsigned int const k = keys[i];
        return k < last ? k : last;
    }
};

//! Where a multi-split reads and writes: n indices, the key-major per-tile
//! histogram table (scratch, num_keys * ntiles words), the sorted indices
//! and the per-key offsets (num_keys + 1 words)
struct SplitRefs
{
    size_type n;
    std::uint32_t* counts;
    std::uint32_t* sorted;
    std::uint32_t* offsets;
};

//---------------------------------------------------------------------------//
//! Per-tile key histogram, stored key-major: counts[key * ntiles + tile]
template<class KeyFn>
__global__ void __launch_bounds__(block_size)
    sort_count_kernel(KeyFn const key_of,
                      SplitRefs const refs,
                      unsigned int const num_keys,
                      unsigned int const ntiles)
{
    __shared__ unsigned int hist[max_sort_keys];
    for (unsigned int k = threadIdx.x; k < num_keys; k += block_size)
        hist[k] = 0;
    __syncthreads();
    size_type const base = blockIdx.x * tile_size;
    for (int i = 0; i < items_per_thread; ++i)
    {
        size_type const idx = base + i * block_size + threadIdx.x;
        if (idx < refs.n)
            atomicAdd(&hist[key_of(idx, num_keys - 1)], 1u);
    }
    __syncthreads();
    for (unsigned int k = threadIdx.x; k < num_keys; k += block_size)
        refs.counts[k * ntiles + blockIdx.x] = hist[k];
}

/*!
 * Exclusive scan of one key's tile counts (workgroup blockIdx.x = key): the
 * counts become the key's thread offsets within its range, and the key's
 * total goes to the totals row after the table (counts[num_keys * ntiles
 * + key]). The keys' ranges are placed by the scatter kernel, which sums
 * the totals itself, so no single workgroup walks the whole table.
 */
__global__ void __launch_bounds__(block_size)
    sort_scan_kernel(SplitRefs const refs,
                     unsigned int const num_keys,
                     unsigned int const ntiles)
{
    __shared__ unsigned int lds[2 * num_waves];
    __shared__ unsigned int carry;
    unsigned int const key = blockIdx.x;
    if (ntiles == 0)
    {
        // Nothing to split (no scatter follows): every key's range is empty
        if (threadIdx.x == 0)
        {
            refs.offsets[key] = 0;
            if (key == 0)
                refs.offsets[num_keys] = 0;
        }
        return;
    }
    if (threadIdx.x == 0)
        carry = 0;
    __syncthreads();
    unsigned int* row = refs.counts + std::size_t(key) * ntiles;
    // Runs of four tiles per thread, read with independent loads
    constexpr unsigned int held = 4;
    for (unsigned int first = 0; first < ntiles; first += block_size * held)
    {
        unsigned int const begin = first + threadIdx.x * held;
        unsigned int v[held];
        unsigned int sum = 0;
#pragma unroll
        for (unsigned int u = 0; u < held; ++u)
        {
            v[u] = begin + u < ntiles ? row[begin + u] : 0u;
            sum += v[u];
        }
        wave::BlockScan2<num_waves> scan(sum, 0u, lds);
        unsigned int running = carry + scan.excl_a;
#pragma unroll
        for (unsigned int u = 0; u < held; ++u)
        {
            if (begin + u < ntiles)
                row[begin + u] = running;
            running += v[u];
        }
        __syncthreads();
        if (threadIdx.x == 0)
            carry += scan.total_a;
        __syncthreads();
    }
    if (threadIdx.x == 0)
        refs.counts[std::size_t(num_keys) * ntiles + key] = carry;
}

/*!
 * Stable scatter of indices to their sorted thread positions.
 *
 * A tile is processed in rounds of one index per thread, in index order. In
 * a round, each wave splits its lanes by key: the lowest remaining lane's
 * key is broadcast, the lanes holding it are found with one ballot and
 * ranked with mbcnt, and they leave the loop (one iteration per distinct
 * key in the wave). Wave totals per key go through LDS to order the waves;
 * a per-key running base carries the order across rounds and tiles.
 */
template<class KeyFn>
__global__ void __launch_bounds__(block_size)
    sort_scatter_kernel(KeyFn const key_of,
                        SplitRefs const refs,
                        unsigned int const num_keys,
                        unsigned int const ntiles)
{
    __shared__ unsigned int base[max_sort_keys];
    __shared__ unsigned int wcount[num_waves][max_sort_keys];
    int const lane = wave::lane_id();
    int const w = threadIdx.x / wave::size;
    // Key ranges from the key totals (one wave scans all of them); the
    // first tile also publishes them as the per-key offsets
    static_assert(max_sort_keys <= wave::size, "one wave scans the key totals");
    if (threadIdx.x < wave::size)
    {
        unsigned int const k = threadIdx.x;
        unsigned int const t
            = k < num_keys ? refs.counts[std::size_t(num_keys) * ntiles + k] : 0u;
        unsigned int const excl = wave::inclusive_scan(t) - t;
        if (k < num_keys)
        {
            base[k] = excl + refs.counts[k * ntiles + blockIdx.x];
            if (blockIdx.x == 0)
                refs.offsets[k] = excl;
        }
        if (blockIdx.x == 0 && k == 0)
            refs.offsets[num_keys] = refs.n;
    }
    size_type const tile_base = blockIdx.x * tile_size;
    for (int r = 0; r < items_per_thread; ++r)
    {
        for (int k = threadIdx.x; k < num_waves * max_sort_keys; k += block_size)
            (&wcount[0][0])[k] = 0;
        __syncthreads();

        size_type const idx = tile_base + r * block_size + threadIdx.x;
        bool const valid = idx < refs.n;
        unsigned int const key = valid ? key_of(idx, num_keys - 1) : 0u;
        unsigned int rank = 0;
        std::uint64_t remaining = wave::ballot(valid);
        while (remaining)
        {
            int const leader = __builtin_ctzll(remaining);
            unsigned int const k = __shfl(key, leader, wave::size);
            bool const mine = valid && key == k;
            std::uint64_t const m = wave::ballot(mine);
            if (mine)
                rank = wave::prefix_count(m);
            if (lane == leader)
                wcount[w][k] = static_cast<unsigned int>(__popcll(m));
            remaining &= ~m;
        }
        __syncthreads();

        if (valid)
        {
            unsigned int pos = base[key] + rank;
            for (int v = 0; v < w; ++v)
                pos += wcount[v][key];
            refs.sorted[pos] = idx;
        }
        __syncthreads();
        for (unsigned int k = threadIdx.x; k < num_keys; k += block_size)
        {
            unsigned int t = 0;
            for (int v = 0; v < num_waves; ++v)
                t += wcount[v][k];
            base[k] += t;
        }
        __syncthreads();
    }
}

template<class KeyFn>
hipError_t launch_split(KeyFn const& key_of,
                        SplitRefs const& refs,
                        unsigned int num_keys,
                        hipStream_t stream)
{
    if (num_keys < 1 || num_keys > max_sort_keys)
        return hipErrorInvalidValue;
    unsigned int const nt = num_tiles(refs.n);
    if (nt > 0)
        sort_count_kernel<<<nt, block_size, 0, stream>>>(key_of, refs, num_keys, nt);
    sort_scan_kernel<<<num_keys, block_size, 0, stream>>>(refs, num_keys, nt);
    if (nt > 0)
        sort_scatter_kernel<<<nt, block_size, 0, stream>>>(key_of, refs, num_keys, nt);
    return hipGetLastError();
}

//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
hipError_t launch_sort_tracks(CoreStateRef const& state,
                              SortKey kind,
                              unsigned int num_keys,
                              hipStream_t stream,
                              std::uint8_t const* along_of_particle)
{
    return launch_split(StateKey{state, kind, along_of_particle},
                        SplitRefs{state.size, state.sort_counts, state.track_slots,
                                  state.action_offsets},
                        num_keys, stream);
}

hipError_t launch_partition_keys(std::uint8_t const* keys,
                                 size_type n,
                                 unsigned int num_keys,
                                 std::uint32_t* counts,
                                 std::uint32_t* sorted,
                                 std::uint32_t* offsets,
                                 hipStream_t stream)
{
    return launch_split(ArrayKey{keys}, SplitRefs{n, counts, sorted, offsets}, num_keys,
                        stream);
}

//---------------------------------------------------------------------------//
// ALLOCATION-FREE PARTITION SCRATCH
//---------------------------------------------------------------------------//
namespace
{
//! Histogram table of the allocation-free partition: max_sort_keys words
//! per tile for up to max_partition_noalloc indices plus a row of key
//! totals (1 MiB + 256 B)
constexpr size_type noalloc_scratch_words
    = max_sort_keys * ((max_partition_noalloc + tile_size - 1) / tile_size + 1);

__device__ __align__(256) std::uint32_t scratch[noalloc_scratch_words];

constexpr int max_devices = 64;
std::mutex scratch_lock;
std::uint32_t* scratch_ptr[max_devices] = {};
}  // namespace

hipError_t partition_noalloc_scratch(std::uint32_t** ptr)
{
    int device = 0;
    hipError_t err = hipGetDevice(&device);
    if (err != hipSuccess)
        return err;
    if (device < 0 || device >= max_devices)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> guard(scratch_lock);
    if (!scratch_ptr[device])
    {
        // The array's name is only a host-side stand-in: its address on
        // this device comes from the runtime, once (each device holds its
        // own copy of the array)
        void* addr = nullptr;
        err = hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch));
        if (err != hipSuccess)
            return err;
        scratch_ptr[device] = static_cast<std::uint32_t*>(addr);
    }
    *ptr = scratch_ptr[device];
    return hipSuccess;
}

size_type sort_counts_size(size_type num_slots)
{
    // Per-tile histograms plus the row of key totals
    return max_sort_keys * (num_tiles(num_slots) + 1);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
