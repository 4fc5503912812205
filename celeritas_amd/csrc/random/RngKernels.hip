//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file random/RngKernels.hip
//! Per-event reseed (celeritas/random/RngReseed.cu:29-72) and test draws.
//---------------------------------------------------------------------------//
#include "track/TrackKernels.hh"

namespace celeritas_amd
{
namespace
{
constexpr int block_size = 256;

//---------------------------------------------------------------------------//
/*!
 * Reseed: SplitMix64 on the seed, then skip event*N + slot subsequences.
 *
 * The jump tables are staged in LDS once per workgroup (2.5 KiB) because
 * every lane walks them with lane-dependent digit counts.
 */
__global__ void __launch_bounds__(block_size)
    reseed_kernel(RngStateRef const state,
                  XorwowJumpTables const* __restrict__ tables,
                  std::uint32_t seed,
                  ull_int event_id)
{
    __shared__ XorwowJumpTables lds;
    {
        auto const* src = reinterpret_cast<std::uint32_t const*>(tables);
        auto* dst = reinterpret_cast<std::uint32_t*>(&lds);
        constexpr int nwords = sizeof(XorwowJumpTables) / 4;
        for (int i = threadIdx.x; i < nwords; i += block_size)
            dst[i] = src[i];
    }
    __syncthreads();
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    if (slot >= state.size)
        return;
    XorwowInitializer init;
    init.seed = seed;
    init.subsequence = event_id * state.size + slot;
    init.offset = 0;
    XorwowEngine e;
    e.initialize(init, lds);
    e.store(state, slot);
}

//! Initialize with explicit per-slot subsequence and offset
__global__ void __launch_bounds__(block_size)
    initialize_kernel(RngStateRef const state,
                      XorwowJumpTables const* __restrict__ tables,
                      std::uint32_t seed,
                      ull_int const* __restrict__ subsequence,
                      ull_int const* __restrict__ offset)
{
    __shared__ XorwowJumpTables lds;
    {
        auto const* src = reinterpret_cast<std::uint32_t const*>(tables);
        auto* dst = reinterpret_cast<std::uint32_t*>(&lds);
        constexpr int nwords = sizeof(XorwowJumpTables) / 4;
        for (int i = threadIdx.x; i < nwords; i += block_size)
            dst[i] = src[i];
    }
    __syncthreads();
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    if (slot >= state.size)
        return;
    XorwowInitializer init;
    init.seed = seed;
    init.subsequence = subsequence[slot];
    init.offset = offset[slot];
    XorwowEngine e;
    e.initialize(init, lds);
    e.store(state, slot);
}

//! Draw values per slot: kind 0 = u32, 1 = double, 2 = float
__global__ void __launch_bounds__(block_size)
    draw_kernel(RngStateRef const state, size_type draws, int kind, void* out)
{
    size_type const slot = blockIdx.x * block_size + threadIdx.x;
    if (slot >= state.size)
        return;
    XorwowEngine e = XorwowEngine::load(state, slot);
    for (size_type i = 0; i < draws; ++i)
    {
        std::size_t const idx = std::size_t(slot) * draws + i;
        if (kind == 0)
            static_cast<std::uint32_t*>(out)[idx] = e();
        else if (kind == 1)
            static_cast<double*>(out)[idx] = generate_canonical<double>(e);
        else
            static_cast<float*>(out)[idx] = generate_canonical<float>(e);
    }
    e.store(state, slot);
}

//---------------------------------------------------------------------------//
}  // namespace

hipError_t launch_reseed_rng(RngStateRef const& state,
                             XorwowJumpTables const* tables,
                             std::uint32_t seed,
                             ull_int event_id,
                             hipStream_t stream)
{
    unsigned int const nb = (state.size + block_size - 1) / block_size;
    reseed_kernel<<<nb, block_size, 0, stream>>>(state, tables, seed, event_id);
    return hipGetLastError();
}

hipError_t launch_rng_initialize(RngStateRef const& state,
                                 XorwowJumpTables const* tables,
                                 std::uint32_t seed,
                                 ull_int const* subsequence,
                                 ull_int const* offset,
                                 hipStream_t stream)
{
    unsigned int const nb = (state.size + block_size - 1) / block_size;
    initialize_kernel<<<nb, block_size, 0, stream>>>(
        state, tables, seed, subsequence, offset);
    return hipGetLastError();
}

hipError_t launch_rng_draw(RngStateRef const& state,
                           size_type draws_per_slot,
                           int kind,
                           void* out,
                           hipStream_t stream)
{
    unsigned int const nb = (state.size + block_size - 1) / block_size;
    draw_kernel<<<nb, block_size, 0, stream>>>(state, draws_per_slot, kind, out);
    return hipGetLastError();
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
