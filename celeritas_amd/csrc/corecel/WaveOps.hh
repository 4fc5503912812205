//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/WaveOps.hh
//! Wave64 and workgroup primitives for gfx950 kernels.
//!
//! Compaction and scans are built from __ballot (a 64-bit lane mask),
//! mbcnt (count of set bits in lower lanes) and DPP-backed shuffles, so a
//! workgroup-wide exclusive scan costs one LDS round trip between waves.
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace celeritas_amd
{
namespace wave
{
//---------------------------------------------------------------------------//
inline constexpr int size = 64;

__device__ __forceinline__ int lane_id()
{
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

//! Lane mask of threads with pred true
__device__ __forceinline__ std::uint64_t ballot(bool pred)
{
    return __ballot(pred);
}

//! Number of set bits in lanes below the calling lane
__device__ __forceinline__ unsigned int prefix_count(std::uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi(
        static_cast<unsigned int>(mask >> 32),
        __builtin_amdgcn_mbcnt_lo(static_cast<unsigned int>(mask), 0u));
}

//! Add the number of currently active lanes to a global counter with one
//! atomic per wave (the lowest active lane issues it)
__device__ __forceinline__ void count_active_lanes(unsigned long long* counter)
{
    std::uint64_t const mask = __ballot(1);
    if (prefix_count(mask) == 0)
        atomicAdd(counter, static_cast<unsigned long long>(__popcll(mask)));
}

//! Add the wave's active lanes to one of \c Spread copies of a counter,
//! chosen by the workgroup index (the copies are summed when read)
template<unsigned int Spread>
__device__ __forceinline__ void count_active_lanes_spread(unsigned long long* copies)
{
    count_active_lanes(copies + (blockIdx.x % Spread));
}

//! Inclusive scan of a 32-bit value over the wave
__device__ __forceinline__ unsigned int inclusive_scan(unsigned int v)
{
    int const lane = lane_id();
#pragma unroll
    for (int offset = 1; offset < size; offset <<= 1)
    {
        unsigned int const n = __shfl_up(v, offset, size);
        if (lane >= offset)
            v += n;
    }
    return v;
}

//! Sum over the wave (result in every lane)
__device__ __forceinline__ unsigned int reduce_sum(unsigned int v)
{
#pragma unroll
    for (int offset = size / 2; offset > 0; offset >>= 1)
        v += __shfl_xor(v, offset, size);
    return v;
}

__device__ __forceinline__ double reduce_sum(double v)
{
#pragma unroll
    for (int offset = size / 2; offset > 0; offset >>= 1)
        v += __shfl_xor(v, offset, size);
    return v;
}

//---------------------------------------------------------------------------//
/*!
 * Workgroup exclusive scan of two counters at once.
 *
 * \tparam NW number of waves in the workgroup
 * \param lds scratch of 2*NW unsigned ints
 * \return exclusive prefixes (a, b) and the workgroup totals
 */
template<int NW>
struct BlockScan2
{
    unsigned int excl_a;
    unsigned int excl_b;
    unsigned int total_a;
    unsigned int total_b;

    __device__ __forceinline__ BlockScan2(unsigned int a,
                                          unsigned int b,
                                          unsigned int* lds)
    {
        int const lane = lane_id();
        int const w = threadIdx.x / size;
        unsigned int const ia = inclusive_scan(a);
        unsigned int const ib = inclusive_scan(b);
        if (lane == size - 1)
        {
            lds[w] = ia;
            lds[NW + w] = ib;
        }
        __syncthreads();
        unsigned int off_a = 0, off_b = 0, tot_a = 0, tot_b = 0;
#pragma unroll
        for (int i = 0; i < NW; ++i)
        {
            unsigned int const wa = lds[i];
            unsigned int const wb = lds[NW + i];
            if (i < w)
            {
                off_a += wa;
                off_b += wb;
            }
            tot_a += wa;
            tot_b += wb;
        }
        excl_a = off_a + ia - a;
        excl_b = off_b + ib - b;
        total_a = tot_a;
        total_b = tot_b;
        __syncthreads();
    }
};

//---------------------------------------------------------------------------//
}  // namespace wave
}  // namespace celeritas_amd
