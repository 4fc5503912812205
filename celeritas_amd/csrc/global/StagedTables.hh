//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/StagedTables.hh
//! LDS staging of the physics tables a step kernel reads per track.
//!
//! The particle and process records, grid headers and grid values, and the
//! material, MSC and fluctuation records live in one contiguous device
//! region (PhysicsParamsRef::hot_base, hot_bytes). A kernel launched with
//! that many bytes of dynamic LDS copies the region once per workgroup and
//! continues with a params view whose table pointers are rebased into LDS,
//! so the chains of dependent table reads of a track (particle -> process
//! -> grid header -> values) are served from LDS instead of L2. The data
//! are the same bytes, so results are bit-identical to the unstaged path.
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "CoreData.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Largest region staged: above it a kernel reads the tables in place (at
//! 32 KiB five 256-thread workgroups still fit a CU's LDS)
inline constexpr unsigned int max_staged_table_bytes = 32 * 1024;

//! Dynamic LDS bytes of a staged launch (0: the tables stay in HBM)
inline unsigned int staged_table_bytes(CoreParamsRef const& p)
{
    unsigned int const n = p.phys.hot_bytes;
    return (n > 0 && n <= max_staged_table_bytes) ? n : 0;
}

/*!
 * Whether any thread of this workgroup maps to a slot.
 *
 * A sorted launch covers all slots but its range is read on the device, so
 * most of its workgroups can be empty: they return before staging (the
 * answer is uniform over the workgroup, so no barrier is left waiting).
 */
__device__ __forceinline__ bool workgroup_has_work(ThreadSlots const& ts,
                                                   size_type size)
{
    size_type const first = blockIdx.x * blockDim.x;
    if (ts.range)
    {
        size_type const end = ts.range_end ? *ts.range_end : ts.range[1];
        return first < end - ts.range[0];
    }
    return first < size;
}

namespace detail
{
template<class T>
__device__ __forceinline__ T const* rebase(T const* p, char const* from, char* to)
{
    return p ? reinterpret_cast<T const*>(to + (reinterpret_cast<char const*>(p) - from))
             : p;
}
}  // namespace detail

/*!
 * Stage the hot tables into the kernel's dynamic LDS and return the params
 * with their table pointers in LDS; without staging, the params unchanged.
 *
 * Every thread of the workgroup must call this before any of them returns
 * (it contains a workgroup barrier).
 *
 * A kernel instantiated with \c Staged true always stages (its launcher
 * checks \c staged_table_bytes first): its table pointers then derive from
 * the LDS array alone, so the compiler reads the tables with LDS
 * instructions and 32-bit addresses rather than generic 64-bit ones.
 */
template<bool Staged>
__device__ __forceinline__ CoreParamsRef stage_tables(CoreParamsRef const& params,
                                                      void* lds)
{
    if constexpr (!Staged)
        return params;
    PhysicsParamsRef const& pp = params.phys;
    unsigned int const n = pp.hot_bytes;
    auto const* src = reinterpret_cast<uint4 const*>(pp.hot_base);
    auto* dst = static_cast<uint4*>(lds);
    for (unsigned int i = threadIdx.x; i < n / 16; i += blockDim.x)
        dst[i] = src[i];
    __syncthreads();

    CoreParamsRef r = params;
    char const* from = pp.hot_base;
    char* to = static_cast<char*>(lds);
    r.phys.particle = detail::rebase(pp.particle, from, to);
    r.phys.grids = detail::rebase(pp.grids, from, to);
    r.phys.grid_reals = detail::rebase(pp.grid_reals, from, to);
    r.phys.mass = detail::rebase(pp.mass, from, to);
    r.phys.charge = detail::rebase(pp.charge, from, to);
    r.phys.materials = detail::rebase(pp.materials, from, to);
    r.phys.msc_material = detail::rebase(pp.msc_material, from, to);
    r.phys.msc_par_mat = detail::rebase(pp.msc_par_mat, from, to);
    r.phys.fluct = detail::rebase(pp.fluct, from, to);
    r.phys.model_class = detail::rebase(pp.model_class, from, to);
    return r;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
