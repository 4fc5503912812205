//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/StagedTables.hh
//! LDS staging of the physics and geometry tables a step kernel reads per
//! track.
//!
//! The particle and process records, grid headers and grid values, and the
//! material, MSC and fluctuation records live in one contiguous device
//! region (PhysicsParamsRef::hot_base, hot_bytes), and so do the surface,
//! face, volume logic, bounding box and connectivity tables of the geometry
//! (GeoParamsRef::hot_base, hot_bytes). A kernel launched with enough
//! dynamic LDS copies the regions it reads once per workgroup and continues
//! with a params view whose table pointers are rebased into LDS, so the
//! chains of dependent table reads of a track (particle -> process -> grid
//! header -> values; volume -> faces -> surfaces -> coefficients) are served
//! from LDS instead of L2. The data are the same bytes, so results are
//! bit-identical to the unstaged path.
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "corecel/WaveOps.hh"
#include "CoreData.hh"
#include "StepExecutors.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Largest region staged: above it a kernel reads the tables in place (at
//! 32 KiB five 256-thread workgroups still fit a CU's LDS)
inline constexpr unsigned int max_staged_table_bytes = 32 * 1024;
//! Largest physics and geometry regions staged together (three 256-thread
//! workgroups per CU, the register limit of the kernel that stages both)
inline constexpr unsigned int max_staged_both_bytes = 52 * 1024;

//! Dynamic LDS bytes of a launch staging the physics tables (0: the tables
//! stay in HBM)
inline unsigned int staged_table_bytes(CoreParamsRef const& p)
{
    unsigned int const n = p.phys.hot_bytes;
    return (n > 0 && n <= max_staged_table_bytes) ? n : 0;
}

//! Dynamic LDS bytes of a launch staging the geometry tables
inline unsigned int staged_geo_bytes(CoreParamsRef const& p)
{
    unsigned int const n = p.geo.hot_bytes;
    return (n > 0 && n <= max_staged_table_bytes) ? n : 0;
}

//! Dynamic LDS bytes of a launch staging both (physics first)
inline unsigned int staged_both_bytes(CoreParamsRef const& p)
{
    unsigned int const a = staged_table_bytes(p);
    unsigned int const b = staged_geo_bytes(p);
    return (a > 0 && b > 0 && a + b <= max_staged_both_bytes) ? a + b : 0;
}

//! What a launch stages in LDS
enum class Staging
{
    none,
    physics,
    geometry,
    both,
};

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

//! Copy a 16-byte-aligned region into LDS (no barrier)
__device__ __forceinline__ void copy_region(char const* from, unsigned int n, char* to)
{
    auto const* src = reinterpret_cast<uint4 const*>(from);
    auto* dst = reinterpret_cast<uint4*>(to);
    for (unsigned int i = threadIdx.x; i < n / 16; i += blockDim.x)
        dst[i] = src[i];
}

//! Physics table pointers rebased into LDS
__device__ __forceinline__ void rebase_physics(PhysicsParamsRef* r, char* to)
{
    PhysicsParamsRef const pp = *r;
    char const* from = pp.hot_base;
    r->particle = rebase(pp.particle, from, to);
    r->grids = rebase(pp.grids, from, to);
    r->grid_reals = rebase(pp.grid_reals, from, to);
    r->mass = rebase(pp.mass, from, to);
    r->charge = rebase(pp.charge, from, to);
    r->materials = rebase(pp.materials, from, to);
    r->msc_material = rebase(pp.msc_material, from, to);
    r->msc_par_mat = rebase(pp.msc_par_mat, from, to);
    r->fluct = rebase(pp.fluct, from, to);
    r->model_class = rebase(pp.model_class, from, to);
}

//! Geometry table pointers rebased into LDS
__device__ __forceinline__ void rebase_geometry(GeoParamsRef* r, char* to)
{
    GeoParamsRef const g = *r;
    char const* from = g.hot_base;
    r->surf_type = rebase(g.surf_type, from, to);
    r->surf_offset = rebase(g.surf_offset, from, to);
    r->surf_data = rebase(g.surf_data, from, to);
    r->vol_face_begin = rebase(g.vol_face_begin, from, to);
    r->vol_num_faces = rebase(g.vol_num_faces, from, to);
    r->vol_max_isect = rebase(g.vol_max_isect, from, to);
    r->faces = rebase(g.faces, from, to);
    r->vol_logic_begin = rebase(g.vol_logic_begin, from, to);
    r->vol_logic_size = rebase(g.vol_logic_size, from, to);
    r->logic = rebase(g.logic, from, to);
    r->vol_flags = rebase(g.vol_flags, from, to);
    r->vol_bbox = rebase(g.vol_bbox, from, to);
    r->conn_begin = rebase(g.conn_begin, from, to);
    r->conn_size = rebase(g.conn_size, from, to);
    r->conn_vols = rebase(g.conn_vols, from, to);
}
}  // namespace detail

/*!
 * Stage the tables a launch reads per track into its dynamic LDS and
 * return the params with those tables' pointers in LDS (the physics region
 * first, then the geometry region).
 *
 * Every thread of the workgroup must call this before any of them returns
 * (it contains a workgroup barrier). The launcher gives the kernel the
 * LDS its staging needs (staged_table_bytes, staged_geo_bytes,
 * staged_both_bytes) and instantiates it with that staging only when the
 * regions fit: the table pointers then derive from the LDS array alone,
 * so the compiler reads them with LDS instructions and 32-bit addresses.
 */
template<Staging S>
__device__ __forceinline__ CoreParamsRef stage(CoreParamsRef const& params, void* lds)
{
    if constexpr (S == Staging::none)
    {
        return params;
    }
    else
    {
        CoreParamsRef r = params;
        char* to = static_cast<char*>(lds);
        unsigned int offset = 0;
        if constexpr (S == Staging::physics || S == Staging::both)
        {
            detail::copy_region(params.phys.hot_base, params.phys.hot_bytes, to);
            detail::rebase_physics(&r.phys, to);
            offset = params.phys.hot_bytes;
        }
        if constexpr (S == Staging::geometry || S == Staging::both)
        {
            detail::copy_region(params.geo.hot_base, params.geo.hot_bytes, to + offset);
            detail::rebase_geometry(&r.geo, to + offset);
        }
        __syncthreads();
        return r;
    }
}

/*!
 * Discrete selection at the end of an along-step kernel (when the params
 * fuse it): the track's step ended in the discrete-select action, so the
 * selection that action would make next draws from the same RNG stream in
 * the same order, and the results are those of a separate launch.
 */
inline namespace CELER_PREC_NS
{
template<bool Charged>
__device__ __forceinline__ void fused_discrete_select(CoreParamsRef const& params,
                                                      CoreStateRef const& state,
                                                      size_type slot)
{
    if (!params.scalars.fuse_discrete_select
        || state.sim.post_step_action[slot] != params.phys.scalars.discrete_action)
        return;
    if (state.diag)
    {
        wave::count_active_lanes_spread<diag_spread>(state.diag + diag_discrete * diag_spread);
        if (Charged)
        {
            wave::count_active_lanes_spread<diag_spread>(
                state.diag + diag_discrete_charged * diag_spread);
        }
    }
    exec::discrete_select_track(params, state, slot);
}
}  // namespace CELER_PREC_NS

//! Stage the physics tables when \c Staged (see stage)
template<bool Staged>
__device__ __forceinline__ CoreParamsRef stage_tables(CoreParamsRef const& params,
                                                      void* lds)
{
    return stage<Staged ? Staging::physics : Staging::none>(params, lds);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
