//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file geo/GeoData.hh
//! Device data for single-unit surface-based CSG geometry (ORANGE "simple
//! unit" semantics: orange/univ/SimpleUnitTracker.hh, orange/OrangeData.hh).
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Types.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Surface types (a subset of orange/OrangeTypes.hh SurfaceType)
enum class SurfaceType : std::uint8_t
{
    px = 0,  //!< Plane normal to x: {x0}
    py,  //!< Plane normal to y: {y0}
    pz,  //!< Plane normal to z: {z0}
    cxc,  //!< Cylinder along x through origin: {r^2}
    cyc,  //!< Cylinder along y through origin: {r^2}
    czc,  //!< Cylinder along z through origin: {r^2}
    sc,  //!< Sphere at origin: {r^2}
    p,  //!< General plane n.x = d: {nx, ny, nz, d}
    size_
};

//! Number of stored reals per surface type
CELER_FORCEINLINE_FUNCTION int surface_data_size(SurfaceType t)
{
    return t == SurfaceType::p ? 4 : 1;
}

//! Logic tokens (postfix); values below lbegin are face indices
namespace logic
{
inline constexpr std::int32_t ltrue = -3;
inline constexpr std::int32_t lor = -4;
inline constexpr std::int32_t land = -5;
inline constexpr std::int32_t lnot = -6;
}  // namespace logic

//! Volume flags (orange/OrangeData.hh VolumeRecord::Flags)
enum VolumeFlags : std::uint8_t
{
    internal_surfaces = 1,  //!< Crossing a face may keep the track inside
    implicit_vol = 2,  //!< Background: enters any volume via its faces
    simple_safety = 4,  //!< All faces support the normal-distance safety
};

//! Maximum faces per volume: senses fit in one 32-bit mask
inline constexpr int max_faces = 32;
//! Maximum candidate intersections in a volume (2 per quadric)
inline constexpr int max_intersections = 2 * max_faces;

//---------------------------------------------------------------------------//
/*!
 * Node of the point-location tree (a bounding interval hierarchy, the
 * structure of orange/detail/BIHData.hh, stored depth-first).
 *
 * Each node carries the one-plane test of the edge that leads into it: a
 * point passes if its coordinate along \c axis is <= \c plane (left child:
 * the plane bounds the left volumes from above) or >= \c plane (right
 * child: it bounds the right volumes from below); axis 3 always passes
 * (the root). A point that passes an inner node continues with the next
 * node (its left child); otherwise, and after a leaf, it continues at
 * \c skip, the first node after the subtree. Leaves list \c num_vols
 * volumes in \c bih_vols from \c vol_begin, in ascending order.
 */
struct BihNode
{
    double plane;
    std::uint32_t skip;
    std::uint32_t vol_begin;
    std::uint16_t num_vols;  //!< 0 for an inner node
    std::uint8_t axis;
    std::uint8_t upper;  //!< 1: pass if x >= plane; 0: pass if x <= plane
    std::uint32_t pad_;
};
static_assert(sizeof(BihNode) == 24, "BihNode layout");

//---------------------------------------------------------------------------//
/*!
 * Immutable geometry data (device pointers).
 */
struct GeoParamsRef
{
    size_type num_surfaces{0};
    size_type num_volumes{0};
    //! Contiguous region of the surface, face, logic, box and connectivity
    //! tables (what a tracking kernel can stage in LDS)
    char const* hot_base{nullptr};
    std::uint32_t hot_bytes{0};

    // Surfaces
    std::uint8_t const* surf_type{nullptr};
    std::uint32_t const* surf_offset{nullptr};  //!< into surf_data
    double const* surf_data{nullptr};

    // Volumes
    std::uint32_t const* vol_face_begin{nullptr};  //!< into faces
    std::uint8_t const* vol_num_faces{nullptr};
    std::uint8_t const* vol_max_isect{nullptr};
    std::uint32_t const* faces{nullptr};  //!< global surface IDs
    std::uint32_t const* vol_logic_begin{nullptr};  //!< into logic
    std::uint16_t const* vol_logic_size{nullptr};
    std::int32_t const* logic{nullptr};
    std::uint8_t const* vol_flags{nullptr};
    double const* vol_bbox{nullptr};  //!< [lo xyz, hi xyz] per volume

    // Connectivity: volumes that have each surface as a face
    std::uint32_t const* conn_begin{nullptr};
    std::uint32_t const* conn_size{nullptr};
    std::uint32_t const* conn_vols{nullptr};

    // Point location: tree over the finite bounding boxes, then the
    // volumes with an unbounded box (bih_vols[inf_begin, +inf_count))
    BihNode const* bih_nodes{nullptr};
    std::uint32_t const* bih_vols{nullptr};
    std::uint32_t num_bih_nodes{0};
    std::uint32_t bih_inf_begin{0};
    std::uint32_t bih_inf_count{0};

    std::uint32_t background{invalid_id};  //!< Implicit fill volume
    std::uint32_t exterior{0};  //!< Volume that ends transport
};

//---------------------------------------------------------------------------//
//! Post-crossing boundary state (orange/OrangeTypes.hh BoundaryResult)
enum class BoundaryResult : std::uint8_t
{
    exiting = 0,
    reentrant = 1,
};

/*!
 * Mutable per-slot geometry state (SoA, device pointers).
 *
 * Position and direction stay double precision in every precision mode.
 */
struct GeoStateRef
{
    double* pos[3]{};
    double* dir[3]{};
    std::uint32_t* volume{nullptr};
    std::uint32_t* surface{nullptr};  //!< current surface (on boundary)
    std::uint8_t* sense{nullptr};  //!< sense of current surface (pre-cross)
    std::uint8_t* boundary{nullptr};  //!< BoundaryResult
    std::uint8_t* failed{nullptr};
    double* next_step{nullptr};
    std::uint32_t* next_surface{nullptr};
    std::uint8_t* next_sense{nullptr};
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
