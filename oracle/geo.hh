//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/geo.hh
//! One-level ORANGE simple unit, restated from orange/OrangeTrackView.hh
//! (initialize :266-351, find_next_step :531-575, move :580-650,
//! cross_boundary :662-775, set_dir :789-840, find_safety :1007-1025),
//! orange/univ/SimpleUnitTracker.hh (:161-575), univ/detail/SenseCalculator,
//! LogicEvaluator, SurfaceFunctors and orange/surf/{PlaneAligned,
//! CylCentered,SphereCentered,Plane}.hh with surf/detail/QuadraticSolver.hh.
//!
//! Senses are kept in a temporary vector and intersections are sorted with
//! std::stable_sort, as in the reference's temp_sense / temp_next buffers.
//!
//! Pinned by: ray traces through testem3-flat (layer widths 0.23/0.57 cm,
//! OrangeGeant.test.cc:102-160), the TestEM3 safety (OrangeJson.test.cc:
//! 648-660) and closed-form plane/cylinder distances.
//---------------------------------------------------------------------------//
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "json.hh"

namespace orc
{
//---------------------------------------------------------------------------//
constexpr std::uint32_t none = 0xffffffffu;

struct Surface
{
    std::string type;  // px py pz cxc cyc czc sc p
    std::vector<double> d;
};

struct Volume
{
    std::vector<std::uint32_t> faces;
    std::vector<int> logic;  // >= 0 face index; -3 true, -4 or, -5 and, -6 not
    int flags = 0;  // 1 internal surfaces, 2 implicit
    bool simple_safety = true;
    std::string label;
};

struct Geometry
{
    std::vector<Surface> surfaces;
    std::vector<Volume> volumes;
    std::vector<std::vector<std::uint32_t>> neighbors;  // per surface
    std::uint32_t background = none;

    static Geometry from_orange_json(Json const& doc);
};

struct GeoState
{
    double pos[3];
    double dir[3];
    std::uint32_t vol = none;
    std::uint32_t surf = none;  // on this surface (if any)
    bool sense = false;
    bool reentrant = false;
    bool failed = false;
    double next_step = 0;
    std::uint32_t next_surf = none;
    bool next_sense = false;
};

// Tracking operations
void geo_initialize(Geometry const& g, GeoState& s, double const* pos, double const* dir);
//! Distance to boundary (max_step <= 0 means unlimited)
double geo_find_next_step(Geometry const& g, GeoState& s, double max_step, bool* boundary);
void geo_move_to_boundary(GeoState& s);
void geo_move_internal(GeoState& s, double dist);
void geo_move_internal_to(GeoState& s, double const* pos);
void geo_cross_boundary(Geometry const& g, GeoState& s);
void geo_set_dir(Geometry const& g, GeoState& s, double const* dir);
double geo_find_safety(Geometry const& g, GeoState const& s);
inline bool geo_is_outside(GeoState const& s)
{
    return s.vol == 0;
}

//---------------------------------------------------------------------------//
}  // namespace orc
