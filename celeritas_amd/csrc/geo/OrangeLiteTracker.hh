//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file geo/OrangeLiteTracker.hh
//! Register-resident track view for single-unit CSG geometry.
//!
//! Semantics follow ORANGE's simple unit tracker for one level:
//! - initialize: orange/univ/SimpleUnitTracker.hh:161-190 (position search,
//!   starting on a surface is a failure)
//! - find_next_step: orange/OrangeTrackView.hh:555-575 and
//!   SimpleUnitTracker.hh:271-285,392-575 (simple and complex volumes)
//! - move_to_boundary / move_internal: OrangeTrackView.hh:580-650
//! - cross_boundary: OrangeTrackView.hh:662-775, SimpleUnitTracker.hh:196-255
//! - set_dir: OrangeTrackView.hh:789-840; find_safety:
//!   SimpleUnitTracker.hh:299-318
//! Surface math: orange/surf/PlaneAligned.hh, CylCentered.hh,
//! SphereCentered.hh, Plane.hh and surf/detail/QuadraticSolver.hh.
//!
//! Unlike the reference, senses for a volume are a 32-bit mask in a register
//! and the logic stack is a 64-bit register, so no per-slot scratch memory
//! is read or written during tracking.
//---------------------------------------------------------------------------//
#pragma once

#include <cmath>

#include "GeoData.hh"

namespace celeritas_amd
{
namespace geo
{
//---------------------------------------------------------------------------//
inline constexpr double no_intersection = HUGE_VAL;
//! Degenerate-quadratic tolerance (orange/OrangeTypes.hh:324-330, double)
inline constexpr double sqrt_quadratic = 1e-5;

//! Signed sense: -1 inside, 0 on, +1 outside
CELER_FORCEINLINE_FUNCTION int real_to_sense(double q)
{
    return static_cast<int>(!(q <= 0)) - static_cast<int>(q < 0);
}

//---------------------------------------------------------------------------//
//! Track state held in registers by a kernel
struct GeoTrack
{
    double pos[3];
    double dir[3];
    std::uint32_t volume;
    std::uint32_t surface;
    std::uint8_t sense;
    std::uint8_t boundary;
    std::uint8_t failed;
    double next_step;
    std::uint32_t next_surface;
    std::uint8_t next_sense;

    CELER_FORCEINLINE_FUNCTION bool is_on_boundary() const
    {
        return surface != invalid_id;
    }
    CELER_FORCEINLINE_FUNCTION bool has_next_step() const
    {
        return next_step != 0;
    }
    CELER_FORCEINLINE_FUNCTION void clear_next()
    {
        next_step = 0;
        next_surface = invalid_id;
    }

    CELER_FORCEINLINE_FUNCTION static GeoTrack
    load(GeoStateRef const& s, size_type i)
    {
        GeoTrack g;
        for (int k = 0; k < 3; ++k)
        {
            g.pos[k] = s.pos[k][i];
            g.dir[k] = s.dir[k][i];
        }
        g.volume = s.volume[i];
        g.surface = s.surface[i];
        g.sense = s.sense[i];
        g.boundary = s.boundary[i];
        g.failed = s.failed[i];
        g.next_step = s.next_step[i];
        g.next_surface = s.next_surface[i];
        g.next_sense = s.next_sense[i];
        return g;
    }
    CELER_FORCEINLINE_FUNCTION void store(GeoStateRef const& s,
                                          size_type i) const
    {
        for (int k = 0; k < 3; ++k)
        {
            s.pos[k][i] = pos[k];
            s.dir[k][i] = dir[k];
        }
        s.volume[i] = volume;
        s.surface[i] = surface;
        s.sense[i] = sense;
        s.boundary[i] = boundary;
        s.failed[i] = failed;
        s.next_step[i] = next_step;
        s.next_surface[i] = next_surface;
        s.next_sense[i] = next_sense;
    }
};

//---------------------------------------------------------------------------//
// SURFACE EVALUATION
//---------------------------------------------------------------------------//
//! Quadric value whose sign gives the sense
CELER_FUNCTION double
eval_surface(GeoParamsRef const& p, std::uint32_t s, double const* x)
{
    double const* d = p.surf_data + p.surf_offset[s];
    switch (static_cast<SurfaceType>(p.surf_type[s]))
    {
        case SurfaceType::px:
            return x[0] - d[0];
        case SurfaceType::py:
            return x[1] - d[0];
        case SurfaceType::pz:
            return x[2] - d[0];
        case SurfaceType::cxc:
            return x[1] * x[1] + x[2] * x[2] - d[0];
        case SurfaceType::cyc:
            return x[0] * x[0] + x[2] * x[2] - d[0];
        case SurfaceType::czc:
            return x[0] * x[0] + x[1] * x[1] - d[0];
        case SurfaceType::sc:
            return x[0] * x[0] + x[1] * x[1] + x[2] * x[2] - d[0];
        case SurfaceType::p:
            return d[0] * x[0] + d[1] * x[1] + d[2] * x[2] - d[3];
        default:
            return 0;
    }
}

CELER_FORCEINLINE_FUNCTION int
calc_sense(GeoParamsRef const& p, std::uint32_t s, double const* x)
{
    return real_to_sense(eval_surface(p, s, x));
}

//! Solve a^-1 scaled quadratic with half-b; 'on' drops the c term
CELER_FUNCTION void
solve_quadratic(double a, double half_b, double c, bool on, double* out)
{
    out[0] = out[1] = no_intersection;
    double const a_inv = 1 / a;
    double const hba = half_b * a_inv;
    if (on)
    {
        double const r = -2 * hba;
        if (r > 0)
            out[0] = r;
        return;
    }
    c *= a_inv;
    double const b2_4 = hba * hba;
    if (b2_4 > c)
    {
        double const t2 = std::sqrt(b2_4 - c);
        double const r0 = -hba - t2;
        double const r1 = -hba + t2;
        if (r1 > 0)
        {
            out[1] = r1;
            if (r0 > 0)
                out[0] = r0;
        }
    }
    else if (b2_4 == c)
    {
        double const r0 = -hba;
        if (r0 > 0)
            out[0] = r0;
    }
}

//! Component i of a 3-vector by selects (a dynamic index would force the
//! vector out of registers into per-lane scratch memory)
CELER_FORCEINLINE_FUNCTION double component(double const* v, int i)
{
    return i == 0 ? v[0] : (i == 1 ? v[1] : v[2]);
}

//! Distances to intersection (up to two; no_intersection when absent)
CELER_FUNCTION void calc_intersections(GeoParamsRef const& p,
                                       std::uint32_t s,
                                       double const* x,
                                       double const* u,
                                       bool on,
                                       double* out)
{
    out[0] = out[1] = no_intersection;
    double const* d = p.surf_data + p.surf_offset[s];
    auto const t = static_cast<SurfaceType>(p.surf_type[s]);
    if (t <= SurfaceType::pz)
    {
        int const ax = static_cast<int>(t);
        double const n_dir = component(u, ax);
        if (!on && n_dir != 0)
        {
            double const dist = (d[0] - component(x, ax)) / n_dir;
            if (dist > 0)
                out[0] = dist;
        }
        return;
    }
    if (t == SurfaceType::p)
    {
        double const n_dir = d[0] * u[0] + d[1] * u[1] + d[2] * u[2];
        if (!on && n_dir != 0)
        {
            double const n_pos = d[0] * x[0] + d[1] * x[1] + d[2] * x[2];
            double const dist = (d[3] - n_pos) / n_dir;
            if (dist > 0)
                out[0] = dist;
        }
        return;
    }
    if (t == SurfaceType::sc)
    {
        double const half_b = u[0] * x[0] + u[1] * x[1] + u[2] * x[2];
        double const c = x[0] * x[0] + x[1] * x[1] + x[2] * x[2] - d[0];
        solve_quadratic(1.0, half_b, c, on, out);
        return;
    }
    // Centered cylinders
    int const ax = static_cast<int>(t) - static_cast<int>(SurfaceType::cxc);
    int const iu = (ax + 1) % 3;
    int const iv = (ax + 2) % 3;
    double const ua = component(u, ax);
    double const a = 1 - ua * ua;
    if (a < sqrt_quadratic * sqrt_quadratic)
        return;
    double const uu = component(u, iu);
    double const uv = component(u, iv);
    double const xu = component(x, iu);
    double const xv = component(x, iv);
    double const half_b = uu * xu + uv * xv;
    double const c = xu * xu + xv * xv - d[0];
    solve_quadratic(a, half_b, c, on, out);
}

//! Outward unit normal at a point
CELER_FUNCTION void calc_normal(GeoParamsRef const& p,
                                std::uint32_t s,
                                double const* x,
                                double* n)
{
    double const* d = p.surf_data + p.surf_offset[s];
    auto const t = static_cast<SurfaceType>(p.surf_type[s]);
    n[0] = n[1] = n[2] = 0;
    if (t <= SurfaceType::pz)
    {
        int const ax = static_cast<int>(t);
        n[0] = ax == 0 ? 1 : 0;
        n[1] = ax == 1 ? 1 : 0;
        n[2] = ax == 2 ? 1 : 0;
        return;
    }
    if (t == SurfaceType::p)
    {
        n[0] = d[0];
        n[1] = d[1];
        n[2] = d[2];
        return;
    }
    if (t == SurfaceType::sc)
    {
        n[0] = x[0];
        n[1] = x[1];
        n[2] = x[2];
    }
    else
    {
        int const ax = static_cast<int>(t) - static_cast<int>(SurfaceType::cxc);
        n[0] = ax == 0 ? 0 : x[0];
        n[1] = ax == 1 ? 0 : x[1];
        n[2] = ax == 2 ? 0 : x[2];
    }
    double const norm
        = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    double const inv = 1 / norm;
    n[0] *= inv;
    n[1] *= inv;
    n[2] *= inv;
}

//! Nearest distance to the surface along its normal (orange CalcSafety)
CELER_FUNCTION double
calc_safety_distance(GeoParamsRef const& p, std::uint32_t s, double const* x)
{
    double n[3];
    calc_normal(p, s, x, n);
    if (CELER_UNLIKELY(std::isnan(n[0])))
        return HUGE_VAL;
    int const sense = calc_sense(p, s, x);
    if (sense > 0)
    {
        n[0] = -n[0];
        n[1] = -n[1];
        n[2] = -n[2];
    }
    else if (sense == 0)
    {
        return 0;
    }
    double isect[2];
    calc_intersections(p, s, x, n, false, isect);
    return isect[0] < isect[1] ? isect[0] : isect[1];
}

//---------------------------------------------------------------------------//
// VOLUME LOGIC
//---------------------------------------------------------------------------//
/*!
 * Senses of all faces of a volume as a bit mask (bit set = positive side).
 *
 * \param known_face face index whose sense is supplied (or -1)
 * \param on_face set to the first face the point lies exactly on (or -1)
 */
CELER_FUNCTION std::uint32_t calc_sense_mask(GeoParamsRef const& p,
                                             std::uint32_t vol,
                                             double const* x,
                                             int known_face,
                                             bool known_sense,
                                             int* on_face)
{
    std::uint32_t const begin = p.vol_face_begin[vol];
    int const nf = p.vol_num_faces[vol];
    std::uint32_t mask = 0;
    *on_face = -1;
    for (int f = 0; f < nf; ++f)
    {
        bool s;
        if (f != known_face)
        {
            int const ss = calc_sense(p, p.faces[begin + f], x);
            s = ss >= 0;
            if (*on_face < 0 && ss == 0)
                *on_face = f;
        }
        else
        {
            s = known_sense;
        }
        mask |= static_cast<std::uint32_t>(s) << f;
    }
    return mask;
}

//! Evaluate the postfix logic expression with a 64-deep bit stack
CELER_FUNCTION bool
eval_logic(GeoParamsRef const& p, std::uint32_t vol, std::uint32_t senses)
{
    std::int32_t const* lg = p.logic + p.vol_logic_begin[vol];
    int const n = p.vol_logic_size[vol];
    std::uint64_t stack = 0;
    for (int i = 0; i < n; ++i)
    {
        std::int32_t const t = lg[i];
        if (t >= 0)
        {
            stack = (stack << 1) | ((senses >> t) & 1u);
        }
        else if (t == logic::lnot)
        {
            stack ^= 1u;
        }
        else if (t == logic::land)
        {
            std::uint64_t const b = stack & 1u;
            stack >>= 1;
            stack &= (~std::uint64_t(1)) | b;
        }
        else if (t == logic::lor)
        {
            std::uint64_t const b = stack & 1u;
            stack >>= 1;
            stack |= b;
        }
        else  // ltrue
        {
            stack = (stack << 1) | 1u;
        }
    }
    return stack & 1u;
}

//! Index of a global surface among the faces of a volume (or -1)
CELER_FUNCTION int
find_face(GeoParamsRef const& p, std::uint32_t vol, std::uint32_t surf)
{
    std::uint32_t const begin = p.vol_face_begin[vol];
    int const nf = p.vol_num_faces[vol];
    for (int f = 0; f < nf; ++f)
    {
        if (p.faces[begin + f] == surf)
            return f;
    }
    return -1;
}

CELER_FORCEINLINE_FUNCTION bool
in_bbox(GeoParamsRef const& p, std::uint32_t vol, double const* x)
{
    double const* b = p.vol_bbox + 6 * vol;
    return x[0] >= b[0] && x[1] >= b[1] && x[2] >= b[2] && x[0] <= b[3]
           && x[1] <= b[4] && x[2] <= b[5];
}

/*!
 * First volume of bih_vols[begin, begin + count) whose box and logic
 * contain the point (the "is_inside" test of orange/detail/
 * BIHTraverser.hh:254-266 with SimpleUnitTracker's sense evaluation).
 */
CELER_FUNCTION bool find_in_list(GeoParamsRef const& p,
                                 std::uint32_t begin,
                                 std::uint32_t count,
                                 double const* x,
                                 std::uint32_t* found,
                                 bool* on_surface)
{
    for (std::uint32_t k = 0; k < count; ++k)
    {
        std::uint32_t const v = p.bih_vols[begin + k];
        if (!in_bbox(p, v, x))
            continue;
        int on_face;
        std::uint32_t const mask = calc_sense_mask(p, v, x, -1, false, &on_face);
        if (eval_logic(p, v, mask))
        {
            *found = v;
            *on_surface = on_face >= 0;
            return true;
        }
    }
    return false;
}

//---------------------------------------------------------------------------//
// TRACKING
//---------------------------------------------------------------------------//
/*!
 * Initialize the volume from the position; "failed" if on a surface.
 *
 * The candidate volumes come from the point-location tree (a node's plane
 * test prunes its subtree; a leaf's volumes are tested in order), then the
 * volumes with an unbounded box. Volumes of a simple unit do not overlap,
 * so the first match is the volume (or, for a point on a face shared by
 * two volumes, either one, and the point is rejected as on a surface).
 */
CELER_FUNCTION void initialize(GeoParamsRef const& p,
                               GeoTrack& g,
                               double const* pos,
                               double const* dir)
{
    for (int k = 0; k < 3; ++k)
    {
        g.pos[k] = pos[k];
        g.dir[k] = dir[k];
    }
    g.failed = 0;
    std::uint32_t found = invalid_id;
    bool on_surface = false;
    std::uint32_t i = 0;
    while (i < p.num_bih_nodes)
    {
        BihNode const node = p.bih_nodes[i];
        bool pass = true;
        if (node.axis < 3)
        {
            double const xa = component(pos, node.axis);
            pass = node.upper ? xa >= node.plane : xa <= node.plane;
        }
        if (pass && node.num_vols == 0)
        {
            ++i;
            continue;
        }
        if (pass
            && find_in_list(p, node.vol_begin, node.num_vols, pos, &found, &on_surface))
        {
            break;
        }
        i = node.skip;
    }
    if (found == invalid_id)
        find_in_list(p, p.bih_inf_begin, p.bih_inf_count, pos, &found, &on_surface);
    if (on_surface)
    {
        found = invalid_id;
    }
    else if (found == invalid_id)
    {
        found = p.background;
    }
    if (found == invalid_id)
    {
        g.failed = 1;
        found = p.exterior;
    }
    g.volume = found;
    g.surface = invalid_id;
    g.sense = 0;
    g.boundary = static_cast<std::uint8_t>(BoundaryResult::exiting);
    g.clear_next();
}

/*!
 * State of a point known to lie inside a volume, away from its faces: what
 * initialize() gives for it, without the search.
 */
CELER_FORCEINLINE_FUNCTION void initialize_interior(GeoTrack& g,
                                                    std::uint32_t volume,
                                                    double const* pos,
                                                    double const* dir)
{
    for (int k = 0; k < 3; ++k)
    {
        g.pos[k] = pos[k];
        g.dir[k] = dir[k];
    }
    g.failed = 0;
    g.volume = volume;
    g.surface = invalid_id;
    g.sense = 0;
    g.boundary = static_cast<std::uint8_t>(BoundaryResult::exiting);
    g.clear_next();
}

//! Distance and whether a boundary limits it
struct Propagation
{
    double distance;
    bool boundary;
};

/*!
 * Find the distance to the next boundary, up to max_step (or HUGE_VAL).
 */
CELER_FUNCTION Propagation find_next_step(GeoParamsRef const& p,
                                          GeoTrack& g,
                                          double max_step)
{
    if (CELER_UNLIKELY(g.boundary
                       == static_cast<std::uint8_t>(BoundaryResult::reentrant)))
    {
        return {0, true};
    }
    std::uint32_t const vol = g.volume;
    std::uint32_t const begin = p.vol_face_begin[vol];
    int const nf = p.vol_num_faces[vol];
    int const on_face = g.is_on_boundary() ? find_face(p, vol, g.surface) : -1;
    bool const simple = !(p.vol_flags[vol] & (internal_surfaces | implicit_vol));

    // Valid candidate intersections of face f, root k (in [0, max_step])
    auto for_each_isect = [&](auto&& visit) {
        for (int f = 0; f < nf; ++f)
        {
            std::uint32_t const s = p.faces[begin + f];
            bool const on = (f == on_face);
            auto const t = static_cast<SurfaceType>(p.surf_type[s]);
            if (on && (t <= SurfaceType::pz || t == SurfaceType::p))
                continue;  // single-intersection surface: cannot reintersect
            double d2[2];
            calc_intersections(p, s, g.pos, g.dir, on, d2);
            for (int k = 0; k < 2; ++k)
            {
                double const d = d2[k];
                if (!(d <= max_step) || d == no_intersection)
                    continue;
                visit(d, f, k);
            }
        }
    };

    std::uint32_t isect_surf = invalid_id;
    bool isect_sense = false;
    double isect_dist = no_intersection;
    if (simple)
    {
        // Simple volume: the nearest intersection exits
        double best = no_intersection;
        int best_face = -1;
        for_each_isect([&](double d, int f, int) {
            if (d < best)
            {
                best = d;
                best_face = f;
            }
        });
        if (best_face >= 0)
        {
            std::uint32_t const s = p.faces[begin + best_face];
            bool cur_sense;
            if (s == g.surface)
                cur_sense = g.sense;
            else
                cur_sense = calc_sense(p, s, g.pos) >= 0;
            isect_surf = s;
            isect_sense = cur_sense;
            isect_dist = best;
        }
    }
    else
    {
        // Complex volume: visit intersections in increasing distance (ties
        // in face, then root, order), flipping the crossed face's sense,
        // until the logic says the track has left. The next candidate is
        // found by a fresh scan each time, so no per-lane intersection list
        // is kept (the usual exit is the first or second candidate).
        int dummy;
        std::uint32_t senses = calc_sense_mask(
            p, vol, g.pos, on_face, static_cast<bool>(g.sense), &dummy);
        double last_d = -1;
        int last_key = -1;
        for (int iter = 0; iter < max_intersections; ++iter)
        {
            double next_d = no_intersection;
            int next_key = 0x7fffffff;
            for_each_isect([&](double d, int f, int k) {
                int const key = 2 * f + k;
                bool const after = d > last_d || (d == last_d && key > last_key);
                bool const before = d < next_d || (d == next_d && key < next_key);
                if (after && before)
                {
                    next_d = d;
                    next_key = key;
                }
            });
            if (next_key == 0x7fffffff)
                break;
            int const f = next_key / 2;
            senses ^= (1u << f);
            if (!eval_logic(p, vol, senses))
            {
                bool const new_sense = (senses >> f) & 1u;
                isect_surf = p.faces[begin + f];
                isect_sense = !new_sense;
                isect_dist = next_d;
                break;
            }
            last_d = next_d;
            last_key = next_key;
        }
    }

    Propagation result;
    if (isect_surf == invalid_id)
    {
        // No boundary within reach
        g.next_step = (max_step < no_intersection) ? max_step
                                                    : no_intersection;
        g.next_surface = invalid_id;
        result.distance = g.next_step;
        result.boundary = false;
    }
    else
    {
        g.next_step = isect_dist;
        g.next_surface = isect_surf;
        g.next_sense = isect_sense;
        result.distance = isect_dist;
        result.boundary = true;
    }
    return result;
}

//! Move to the next boundary (found by find_next_step)
CELER_FORCEINLINE_FUNCTION void move_to_boundary(GeoTrack& g)
{
    double const dist = g.next_step;
    for (int k = 0; k < 3; ++k)
        g.pos[k] = std::fma(dist, g.dir[k], g.pos[k]);
    g.surface = g.next_surface;
    g.sense = g.next_sense;
    g.clear_next();
}

//! Move within the volume by a distance less than the boundary distance
CELER_FORCEINLINE_FUNCTION void move_internal(GeoTrack& g, double dist)
{
    for (int k = 0; k < 3; ++k)
        g.pos[k] = std::fma(dist, g.dir[k], g.pos[k]);
    g.next_step -= dist;
    g.surface = invalid_id;
}

//! Move to a nearby point inside the current volume
CELER_FORCEINLINE_FUNCTION void move_internal(GeoTrack& g, double const* x)
{
    for (int k = 0; k < 3; ++k)
        g.pos[k] = x[k];
    g.surface = invalid_id;
    g.clear_next();
}

/*!
 * Cross the current surface into the neighboring volume.
 */
CELER_FUNCTION void cross_boundary(GeoParamsRef const& p, GeoTrack& g)
{
    if (CELER_UNLIKELY(g.boundary
                       == static_cast<std::uint8_t>(BoundaryResult::reentrant)))
    {
        g.boundary = static_cast<std::uint8_t>(BoundaryResult::exiting);
        return;
    }
    g.sense = !g.sense;
    g.boundary = static_cast<std::uint8_t>(BoundaryResult::exiting);

    std::uint32_t const surf = g.surface;
    std::uint32_t const cb = p.conn_begin[surf];
    std::uint32_t const cn = p.conn_size[surf];
    std::uint32_t found = invalid_id;
    for (std::uint32_t i = 0; i < cn; ++i)
    {
        std::uint32_t const v = p.conn_vols[cb + i];
        if (v == g.volume)
            continue;
        if (cn >= 3 && !in_bbox(p, v, g.pos))
            continue;
        int const face = find_face(p, v, surf);
        int on_face;
        std::uint32_t const mask = calc_sense_mask(
            p, v, g.pos, face, static_cast<bool>(g.sense), &on_face);
        if (eval_logic(p, v, mask))
        {
            found = v;
            break;
        }
    }
    if (found == invalid_id)
        found = p.background;
    if (CELER_UNLIKELY(found == invalid_id))
    {
        g.failed = 1;
        found = p.exterior;
    }
    g.volume = found;
}

/*!
 * Change direction; on a boundary this may make the crossing reentrant.
 */
CELER_FUNCTION void
set_dir(GeoParamsRef const& p, GeoTrack& g, double const* newdir)
{
    if (g.is_on_boundary())
    {
        double n[3];
        calc_normal(p, g.surface, g.pos, n);
        double const dn_new = n[0] * newdir[0] + n[1] * newdir[1]
                              + n[2] * newdir[2];
        double const dn_old = n[0] * g.dir[0] + n[1] * g.dir[1]
                              + n[2] * g.dir[2];
        if ((dn_new >= 0) != (dn_old >= 0))
            g.boundary ^= 1u;
    }
    for (int k = 0; k < 3; ++k)
        g.dir[k] = newdir[k];
    g.clear_next();
}

/*!
 * Isotropic safety distance (conservative; 0 for unsupported volumes).
 */
CELER_FUNCTION double find_safety(GeoParamsRef const& p, GeoTrack const& g)
{
    std::uint32_t const vol = g.volume;
    if (!(p.vol_flags[vol] & simple_safety))
        return 0;
    std::uint32_t const begin = p.vol_face_begin[vol];
    int const nf = p.vol_num_faces[vol];
    double result = HUGE_VAL;
    for (int f = 0; f < nf; ++f)
    {
        double const d = calc_safety_distance(p, p.faces[begin + f], g.pos);
        result = d < result ? d : result;
    }
    return result;
}

//! Whether the track has left the world
CELER_FORCEINLINE_FUNCTION bool
is_outside(GeoParamsRef const& p, GeoTrack const& g)
{
    return g.volume == p.exterior;
}

//---------------------------------------------------------------------------//
}  // namespace geo
}  // namespace celeritas_amd
