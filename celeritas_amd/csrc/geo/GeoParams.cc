//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file geo/GeoParams.cc
//---------------------------------------------------------------------------//
#include "GeoParams.hh"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>
#include <type_traits>

namespace celeritas_amd
{
namespace
{
//---------------------------------------------------------------------------//
SurfaceType surface_type_from_string(std::string const& s)
{
    static char const* const names[] = {"px", "py", "pz", "cxc", "cyc", "czc", "sc", "p"};
    for (int i = 0; i < static_cast<int>(SurfaceType::size_); ++i)
    {
        if (s == names[i])
            return static_cast<SurfaceType>(i);
    }
    throw RuntimeError("unsupported ORANGE surface type '" + s + "'");
}

//! Axis-aligned box used for interval propagation through the logic
struct Box
{
    double lo[3];
    double hi[3];

    static Box infinite()
    {
        return {{-HUGE_VAL, -HUGE_VAL, -HUGE_VAL}, {HUGE_VAL, HUGE_VAL, HUGE_VAL}};
    }
    static Box empty()
    {
        return {{HUGE_VAL, HUGE_VAL, HUGE_VAL}, {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL}};
    }
};
Box intersect(Box const& a, Box const& b)
{
    Box r;
    for (int k = 0; k < 3; ++k)
    {
        r.lo[k] = std::max(a.lo[k], b.lo[k]);
        r.hi[k] = std::min(a.hi[k], b.hi[k]);
    }
    return r;
}
Box unite(Box const& a, Box const& b)
{
    Box r;
    for (int k = 0; k < 3; ++k)
    {
        r.lo[k] = std::min(a.lo[k], b.lo[k]);
        r.hi[k] = std::max(a.hi[k], b.hi[k]);
    }
    return r;
}

//! Bounding boxes of the (negative, positive) sides of a surface
void halfspace_boxes(SurfaceType t, std::vector<double> const& d, Box* neg, Box* pos)
{
    *neg = Box::infinite();
    *pos = Box::infinite();
    if (t <= SurfaceType::pz)
    {
        int ax = static_cast<int>(t);
        neg->hi[ax] = d[0];
        pos->lo[ax] = d[0];
    }
    else if (t >= SurfaceType::cxc && t <= SurfaceType::czc)
    {
        int ax = static_cast<int>(t) - static_cast<int>(SurfaceType::cxc);
        double r = std::sqrt(d[0]);
        for (int k : {(ax + 1) % 3, (ax + 2) % 3})
        {
            neg->lo[k] = -r;
            neg->hi[k] = r;
        }
    }
    else if (t == SurfaceType::sc)
    {
        double r = std::sqrt(d[0]);
        for (int k = 0; k < 3; ++k)
        {
            neg->lo[k] = -r;
            neg->hi[k] = r;
        }
    }
}

bool is_finite(Box const& b)
{
    for (int k = 0; k < 3; ++k)
    {
        if (!std::isfinite(b.lo[k]) || !std::isfinite(b.hi[k]))
            return false;
    }
    return true;
}

bool is_empty(Box const& b)
{
    for (int k = 0; k < 3; ++k)
    {
        if (!(b.lo[k] <= b.hi[k]))
            return true;
    }
    return false;
}

//! Half the surface area of a box
double half_area(Box const& b)
{
    double d[3];
    for (int k = 0; k < 3; ++k)
        d[k] = std::max(b.hi[k] - b.lo[k], 0.0);
    return d[0] * d[1] + d[1] * d[2] + d[2] * d[0];
}

/*!
 * Depth-first builder of the point-location tree.
 *
 * Partitions as the reference does (orange/detail/BIHPartitioner.cc:
 * 43-80): candidate planes lie halfway between consecutive distinct box
 * centres along each axis (at most max_candidates per axis, equally
 * spaced), a volume goes left when its centre is below the plane, and the
 * cost of a split is the sum over both sides of the volume count times the
 * surface area of their union box. A node becomes a leaf at max_leaf
 * volumes or when no plane separates the centres.
 */
struct BihBuilder
{
    static constexpr std::size_t max_leaf = 2;
    static constexpr std::size_t max_candidates = 64;

    std::vector<Box> const& boxes;
    std::vector<BihNode> nodes;
    std::vector<std::uint32_t> vols;

    double center(std::uint32_t v, int ax) const
    {
        return 0.5 * (boxes[v].lo[ax] + boxes[v].hi[ax]);
    }

    void make(std::vector<std::uint32_t> ids, int axis, bool upper, double plane)
    {
        auto const self = static_cast<std::uint32_t>(nodes.size());
        BihNode node{};
        node.plane = plane;
        node.axis = static_cast<std::uint8_t>(axis);
        node.upper = upper;
        nodes.push_back(node);

        double best_cost = HUGE_VAL;
        int best_axis = -1;
        double best_pos = 0;
        if (ids.size() > max_leaf)
        {
            for (int ax = 0; ax < 3; ++ax)
            {
                std::vector<double> c;
                for (auto v : ids)
                    c.push_back(this->center(v, ax));
                std::sort(c.begin(), c.end());
                c.erase(std::unique(c.begin(), c.end()), c.end());
                std::size_t const step
                    = std::max<std::size_t>(c.size() / (max_candidates + 1), 1);
                for (std::size_t i = step; i < c.size(); i += step)
                {
                    double const pos = 0.5 * (c[i - 1] + c[i]);
                    Box l = Box::empty(), r = Box::empty();
                    double nl = 0, nr = 0;
                    for (auto v : ids)
                    {
                        if (this->center(v, ax) < pos)
                        {
                            l = unite(l, boxes[v]);
                            nl += 1;
                        }
                        else
                        {
                            r = unite(r, boxes[v]);
                            nr += 1;
                        }
                    }
                    double const cost = nl * half_area(l) + nr * half_area(r);
                    if (cost < best_cost)
                    {
                        best_cost = cost;
                        best_axis = ax;
                        best_pos = pos;
                    }
                }
            }
        }
        if (best_axis < 0)
        {
            CELER_VALIDATE(ids.size() < 65536,
                           << "too many volumes share one bounding-box centre");
            std::sort(ids.begin(), ids.end());
            nodes[self].vol_begin = static_cast<std::uint32_t>(vols.size());
            nodes[self].num_vols = static_cast<std::uint16_t>(ids.size());
            vols.insert(vols.end(), ids.begin(), ids.end());
        }
        else
        {
            std::vector<std::uint32_t> left, right;
            Box lb = Box::empty(), rb = Box::empty();
            for (auto v : ids)
            {
                if (this->center(v, best_axis) < best_pos)
                {
                    left.push_back(v);
                    lb = unite(lb, boxes[v]);
                }
                else
                {
                    right.push_back(v);
                    rb = unite(rb, boxes[v]);
                }
            }
            this->make(std::move(left), best_axis, false, lb.hi[best_axis]);
            this->make(std::move(right), best_axis, true, rb.lo[best_axis]);
        }
        nodes[self].skip = static_cast<std::uint32_t>(nodes.size());
    }
};
//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
std::vector<std::int32_t> parse_logic_string(std::string const& s)
{
    std::vector<std::int32_t> result;
    std::istringstream is(s);
    std::string tok;
    while (is >> tok)
    {
        if (tok == "~")
            result.push_back(logic::lnot);
        else if (tok == "&")
            result.push_back(logic::land);
        else if (tok == "|")
            result.push_back(logic::lor);
        else if (tok == "*")
            result.push_back(logic::ltrue);
        else
            result.push_back(static_cast<std::int32_t>(std::stol(tok)));
    }
    return result;
}

//---------------------------------------------------------------------------//
GeoInput geo_input_from_orange_json(JsonValue const& doc)
{
    CELER_VALIDATE(!doc.has("universes") || doc.at("universes").size() == 1,
                   << "multi-universe geometry is not supported (the file has "
                   << doc.at("universes").size() << " universes)");
    JsonValue const& univ = doc.has("universes") ? doc.at("universes")[0] : doc;
    CELER_VALIDATE(univ.get_str("_type", "simple unit") == "simple unit",
                   << "only a single simple unit is supported");
    GeoInput inp;
    JsonValue const& surfs = univ.at("surfaces");
    auto const& types = surfs.at("types");
    auto const data = surfs.at("data").as_doubles();
    auto const sizes = surfs.at("sizes").as_ints();
    std::size_t off = 0;
    for (std::size_t i = 0; i < types.size(); ++i)
    {
        SurfaceType t = surface_type_from_string(types[i].as_string());
        std::size_t n = static_cast<std::size_t>(sizes[i]);
        CELER_VALIDATE(static_cast<int>(n) == surface_data_size(t),
                       << "bad data size for surface " << i);
        inp.surf_types.push_back(static_cast<std::uint8_t>(t));
        inp.surf_data.emplace_back(data.begin() + off, data.begin() + off + n);
        off += n;
    }
    if (univ.has("surface_names"))
    {
        for (auto const& s : univ.at("surface_names").arr)
            inp.surface_labels.push_back(s.as_string());
    }
    auto const& cells = univ.at("cells");
    for (std::size_t i = 0; i < cells.size(); ++i)
    {
        JsonValue const& c = cells[i];
        GeoInput::Volume v;
        for (auto f : c.at("faces").as_ints())
            v.faces.push_back(static_cast<std::uint32_t>(f));
        v.logic = parse_logic_string(c.at("logic").as_string());
        v.flags = static_cast<std::uint8_t>(c.get_int("flags", 0)
                                            & (internal_surfaces | implicit_vol));
        v.background = (c.get_int("zorder", 2) == 1);
        if (univ.has("cell_names"))
            v.label = univ.at("cell_names")[i].as_string();
        inp.volumes.push_back(std::move(v));
    }
    return inp;
}

//---------------------------------------------------------------------------//
GeoParams::GeoParams(GeoInput const& inp)
{
    num_surfaces_ = static_cast<size_type>(inp.surf_types.size());
    num_volumes_ = static_cast<size_type>(inp.volumes.size());
    CELER_VALIDATE(num_volumes_ > 0, << "geometry has no volumes");

    // Surfaces
    for (size_type s = 0; s < num_surfaces_; ++s)
    {
        h_surf_type.push_back(inp.surf_types[s]);
        h_surf_offset.push_back(static_cast<std::uint32_t>(h_surf_data.size()));
        for (double d : inp.surf_data[s])
            h_surf_data.push_back(d);
    }

    // Volumes
    std::vector<std::vector<std::uint32_t>> conn(num_surfaces_);
    for (size_type v = 0; v < num_volumes_; ++v)
    {
        auto const& vol = inp.volumes[v];
        CELER_VALIDATE(vol.faces.size() <= static_cast<std::size_t>(max_faces),
                       << "volume " << v << " has " << vol.faces.size()
                       << " faces (max " << max_faces << ")");
        h_face_begin.push_back(static_cast<std::uint32_t>(h_faces.size()));
        h_num_faces.push_back(static_cast<std::uint8_t>(vol.faces.size()));
        int max_isect = 0;
        bool simple = true;
        for (auto s : vol.faces)
        {
            CELER_VALIDATE(s < num_surfaces_, << "bad face in volume " << v);
            h_faces.push_back(s);
            conn[s].push_back(v);
            auto t = static_cast<SurfaceType>(inp.surf_types[s]);
            max_isect += (t <= SurfaceType::pz || t == SurfaceType::p) ? 1 : 2;
            simple = simple && (t != SurfaceType::p || true);
        }
        h_max_isect.push_back(static_cast<std::uint8_t>(max_isect));
        h_logic_begin.push_back(static_cast<std::uint32_t>(h_logic.size()));
        h_logic_size.push_back(static_cast<std::uint16_t>(vol.logic.size()));
        for (auto t : vol.logic)
        {
            CELER_VALIDATE(t < 0 || t < static_cast<std::int32_t>(vol.faces.size()),
                           << "bad logic token in volume " << v);
            h_logic.push_back(t);
        }
        std::uint8_t flags = vol.flags;
        if (simple && !(flags & implicit_vol))
            flags |= simple_safety;
        h_flags.push_back(flags);
        if (vol.background)
            background = v;
        labels_.push_back(vol.label);

        // Bounding box by interval propagation of the postfix logic
        std::vector<std::pair<Box, Box>> stack;  // (true region, false region)
        for (auto t : vol.logic)
        {
            if (t >= 0)
            {
                std::uint32_t s = vol.faces[t];
                Box neg, pos;
                halfspace_boxes(static_cast<SurfaceType>(inp.surf_types[s]),
                                inp.surf_data[s], &neg, &pos);
                stack.push_back({pos, neg});
            }
            else if (t == logic::lnot)
            {
                std::swap(stack.back().first, stack.back().second);
            }
            else if (t == logic::ltrue)
            {
                stack.push_back({Box::infinite(), Box::empty()});
            }
            else
            {
                CELER_VALIDATE(stack.size() >= 2, << "bad logic in volume " << v);
                auto b = stack.back();
                stack.pop_back();
                auto a = stack.back();
                stack.pop_back();
                if (t == logic::land)
                    stack.push_back({intersect(a.first, b.first),
                                     unite(a.second, b.second)});
                else
                    stack.push_back({unite(a.first, b.first),
                                     intersect(a.second, b.second)});
            }
        }
        CELER_VALIDATE(stack.size() == 1, << "unbalanced logic in volume " << v);
        // The boxes only prune candidates, so they are bumped outward (as
        // ORANGE bumps its BIH boxes): a position moved onto a face can
        // land a few ulps outside the exact box of the volume it enters
        Box const& bb = stack.back().first;
        auto bump = [](double x) {
            return std::isfinite(x) ? 1e-8 * std::max(1.0, std::fabs(x)) : 0.0;
        };
        for (int k = 0; k < 3; ++k)
            h_bbox.push_back(bb.lo[k] - bump(bb.lo[k]));
        for (int k = 0; k < 3; ++k)
            h_bbox.push_back(bb.hi[k] + bump(bb.hi[k]));
    }
    // Point-location tree over the non-implicit volumes with a finite,
    // non-empty box; unbounded ones are tested after it, empty ones never
    // contain a point
    {
        std::vector<Box> boxes(num_volumes_);
        std::vector<std::uint32_t> finite, unbounded;
        for (size_type v = 0; v < num_volumes_; ++v)
        {
            for (int k = 0; k < 3; ++k)
            {
                boxes[v].lo[k] = h_bbox[6 * v + k];
                boxes[v].hi[k] = h_bbox[6 * v + 3 + k];
            }
            if ((h_flags[v] & implicit_vol) || is_empty(boxes[v]))
                continue;
            (is_finite(boxes[v]) ? finite : unbounded).push_back(v);
        }
        BihBuilder b{boxes, {}, {}};
        if (!finite.empty())
            b.make(std::move(finite), 3, false, 0);
        h_bih_nodes = std::move(b.nodes);
        h_bih_vols = std::move(b.vols);
        bih_inf_begin_ = static_cast<std::uint32_t>(h_bih_vols.size());
        bih_inf_count_ = static_cast<std::uint32_t>(unbounded.size());
        h_bih_vols.insert(h_bih_vols.end(), unbounded.begin(), unbounded.end());
    }

    for (size_type s = 0; s < num_surfaces_; ++s)
    {
        h_conn_begin.push_back(static_cast<std::uint32_t>(h_conn_vols.size()));
        h_conn_size.push_back(static_cast<std::uint32_t>(conn[s].size()));
        for (auto v : conn[s])
            h_conn_vols.push_back(v);
    }

    // Upload. The tables that tracking reads per track (surfaces, faces,
    // volume logic and boxes, surface connectivity) share one contiguous
    // region that a kernel can copy into LDS and rebase into; the point-
    // location tree stays apart.
    std::vector<char> hot;
    auto place = [&hot](auto const& v) {
        std::size_t const off = (hot.size() + 15) / 16 * 16;
        std::size_t const n = v.size() * sizeof(v[0]);
        hot.resize(off + n);
        if (n)
            std::memcpy(hot.data() + off, v.data(), n);
        return off;
    };
    std::size_t const o_surf_type = place(h_surf_type);
    std::size_t const o_surf_offset = place(h_surf_offset);
    std::size_t const o_surf_data = place(h_surf_data);
    std::size_t const o_face_begin = place(h_face_begin);
    std::size_t const o_num_faces = place(h_num_faces);
    std::size_t const o_max_isect = place(h_max_isect);
    std::size_t const o_faces = place(h_faces);
    std::size_t const o_logic_begin = place(h_logic_begin);
    std::size_t const o_logic_size = place(h_logic_size);
    std::size_t const o_logic = place(h_logic);
    std::size_t const o_flags = place(h_flags);
    std::size_t const o_bbox = place(h_bbox);
    std::size_t const o_conn_begin = place(h_conn_begin);
    std::size_t const o_conn_size = place(h_conn_size);
    std::size_t const o_conn_vols = place(h_conn_vols);
    hot.resize((hot.size() + 15) / 16 * 16);
    d_hot_ = to_device(hot);
    char const* base = d_hot_.data();
    auto at = [base](std::size_t off, auto const& v) {
        using T = std::remove_const_t<std::remove_reference_t<decltype(v[0])>>;
        return v.empty() ? static_cast<T const*>(nullptr)
                         : reinterpret_cast<T const*>(base + off);
    };
    d_bih_nodes_ = to_device(h_bih_nodes);
    d_bih_vols_ = to_device(h_bih_vols);

    ref_.num_surfaces = num_surfaces_;
    ref_.num_volumes = num_volumes_;
    ref_.hot_base = base;
    ref_.hot_bytes = static_cast<std::uint32_t>(hot.size());
    ref_.surf_type = at(o_surf_type, h_surf_type);
    ref_.surf_offset = at(o_surf_offset, h_surf_offset);
    ref_.surf_data = at(o_surf_data, h_surf_data);
    ref_.vol_face_begin = at(o_face_begin, h_face_begin);
    ref_.vol_num_faces = at(o_num_faces, h_num_faces);
    ref_.vol_max_isect = at(o_max_isect, h_max_isect);
    ref_.faces = at(o_faces, h_faces);
    ref_.vol_logic_begin = at(o_logic_begin, h_logic_begin);
    ref_.vol_logic_size = at(o_logic_size, h_logic_size);
    ref_.logic = at(o_logic, h_logic);
    ref_.vol_flags = at(o_flags, h_flags);
    ref_.vol_bbox = at(o_bbox, h_bbox);
    ref_.conn_begin = at(o_conn_begin, h_conn_begin);
    ref_.conn_size = at(o_conn_size, h_conn_size);
    ref_.conn_vols = at(o_conn_vols, h_conn_vols);
    ref_.bih_nodes = d_bih_nodes_.data();
    ref_.bih_vols = d_bih_vols_.data();
    ref_.num_bih_nodes = static_cast<std::uint32_t>(h_bih_nodes.size());
    ref_.bih_inf_begin = bih_inf_begin_;
    ref_.bih_inf_count = bih_inf_count_;
    ref_.background = background;
    ref_.exterior = 0;
}

//---------------------------------------------------------------------------//
std::uint32_t GeoParams::find_volume(std::string const& label) const
{
    for (size_type v = 0; v < num_volumes_; ++v)
    {
        if (labels_[v] == label)
            return v;
    }
    return invalid_id;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
