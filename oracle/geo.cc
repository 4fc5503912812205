//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see geo.hh).
//---------------------------------------------------------------------------//
#include "geo.hh"

#include <algorithm>
#include <cmath>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace orc
{
namespace
{
//---------------------------------------------------------------------------//
constexpr double inf = std::numeric_limits<double>::infinity();

enum class SignedSense
{
    inside = -1,
    on = 0,
    outside = 1
};

SignedSense real_to_sense(double q)
{
    if (q < 0)
        return SignedSense::inside;
    if (q > 0)
        return SignedSense::outside;
    return SignedSense::on;
}

bool to_sense(SignedSense s)
{
    return static_cast<int>(s) >= 0;
}

double dot(double const* a, double const* b)
{
    double r = 0;
    for (int i = 0; i < 3; ++i)
        r = std::fma(a[i], b[i], r);
    return r;
}

int axis_of(std::string const& t)
{
    char c = t[t.size() == 2 ? 1 : 1];
    return c - 'x';
}

SignedSense calc_sense(Surface const& s, double const* x)
{
    auto const& t = s.type;
    if (t == "px" || t == "py" || t == "pz")
        return real_to_sense(x[axis_of(t)] - s.d[0]);
    if (t == "cxc" || t == "cyc" || t == "czc")
    {
        int ax = axis_of(t);
        int u = (ax + 1) % 3, v = (ax + 2) % 3;
        return real_to_sense(x[u] * x[u] + x[v] * x[v] - s.d[0]);
    }
    if (t == "sc")
        return real_to_sense(x[0] * x[0] + x[1] * x[1] + x[2] * x[2] - s.d[0]);
    if (t == "p")
        return real_to_sense(s.d[0] * x[0] + s.d[1] * x[1] + s.d[2] * x[2] - s.d[3]);
    throw std::runtime_error("oracle: bad surface " + t);
}

//! Quadratic solver following surf/detail/QuadraticSolver.hh
void quadratic(double a, double half_b, double c, bool on, double out[2])
{
    out[0] = out[1] = inf;
    double a_inv = 1 / a;
    double hba = half_b * a_inv;
    if (on)
    {
        double r = -2 * hba;
        out[0] = r > 0 ? r : inf;
        return;
    }
    c *= a_inv;
    double b2_4 = hba * hba;
    if (b2_4 > c)
    {
        double t2 = std::sqrt(b2_4 - c);
        out[0] = -hba - t2;
        out[1] = -hba + t2;
        if (out[1] <= 0)
        {
            out[0] = out[1] = inf;
        }
        else if (out[0] <= 0)
        {
            out[0] = inf;
        }
    }
    else if (b2_4 == c)
    {
        out[0] = -hba;
        if (out[0] <= 0)
            out[0] = inf;
    }
}

//! Intersections: returns count of entries written (1 for planes, 2 else)
int calc_intersections(Surface const& s, double const* x, double const* u,
                       bool on, double out[2])
{
    auto const& t = s.type;
    out[0] = out[1] = inf;
    if (t == "px" || t == "py" || t == "pz" || t == "p")
    {
        double n_dir, n_pos, d;
        if (t == "p")
        {
            n_dir = s.d[0] * u[0] + s.d[1] * u[1] + s.d[2] * u[2];
            n_pos = s.d[0] * x[0] + s.d[1] * x[1] + s.d[2] * x[2];
            d = s.d[3];
        }
        else
        {
            int ax = axis_of(t);
            n_dir = u[ax];
            n_pos = x[ax];
            d = s.d[0];
        }
        if (!on && n_dir != 0)
        {
            double dist = (d - n_pos) / n_dir;
            if (dist > 0)
                out[0] = dist;
        }
        return 1;
    }
    if (t == "sc")
    {
        double hb = u[0] * x[0] + u[1] * x[1] + u[2] * x[2];
        quadratic(1.0, hb, x[0] * x[0] + x[1] * x[1] + x[2] * x[2] - s.d[0], on, out);
        return 2;
    }
    int ax = axis_of(t);
    int iu = (ax + 1) % 3, iv = (ax + 2) % 3;
    double a = 1 - u[ax] * u[ax];
    if (a < 1e-5 * 1e-5)
        return 2;
    double hb = u[iu] * x[iu] + u[iv] * x[iv];
    double c = x[iu] * x[iu] + x[iv] * x[iv] - s.d[0];
    quadratic(a, hb, c, on, out);
    return 2;
}

void calc_normal(Surface const& s, double const* x, double n[3])
{
    auto const& t = s.type;
    n[0] = n[1] = n[2] = 0;
    if (t == "px" || t == "py" || t == "pz")
    {
        n[axis_of(t)] = 1;
        return;
    }
    if (t == "p")
    {
        for (int i = 0; i < 3; ++i)
            n[i] = s.d[i];
        return;
    }
    if (t == "sc")
    {
        for (int i = 0; i < 3; ++i)
            n[i] = x[i];
    }
    else
    {
        int ax = axis_of(t);
        n[(ax + 1) % 3] = x[(ax + 1) % 3];
        n[(ax + 2) % 3] = x[(ax + 2) % 3];
    }
    double len = std::sqrt(dot(n, n));
    for (int i = 0; i < 3; ++i)
        n[i] /= len;
}

//! Sense vector for a volume; optional known face; first "on" face out
std::vector<bool> senses_for(Geometry const& g, Volume const& v, double const* x,
                             int known_face, bool known_sense, int* on_face)
{
    std::vector<bool> r(v.faces.size());
    *on_face = -1;
    for (std::size_t f = 0; f < v.faces.size(); ++f)
    {
        if (static_cast<int>(f) == known_face)
        {
            r[f] = known_sense;
            continue;
        }
        SignedSense ss = calc_sense(g.surfaces[v.faces[f]], x);
        r[f] = to_sense(ss);
        if (*on_face < 0 && ss == SignedSense::on)
            *on_face = static_cast<int>(f);
    }
    return r;
}

bool evaluate(std::vector<int> const& logic, std::vector<bool> const& senses)
{
    std::vector<bool> st;
    for (int t : logic)
    {
        if (t >= 0)
        {
            st.push_back(senses.at(t));
        }
        else if (t == -3)
        {
            st.push_back(true);
        }
        else if (t == -6)
        {
            st.back() = !st.back();
        }
        else
        {
            bool b = st.back();
            st.pop_back();
            bool a = st.back();
            st.back() = (t == -5) ? (a && b) : (a || b);
        }
    }
    return st.back();
}

int find_face(Volume const& v, std::uint32_t surf)
{
    for (std::size_t f = 0; f < v.faces.size(); ++f)
        if (v.faces[f] == surf)
            return static_cast<int>(f);
    return -1;
}
//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
Geometry Geometry::from_orange_json(Json const& doc)
{
    Json const& u = doc.find("universes") ? doc("universes").a.at(0) : doc;
    Geometry g;
    auto const& types = u("surfaces")("types").a;
    auto data = u("surfaces")("data").nums();
    auto sizes = u("surfaces")("sizes").nums();
    std::size_t off = 0;
    for (std::size_t i = 0; i < types.size(); ++i)
    {
        Surface s;
        s.type = types[i].s;
        auto n = static_cast<std::size_t>(sizes[i]);
        s.d.assign(data.begin() + off, data.begin() + off + n);
        off += n;
        g.surfaces.push_back(s);
    }
    auto const& cells = u("cells").a;
    g.neighbors.resize(g.surfaces.size());
    for (std::size_t i = 0; i < cells.size(); ++i)
    {
        Volume v;
        for (auto f : cells[i]("faces").nums())
            v.faces.push_back(static_cast<std::uint32_t>(f));
        std::istringstream is(cells[i]("logic").s);
        std::string tok;
        while (is >> tok)
        {
            if (tok == "~")
                v.logic.push_back(-6);
            else if (tok == "&")
                v.logic.push_back(-5);
            else if (tok == "|")
                v.logic.push_back(-4);
            else if (tok == "*")
                v.logic.push_back(-3);
            else
                v.logic.push_back(std::stoi(tok));
        }
        v.flags = static_cast<int>(cells[i].num("flags", 0)) & 3;
        if (cells[i].num("zorder", 2) == 1)
            g.background = static_cast<std::uint32_t>(i);
        if (auto names = u.find("cell_names"))
            v.label = names->a.at(i).s;
        for (auto f : v.faces)
            g.neighbors[f].push_back(static_cast<std::uint32_t>(i));
        g.volumes.push_back(v);
    }
    return g;
}

//---------------------------------------------------------------------------//
void geo_initialize(Geometry const& g, GeoState& s, double const* pos, double const* dir)
{
    for (int i = 0; i < 3; ++i)
    {
        s.pos[i] = pos[i];
        s.dir[i] = dir[i];
    }
    s.failed = false;
    std::uint32_t found = none;
    bool on_surface = false;
    for (std::uint32_t v = 0; v < g.volumes.size(); ++v)
    {
        if (g.volumes[v].flags & 2)
            continue;
        int on_face;
        auto sen = senses_for(g, g.volumes[v], pos, -1, false, &on_face);
        if (evaluate(g.volumes[v].logic, sen))
        {
            found = v;
            on_surface = on_face >= 0;
            break;
        }
    }
    if (on_surface)
        found = none;
    else if (found == none)
        found = g.background;
    if (found == none)
    {
        s.failed = true;
        found = 0;
    }
    s.vol = found;
    s.surf = none;
    s.reentrant = false;
    s.next_step = 0;
    s.next_surf = none;
}

double geo_find_next_step(Geometry const& g, GeoState& s, double max_step, bool* boundary)
{
    if (s.reentrant)
    {
        *boundary = true;
        return 0;
    }
    bool const limited = max_step > 0;
    Volume const& v = g.volumes[s.vol];
    int on_face = (s.surf != none) ? find_face(v, s.surf) : -1;
    bool const simple = !(v.flags & 3);
    struct Isect
    {
        double dist;
        int face;
    };
    std::vector<Isect> isects;
    for (std::size_t f = 0; f < v.faces.size(); ++f)
    {
        Surface const& surf = g.surfaces[v.faces[f]];
        bool on = (static_cast<int>(f) == on_face);
        bool single = surf.type[0] == 'p';
        if (on && single)
            continue;
        double d[2];
        int n = calc_intersections(surf, s.pos, s.dir, on, d);
        for (int k = 0; k < n; ++k)
        {
            bool valid = limited ? (d[k] <= max_step) : std::isfinite(d[k]);
            if (valid)
                isects.push_back({d[k], static_cast<int>(f)});
        }
    }
    std::uint32_t next_surf = none;
    bool next_sense = false;
    double dist = inf;
    if (!isects.empty())
    {
        if (simple)
        {
            auto it = std::min_element(isects.begin(), isects.end(),
                                       [](Isect a, Isect b) { return a.dist < b.dist; });
            std::uint32_t sid = v.faces[it->face];
            bool cur = (sid == s.surf) ? s.sense
                                       : to_sense(calc_sense(g.surfaces[sid], s.pos));
            next_surf = sid;
            next_sense = cur;
            dist = it->dist;
        }
        else
        {
            std::stable_sort(isects.begin(), isects.end(),
                             [](Isect a, Isect b) { return a.dist < b.dist; });
            int dummy;
            auto sen = senses_for(g, v, s.pos, on_face, s.sense, &dummy);
            for (auto const& is : isects)
            {
                sen[is.face] = !sen[is.face];
                if (!evaluate(v.logic, sen))
                {
                    next_surf = v.faces[is.face];
                    next_sense = !sen[is.face];
                    dist = is.dist;
                    break;
                }
            }
        }
    }
    if (next_surf == none)
    {
        s.next_step = limited ? max_step : inf;
        s.next_surf = none;
        *boundary = false;
        return s.next_step;
    }
    s.next_step = dist;
    s.next_surf = next_surf;
    s.next_sense = next_sense;
    *boundary = true;
    return dist;
}

void geo_move_to_boundary(GeoState& s)
{
    for (int i = 0; i < 3; ++i)
        s.pos[i] = std::fma(s.next_step, s.dir[i], s.pos[i]);
    s.surf = s.next_surf;
    s.sense = s.next_sense;
    s.next_step = 0;
    s.next_surf = none;
}

void geo_move_internal(GeoState& s, double dist)
{
    for (int i = 0; i < 3; ++i)
        s.pos[i] = std::fma(dist, s.dir[i], s.pos[i]);
    s.next_step -= dist;
    s.surf = none;
}

void geo_move_internal_to(GeoState& s, double const* pos)
{
    for (int i = 0; i < 3; ++i)
        s.pos[i] = pos[i];
    s.surf = none;
    s.next_step = 0;
    s.next_surf = none;
}

void geo_cross_boundary(Geometry const& g, GeoState& s)
{
    if (s.reentrant)
    {
        s.reentrant = false;
        return;
    }
    s.sense = !s.sense;
    std::uint32_t found = none;
    for (std::uint32_t nv : g.neighbors[s.surf])
    {
        if (nv == s.vol)
            continue;
        Volume const& v = g.volumes[nv];
        int on_face;
        auto sen = senses_for(g, v, s.pos, find_face(v, s.surf), s.sense, &on_face);
        if (evaluate(v.logic, sen))
        {
            found = nv;
            break;
        }
    }
    if (found == none)
        found = g.background;
    if (found == none)
    {
        s.failed = true;
        found = 0;
    }
    s.vol = found;
}

void geo_set_dir(Geometry const& g, GeoState& s, double const* dir)
{
    if (s.surf != none)
    {
        double n[3];
        calc_normal(g.surfaces[s.surf], s.pos, n);
        if ((dot(n, dir) >= 0) != (dot(n, s.dir) >= 0))
            s.reentrant = !s.reentrant;
    }
    for (int i = 0; i < 3; ++i)
        s.dir[i] = dir[i];
    s.next_step = 0;
    s.next_surf = none;
}

double geo_find_safety(Geometry const& g, GeoState const& s)
{
    Volume const& v = g.volumes[s.vol];
    if (!v.simple_safety)
        return 0;
    double result = inf;
    for (auto sid : v.faces)
    {
        Surface const& surf = g.surfaces[sid];
        double n[3];
        calc_normal(surf, s.pos, n);
        double d;
        if (std::isnan(n[0]))
        {
            d = inf;
        }
        else
        {
            SignedSense ss = calc_sense(surf, s.pos);
            if (ss == SignedSense::on)
            {
                d = 0;
            }
            else
            {
                if (ss == SignedSense::outside)
                    for (double& c : n)
                        c = -c;
                double isect[2];
                calc_intersections(surf, s.pos, n, false, isect);
                d = std::min(isect[0], isect[1]);
            }
        }
        result = std::min(result, d);
    }
    return result;
}

//---------------------------------------------------------------------------//
}  // namespace orc
