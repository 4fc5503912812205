//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see physics.hh).
//---------------------------------------------------------------------------//
#include "physics.hh"

#include <limits>

#include "problem.hh"

namespace orc
{
//---------------------------------------------------------------------------//
void build_physics(Json const& inp, PhysicsData& out)
{
    if (auto parts = inp.find("particles"))
    {
        for (auto const& p : parts->a)
        {
            ParticleDef d;
            d.label = p.str("label", "");
            d.pdg = static_cast<int>(p.num("pdg", 0));
            d.mass = p.num("mass", 0);
            d.charge = p.num("charge", 0);
            out.particles.push_back(d);
        }
    }
    if (auto opts = inp.find("options"))
        out.secondary_stack_factor = opts->num("secondary_stack_factor", 3);
}

void PhysicsData::finalize_actions(ActionIds const& ids)
{
    pre_step_action = ids.pre_step;
}

//---------------------------------------------------------------------------//
void pre_step(State& s)
{
    s.stack_size = 0;
    for (auto& t : s.tracks)
    {
        if (t.status == inactive)
        {
            t.step_length = std::numeric_limits<double>::infinity();
            t.post_action = none;
            t.along_action = none;
            continue;
        }
        t.edep = 0;
        t.sec_start = 0;
        t.sec_count = 0;
        t.element = none;
        if (t.status == errored)
            continue;
        t.status = alive;
    }
}

//---------------------------------------------------------------------------//
}  // namespace orc
