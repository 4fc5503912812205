//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see loop.hh).
//---------------------------------------------------------------------------//
#include "loop.hh"

#include <stdexcept>

#include "physics.hh"
#include "problem.hh"

namespace orc
{
//---------------------------------------------------------------------------//
void run_action(State& s, std::string const& label)
{
    if (label == "extend-from-primaries")
        extend_from_primaries(s);
    else if (label == "initialize-tracks")
        initialize_tracks(s);
    else if (label == "pre-step")
        pre_step(s);
    else if (label == "extend-from-secondaries")
        extend_from_secondaries(s);
    else
        throw std::runtime_error("oracle: unknown action '" + label + "'");
}
//---------------------------------------------------------------------------//
}  // namespace orc
