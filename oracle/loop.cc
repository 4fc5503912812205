//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see loop.hh).
//---------------------------------------------------------------------------//
#include "loop.hh"

#include <algorithm>
#include <stdexcept>

#include "physics.hh"
#include "problem.hh"
#include "user.hh"

namespace orc
{
//---------------------------------------------------------------------------//
void run_action(State& s, std::string const& label)
{
    ActionIds const& ids = s.problem->actions;
    if (label == "extend-from-primaries")
        return extend_from_primaries(s);
    if (label == "initialize-tracks")
        return initialize_tracks(s);
    if (label == "pre-step")
        return pre_step(s);
    if (label == "along-step-neutral")
        return along_step_neutral(s);
    if (label == "physics-discrete-select")
        return discrete_select(s);
    if (label == "geo-boundary")
        return boundary(s);
    if (label == "tracking-cut")
        return tracking_cut(s);
    if (label == "extend-from-secondaries")
        return extend_from_secondaries(s);
    for (std::size_t i = 0; i < ids.labels.size(); ++i)
    {
        if (ids.labels[i] == label
            && std::find(ids.model.begin(), ids.model.end(), i) != ids.model.end())
            return interact(s, static_cast<std::uint32_t>(i));
    }
    throw std::runtime_error("oracle: unknown action '" + label + "'");
}

//---------------------------------------------------------------------------//
StepResult step_once(State& s)
{
    ActionIds const& ids = s.problem->actions;
    s.counters.num_generated = 0;
    extend_from_primaries(s);
    initialize_tracks(s);
    pre_step(s);
    user_gather_pre(s);
    along_step_neutral(s);
    if (ids.along_user != ids.along_neutral)
        along_step_charged(s);
    discrete_select(s);
    std::vector<std::uint32_t> models = ids.model;
    std::sort(models.begin(), models.end());
    for (auto m : models)
        interact(s, m);
    boundary(s);
    tracking_cut(s);
    user_action_diag(s);
    calo_tally(s);
    user_post(s);
    extend_from_secondaries(s);
    return {s.counters.num_generated, s.counters.num_initializers,
            s.counters.num_active, s.counters.num_alive};
}

//---------------------------------------------------------------------------//
}  // namespace orc
