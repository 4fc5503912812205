//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see user.hh).
//---------------------------------------------------------------------------//
#include "user.hh"

#include <algorithm>

#include "problem.hh"

namespace orc
{
//---------------------------------------------------------------------------//
int g_user_mutation = 0;

void user_gather_pre(State& s)
{
    if (!s.user || !s.user->collect)
        return;
    UserObservers& u = *s.user;
    u.slot_detector.assign(s.tracks.size(), none);
    u.pre.resize(s.tracks.size());
    for (std::size_t i = 0; i < s.tracks.size(); ++i)
    {
        Track const& t = s.tracks[i];
        if (t.status == inactive)
            continue;
        std::uint32_t const det
            = u.detector_map.empty() ? 0u : u.detector_map.at(t.geo.vol);
        u.slot_detector[i] = det;
        if (det == none)
            continue;
        StepHit& h = u.pre[i];
        h.time[0] = t.time;
        std::copy(t.geo.pos, t.geo.pos + 3, h.pos[0]);
        std::copy(t.geo.dir, t.geo.dir + 3, h.dir[0]);
        h.volume[0] = t.geo.vol;
        h.energy[0] = t.energy;
    }
}

void user_action_diag(State& s)
{
    if (!s.user || !s.user->action_diag)
        return;
    UserObservers& u = *s.user;
    for (Track const& t : s.tracks)
    {
        if (t.status == inactive || t.post_action == none || t.post_action >= u.num_actions)
            continue;
        if (g_user_mutation == 1 && t.status == killed)
            continue;
        ++u.action_counts.at(std::size_t(t.particle) * u.num_actions + t.post_action);
    }
}

void user_post(State& s)
{
    if (!s.user)
        return;
    UserObservers& u = *s.user;
    if (u.step_diag)
    {
        for (Track const& t : s.tracks)
        {
            if (t.status != killed)
                continue;
            std::uint32_t bin = std::min(t.num_steps, u.step_bins - 1);
            if (g_user_mutation == 2 && bin > 0)
                --bin;
            ++u.step_counts.at(std::size_t(t.particle) * u.step_bins + bin);
        }
    }
    if (u.collect)
    {
        for (std::size_t i = 0; i < s.tracks.size(); ++i)
        {
            Track const& t = s.tracks[i];
            if (t.status == inactive || u.slot_detector.at(i) == none)
                continue;
            if (u.nonzero_edep && t.edep == 0)
                continue;
            StepHit h = u.pre[i];
            if (g_user_mutation == 3)
                std::copy(t.geo.pos, t.geo.pos + 3, h.pos[0]);
            h.detector = u.slot_detector[i];
            h.track_id = t.track_id;
            h.event_id = t.event_id;
            h.parent_id = t.parent_id;
            h.track_step_count = t.num_steps;
            h.action_id = t.post_action;
            h.step_length = t.step_length;
            h.particle = t.particle;
            h.edep = t.edep;
            h.time[1] = t.time;
            std::copy(t.geo.pos, t.geo.pos + 3, h.pos[1]);
            std::copy(t.geo.dir, t.geo.dir + 3, h.dir[1]);
            h.volume[1] = geo_is_outside(t.geo) ? none : t.geo.vol;
            h.energy[1] = t.energy;
            u.hits.push_back(h);
        }
    }
}

//---------------------------------------------------------------------------//
}  // namespace orc
