//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see core.hh).
//---------------------------------------------------------------------------//
#include "core.hh"

#include <algorithm>
#include <stdexcept>

#include "physics.hh"
#include "problem.hh"

namespace orc
{
namespace
{
//---------------------------------------------------------------------------//
//! SimTrackView::operator=(SimTrackInitializer) (SimTrackView.hh:148-164)
void assign_sim(Track& t, Initializer const& init)
{
    t.track_id = init.track_id;
    t.parent_id = init.parent_id;
    t.event_id = init.event_id;
    t.num_steps = 0;
    t.time = init.time;
    t.status = initializing;
    t.step_length = 0;
    t.post_action = none;
    t.along_action = none;
    t.num_looping_steps = 0;
}

//! PhysicsTrackView::operator=(Initializer) (PhysicsTrackView.hh:207-212)
void reset_physics(Track& t)
{
    t.interaction_mfp = 0;
    t.msc_range_init = t.msc_range_factor = t.msc_limit_min = 0;
}

//! CoreTrackView::apply_errored (CoreTrackView.hh:340-347)
void apply_errored(State const& s, Track& t)
{
    t.status = errored;
    t.along_action = none;
    t.post_action = s.problem->actions.tracking_cut;
}

std::uint32_t make_track_id(State& s, std::uint32_t event)
{
    if (event >= s.track_counters.size())
        throw std::runtime_error("event id exceeds max_events");
    return s.track_counters[event]++;
}
//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
std::unique_ptr<State> make_state(std::shared_ptr<Problem const> problem,
                                  std::uint32_t num_slots,
                                  std::uint32_t stream_id)
{
    auto s = std::make_unique<State>();
    s->problem = problem;
    s->tracks.resize(num_slots);
    auto rng = seeded_states(problem->seed, stream_id, num_slots);
    for (std::uint32_t i = 0; i < num_slots; ++i)
        s->tracks[i].rng = rng[i];
    s->initializers.resize(problem->init_capacity);
    s->vacancies.resize(num_slots);
    for (std::uint32_t i = 0; i < num_slots; ++i)
        s->vacancies[i] = i;
    s->secondary_counts.assign(num_slots + 1, 0);
    s->track_counters.assign(problem->max_events, 0);
    s->stack.resize(static_cast<std::size_t>(
        std::max(1.0, problem->secondary_stack_factor) * num_slots));
    s->counters.num_vacancies = num_slots;
    s->in_place = problem->in_place;
    s->calo_edep.assign(std::max<std::uint32_t>(problem->num_detectors, 1), 0.0);
    return s;
}

//---------------------------------------------------------------------------//
void insert_primaries(State& s, std::vector<PrimaryP> const& p)
{
    if (!s.pending_primaries.empty())
        throw std::runtime_error("multiple consecutive primary insertions");
    if (p.size() + s.counters.num_initializers > s.initializers.size())
        throw std::runtime_error("insufficient initializer capacity");
    s.pending_primaries = p;
}

//! ExtendFromPrimariesAction::step_impl + ProcessPrimariesExecutor
void extend_from_primaries(State& s)
{
    auto const count = static_cast<std::uint32_t>(s.pending_primaries.size());
    s.counters.num_initializers += count;
    for (std::uint32_t i = 0; i < count; ++i)
    {
        PrimaryP const& p = s.pending_primaries[i];
        Initializer& ti = s.initializers[s.counters.num_initializers - count + i];
        ti.track_id = make_track_id(s, p.event);
        ti.parent_id = none;
        ti.event_id = p.event;
        ti.time = p.time;
        for (int k = 0; k < 3; ++k)
        {
            ti.pos[k] = p.pos[k];
            ti.dir[k] = p.dir[k];
        }
        ti.particle = p.particle;
        ti.energy = p.energy;
    }
    s.counters.num_generated += count;
    s.pending_primaries.clear();
}

/*!
 * InitializeTracksAction::step_impl + InitTracksExecutor.
 *
 * TrackOrder::none pops initializers and vacancies from the back
 * (InitTracksExecutor.hh:71-104). With init_charge (in_place false) the new
 * initializers' indices are stably partitioned neutral-first
 * (InitializeTracksAction.cc:66-78, TrackInitAlgorithms.cc:80-94); thread
 * tid takes initializer indices[n-1-tid] and a neutral track fills vacancy
 * n-1-tid from the front, a charged one vacancy V-1-tid from the back
 * (Utils.hh:88-98).
 */
void initialize_tracks(State& s)
{
    Counters& c = s.counters;
    std::uint32_t const n = std::min(c.num_vacancies, c.num_initializers);
    Problem const& prob = *s.problem;
    bool const by_charge = !s.in_place;
    std::vector<std::uint32_t> indices;
    auto is_neutral = [&prob](Initializer const& ti) {
        return prob.physics->particles[ti.particle].charge == 0;
    };
    if (by_charge)
    {
        indices.resize(n);
        for (std::uint32_t i = 0; i < n; ++i)
            indices[i] = i;
        Initializer const* stencil = s.initializers.data() + c.num_initializers - n;
        std::stable_partition(indices.begin(), indices.end(),
                              [&](std::uint32_t i) { return is_neutral(stencil[i]); });
    }
    for (std::uint32_t tid = 0; tid < n; ++tid)
    {
        std::uint32_t const init_idx
            = by_charge ? indices[n - 1 - tid] + c.num_initializers - n
                        : c.num_initializers - tid - 1;
        Initializer const& init = s.initializers[init_idx];
        std::uint32_t vac_idx = c.num_vacancies - tid - 1;
        if (by_charge && is_neutral(init))
            vac_idx = n - 1 - tid;
        Track& t = s.tracks[s.vacancies[vac_idx]];
        assign_sim(t, init);
        t.particle = init.particle;
        t.energy = init.energy;
        // Parent slots are always cleared by the generate action, so the
        // geometry is initialized from the position
        geo_initialize(prob.geo, t.geo, init.pos, init.dir);
        if (t.geo.failed || geo_is_outside(t.geo))
        {
            apply_errored(s, t);
            continue;
        }
        std::uint32_t mat = t.geo.vol < prob.volume_to_material.size()
                                ? prob.volume_to_material[t.geo.vol]
                                : none;
        if (mat == none)
        {
            apply_errored(s, t);
            continue;
        }
        reset_physics(t);
    }
    c.num_initializers -= n;
    c.num_vacancies -= n;
    c.num_active = static_cast<std::uint32_t>(s.tracks.size()) - c.num_vacancies;
}

//! ExtendFromSecondariesAction::step_impl with LocateAlive/ProcessSecondaries
void extend_from_secondaries(State& s)
{
    std::uint32_t const size = static_cast<std::uint32_t>(s.tracks.size());
    // Locate alive
    for (std::uint32_t i = 0; i < size; ++i)
    {
        Track const& t = s.tracks[i];
        std::uint32_t nsec = 0;
        if (t.status != inactive)
        {
            for (std::uint32_t k = 0; k < t.sec_count; ++k)
                if (s.stack[t.sec_start + k].particle != none)
                    ++nsec;
        }
        if (t.status == alive)
        {
            s.vacancies[i] = none;
        }
        else if (nsec > 0 && s.in_place)
        {
            --nsec;
            s.vacancies[i] = none;
        }
        else
        {
            s.vacancies[i] = i;
        }
        s.secondary_counts[i] = nsec;
    }
    // remove_if(vacancies == occupied): stable
    auto stop = std::remove(s.vacancies.begin(), s.vacancies.end(), none);
    s.counters.num_vacancies = static_cast<std::uint32_t>(stop - s.vacancies.begin());
    std::fill(stop, s.vacancies.end(), none);
    // exclusive scan over size+1 entries
    s.secondary_counts[size] = 0;
    std::uint32_t acc = 0;
    for (auto& v : s.secondary_counts)
    {
        std::uint32_t cur = v;
        v = acc;
        acc += cur;
    }
    s.counters.num_secondaries = s.secondary_counts[size];
    s.counters.num_initializers += s.counters.num_secondaries;
    if (s.counters.num_initializers > s.initializers.size())
        throw std::runtime_error("insufficient capacity for track initializers");
    s.counters.num_alive = size - s.counters.num_vacancies;

    // Process secondaries
    Counters const& c = s.counters;
    for (std::uint32_t i = 0; i < size; ++i)
    {
        Track& t = s.tracks[i];
        if (t.status == inactive)
            continue;
        std::uint32_t offset = c.num_secondaries - s.secondary_counts[i];
        bool initialized = false;
        std::uint32_t const parent_id = t.track_id;
        for (std::uint32_t k = 0; k < t.sec_count; ++k)
        {
            SecondaryP const& sec = s.stack[t.sec_start + k];
            if (sec.particle == none)
                continue;
            Initializer ti;
            ti.track_id = make_track_id(s, t.event_id);
            ti.parent_id = parent_id;
            ti.event_id = t.event_id;
            ti.time = t.time;
            for (int d = 0; d < 3; ++d)
            {
                ti.pos[d] = t.geo.pos[d];
                ti.dir[d] = sec.dir[d];
            }
            ti.particle = sec.particle;
            ti.energy = sec.energy;
            if (!initialized && t.status != alive && s.in_place)
            {
                assign_sim(t, ti);
                for (int d = 0; d < 3; ++d)
                    t.geo.dir[d] = ti.dir[d];
                t.geo.failed = false;
                t.geo.next_step = 0;
                t.geo.next_surf = none;
                t.particle = ti.particle;
                t.energy = ti.energy;
                reset_physics(t);
                initialized = true;
            }
            else
            {
                s.initializers[c.num_initializers - offset] = ti;
                --offset;
            }
        }
        if (!initialized && t.status == killed)
            t.status = inactive;
    }
}

//---------------------------------------------------------------------------//
void mock_interact(State& s,
                   std::vector<std::uint32_t> const& num_secondaries,
                   std::vector<char> const& alive_flags)
{
    for (std::uint32_t i = 0; i < s.tracks.size(); ++i)
    {
        Track& t = s.tracks[i];
        if (t.status != alive)
            continue;
        if (!alive_flags[i])
            t.status = killed;
        std::uint32_t n = num_secondaries[i];
        if (n > 0)
        {
            t.sec_start = s.stack_size;
            t.sec_count = n;
            s.stack_size += n;
            for (std::uint32_t k = 0; k < n; ++k)
            {
                SecondaryP& sec = s.stack[t.sec_start + k];
                sec.particle = 0;
                sec.energy = 5.0;
                sec.dir[0] = 1;
                sec.dir[1] = sec.dir[2] = 0;
            }
        }
        else
        {
            t.sec_start = 0;
            t.sec_count = 0;
        }
    }
}

//---------------------------------------------------------------------------//
}  // namespace orc
