//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/core.hh
//! Array-of-structs track state and the track-management actions, restated
//! from celeritas/track/detail/{ProcessPrimaries,InitTracks,LocateAlive,
//! ProcessSecondaries}Executor.hh, track/detail/TrackInitAlgorithms.cc
//! (std::remove_if / exclusive scan on the host), track/*Action.cc and
//! celeritas/track/SimTrackView.hh. Loops run serially over slots in slot
//! order, which fixes the order of track IDs from the atomic counters.
//!
//! Pinned by: TrackInit.test.cc:228-500 (vacancies, initializer IDs, track
//! and parent IDs for the 'run', 'primaries' and 'extend_from_secondaries'
//! scenarios with MockInteractExecutor).
//---------------------------------------------------------------------------//
#pragma once

#include <cstdint>
#include <memory>
#include <vector>

#include "geo.hh"
#include "rng.hh"

namespace orc
{
//---------------------------------------------------------------------------//
enum Status : std::uint8_t
{
    inactive = 0,
    initializing = 1,
    alive = 2,
    errored = 3,
    killed = 4
};

struct Initializer
{
    std::uint32_t track_id = none;
    std::uint32_t parent_id = none;
    std::uint32_t event_id = none;
    std::uint32_t particle = none;
    double time = 0;
    double pos[3] = {0, 0, 0};
    double dir[3] = {0, 0, 0};
    double energy = 0;
};

struct SecondaryP
{
    std::uint32_t particle = none;  // none: cut
    double energy = 0;
    double dir[3] = {0, 0, 0};
};

struct PrimaryP
{
    std::uint32_t particle;
    std::uint32_t event;
    double energy;
    double pos[3];
    double dir[3];
    double time;
};

//! One track slot (all fields of the reference's SoA collections)
struct Track
{
    // sim
    std::uint32_t track_id = none, parent_id = none, event_id = none;
    std::uint32_t num_steps = 0;
    double time = 0;
    Status status = inactive;
    double step_length = 0;
    std::uint32_t post_action = none, along_action = none;
    std::uint32_t num_looping_steps = 0;
    // particle
    std::uint32_t particle = none;
    double energy = 0;
    // geometry
    GeoState geo;
    // physics (persistent)
    double interaction_mfp = 0;
    double msc_range_init = 0, msc_range_factor = 0, msc_limit_min = 0;
    // physics (step)
    double macro_xs = 0;
    double per_process_xs[4] = {0, 0, 0, 0};
    double dedx_range = 0;
    double edep = 0;
    std::uint32_t sec_start = 0, sec_count = 0;
    std::uint32_t element = none;
    std::uint32_t pre_volume = none;  // volume at the start of the step
    // msc step
    double msc_true_path = 0, msc_geom_path = 0, msc_alpha = -1;
    bool msc_is_displaced = false, msc_enabled = false;
    // rng
    XorwowState rng;
};

struct Counters
{
    std::uint32_t num_initializers = 0, num_vacancies = 0, num_active = 0;
    std::uint32_t num_secondaries = 0, num_alive = 0, num_generated = 0;
};

struct Problem;  // geometry + physics (defined by the stepping-loop oracle)
struct UserObservers;  // diagnostics and step gather (user.hh)

struct State
{
    std::shared_ptr<Problem const> problem;
    std::vector<Track> tracks;
    std::vector<Initializer> initializers;  // capacity-sized
    std::vector<std::uint32_t> vacancies;
    std::vector<std::uint32_t> secondary_counts;  // size+1
    std::vector<std::uint32_t> track_counters;
    std::vector<SecondaryP> stack;  // secondary storage
    std::uint32_t stack_size = 0;
    std::vector<PrimaryP> pending_primaries;
    Counters counters;
    bool in_place = true;  // TrackOrder != init_charge
    std::vector<double> calo_edep;
    double total_edep = 0;  //!< deposited in all volumes [MeV]
    //! Optional user observers (null: none)
    std::shared_ptr<UserObservers> user;
};

//! Build an empty state (all slots vacant, RNG seeded from the MT stream)
std::unique_ptr<State> make_state(std::shared_ptr<Problem const> problem,
                                  std::uint32_t num_slots,
                                  std::uint32_t stream_id);

// Actions
void insert_primaries(State& s, std::vector<PrimaryP> const& p);
void extend_from_primaries(State& s);
void initialize_tracks(State& s);
void extend_from_secondaries(State& s);
//! MockInteractExecutor (test/celeritas/track/MockInteractExecutor.hh)
void mock_interact(State& s,
                   std::vector<std::uint32_t> const& num_secondaries,
                   std::vector<char> const& alive);

//---------------------------------------------------------------------------//
}  // namespace orc
