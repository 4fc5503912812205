//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/CoreData.hh
//! Plain-old-data views of the core params and per-slot state.
//!
//! Counterparts of CoreParamsData / CoreStateData
//! (celeritas/global/CoreTrackData.hh:66-160). Every per-slot field is its
//! own 256-byte-aligned array so that a wave64 access is coalesced; kernels
//! receive these structs by value (kernel argument segment), so there is no
//! device-side indirection to reach a field.
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Types.hh"
#include "field/FieldData.hh"
#include "geo/GeoData.hh"
#include "phys/PhysicsData.hh"
#include "random/XorwowEngine.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Track initializer (celeritas/track/TrackInitData.hh:66-80)
struct TrackInitializer
{
    std::uint32_t track_id;
    std::uint32_t parent_id;
    std::uint32_t event_id;
    std::uint32_t particle_id;
    double time;
    double pos[3];
    double dir[3];
    double energy;
    //! Volume of the parent's interior it was made in (invalid: locate the
    //! track from its position)
    std::uint32_t volume;
    std::uint32_t reserved;
};

//! Secondary particle (celeritas/phys/Secondary.hh)
struct Secondary
{
    std::uint32_t particle_id;  //!< invalid_id: rejected by cutoff
    std::uint32_t pad_;
    double energy;
    double dir[3];
};

//! Primary particle (celeritas/phys/Primary.hh:23-31)
struct Primary
{
    std::uint32_t particle_id;
    std::uint32_t event_id;
    double energy;
    double pos[3];
    double dir[3];
    double time;
};

//---------------------------------------------------------------------------//
/*!
 * Device-resident counters (celeritas/track/CoreStateCounters.hh:24-47).
 *
 * The host reads them asynchronously; kernels read and update them so that
 * no step requires a host round trip.
 */
struct CoreCounters
{
    std::uint32_t num_initializers;
    std::uint32_t num_vacancies;
    std::uint32_t num_active;
    std::uint32_t num_secondaries;
    std::uint32_t num_alive;
    std::uint32_t num_generated;
    // Launch parameters computed on device for the next kernel
    std::uint32_t num_new_tracks;  //!< this step's min(vacancies, inits)
    std::uint32_t init_base;  //!< initializer count before the launch
    std::uint32_t vacancy_base;  //!< vacancy count before the launch
    std::uint32_t prev_secondaries;  //!< secondaries at the back of inits
    std::uint32_t secondary_stack_size;  //!< allocated secondaries
    std::uint32_t error_flags;  //!< bit 0: initializer overflow, bit 1: stack
    std::uint32_t steps_taken;  //!< step iterations run on the device
    std::uint32_t num_new_neutral;  //!< neutral initializers this launch
    //! Tracks failed because a sampling loop ran out of its iteration budget
    std::uint32_t num_sampling_failures;
    //! Events completed on the device (entries of InitStateRef::completed)
    std::uint32_t num_completed;
    //! Sum over step iterations of the active tracks (track-steps)
    unsigned long long track_steps;
};

enum CoreErrorFlags : std::uint32_t
{
    err_initializer_overflow = 1u,
    err_secondary_overflow = 2u,
    err_event_overflow = 4u,
    //! A rejection loop exceeded its iteration budget (the track failed;
    //! not fatal to the run)
    err_sampling_budget = 8u,
};

//---------------------------------------------------------------------------//
//! Simulation state (celeritas/track/SimData.hh:119-175)
struct SimStateRef
{
    std::uint32_t* track_id{nullptr};
    std::uint32_t* parent_id{nullptr};
    std::uint32_t* event_id{nullptr};
    std::uint32_t* num_steps{nullptr};
    double* time{nullptr};
    std::uint8_t* status{nullptr};
    double* step_length{nullptr};
    std::uint8_t* post_step_action{nullptr};
    std::uint8_t* along_step_action{nullptr};
    std::uint32_t* num_looping_steps{nullptr};  //!< consecutive looping steps
};

//! Particle state (celeritas/phys/ParticleData.hh:119-135)
struct ParticleStateRef
{
    std::uint8_t* particle_id{nullptr};
    double* energy{nullptr};
};

//! Track initialization state (celeritas/track/TrackInitData.hh:100-140)
struct InitStateRef
{
    std::uint32_t* vacancies{nullptr};
    std::uint32_t* secondary_counts{nullptr};  //!< size+1
    std::uint32_t* indices{nullptr};
    std::uint32_t* track_counters{nullptr};  //!< per event
    //! Per event: initializers plus live tracks (0 once the event is done)
    std::int32_t* event_live{nullptr};
    double* event_edep{nullptr};  //!< per event: deposited energy [MeV]
    //! Events in the order their live count reached zero (num_completed)
    std::uint32_t* completed{nullptr};
    TrackInitializer* initializers{nullptr};
    std::uint32_t* block_scratch{nullptr};  //!< compaction partials
    size_type capacity{0};
    size_type max_events{0};
    //! 1 if event_live and event_edep are maintained (per-event completion
    //! for the serving loop); plain transport skips their atomics
    std::uint32_t event_tally{0};
};

//! Secondary storage (corecel/data/StackAllocator.hh)
struct SecondaryStackRef
{
    Secondary* storage{nullptr};
    size_type capacity{0};
};

//---------------------------------------------------------------------------//
//! Admission counters of the serving queue
struct ServeCounters
{
    std::uint32_t arrived;  //!< events staged so far (written by the host)
    std::uint32_t admitted;  //!< events admitted so far (device)
    std::uint32_t first;  //!< first primary admitted this step
    std::uint32_t count;  //!< primaries admitted this step
};

/*!
 * Events waiting for the device to admit them (the serving loop's queue).
 *
 * The host stages an arriving event's primaries at its place in
 * \c primaries (events lie in arrival order, each contiguous, with the
 * event index as event ID) and then raises \c arrived; at the start of
 * every step the device admits arrived events in order while the queued
 * initializers stay within \c max_queued (or the first of them when
 * nothing is queued or alive) and turns their primaries into initializers.
 */
struct ServeQueueRef
{
    Primary const* primaries{nullptr};
    std::uint32_t const* event_first{nullptr};  //!< per event: first primary
    std::uint32_t const* event_count{nullptr};  //!< per event: primaries
    std::uint32_t* admit_step{nullptr};  //!< per event: step it was admitted
    ServeCounters* counters{nullptr};  //!< null: no queue attached
    size_type num_events{0};  //!< capacity of the event tables
    size_type max_queued{0};
};

//---------------------------------------------------------------------------//
/*!
 * All per-slot mutable state.
 */
struct CoreStateRef
{
    size_type size{0};
    std::uint32_t stream_id{0};
    //! Workgroup size of the step kernels (64, 128 or 256; autotuned)
    std::uint32_t block_size{256};
    SimStateRef sim;
    ParticleStateRef particle;
    GeoStateRef geo;
    PhysicsStateRef phys;
    RngStateRef rng;
    InitStateRef init;
    SecondaryStackRef secondaries;
    CoreCounters* counters{nullptr};
    std::uint32_t* pre_volume{nullptr};  //!< volume at the start of the step
    std::uint32_t* track_slots{nullptr};  //!< thread -> slot indirection
    std::uint32_t* action_offsets{nullptr};  //!< sorted ranges per key
    std::uint32_t* sort_counts{nullptr};  //!< key histogram per sort tile
    double* calo_edep{nullptr};  //!< per-volume energy deposition tally
    //! Per-workgroup partial tallies [detectors][workgroups of the pass]
    double* calo_partial{nullptr};
    //! Optional work counters (null unless diagnostics are enabled), see
    //! DiagCounter
    unsigned long long* diag{nullptr};
    //! Serving queue admitted from on the device (empty unless attached)
    ServeQueueRef serve;
};

//! Largest number of sort keys (actions + 1 for inactive slots)
inline constexpr unsigned int max_sort_keys = 64;

//! Workgroups of the calorimeter tally's first pass
inline constexpr unsigned int calo_tally_blocks = 1024;
//! Detector bins that the tally reduces in LDS and in two passes
inline constexpr unsigned int max_lds_detectors = 512;

//! Per-slot key of a track reindexing
enum class SortKey : std::uint8_t
{
    status,  //!< active slots first (TrackOrder::reindex_status)
    particle,  //!< by particle type (reindex_particle_type)
    along_action,  //!< by along-step action (reindex_along_step_action)
    post_action,  //!< by post-step action (reindex_step_limit_action)
    //! By the along-step action the pre-step will choose: the particle's
    //! (a function of its charge) for an alive or initializing track, none
    //! for an inactive or errored one. Sorting before the pre-step with it
    //! gives the permutation that sorting by along_action after it gives.
    along_action_of_particle,
};

/*!
 * Thread-to-slot map of a launch (ActionLauncher.device.hh:120-135 with the
 * track_slots indirection, ThreadId -> TrackSlotId).
 *
 * With a range (an action-sorted launch) the thread count is read from the
 * device-resident offsets, so launches never wait for the host; threads
 * beyond the range exit at once.
 */
struct ThreadSlots
{
    std::uint32_t const* range{nullptr};  //!< &action_offsets[action], or null
    std::uint32_t const* slots{nullptr};  //!< thread -> slot, or null (identity)
    //! End of a range spanning several keys (&action_offsets[last + 1]);
    //! null: the range is the one key's, [range[0], range[1])
    std::uint32_t const* range_end{nullptr};
};

//! Number of workgroups covering the state's slots
inline unsigned int grid_of(CoreStateRef const& s)
{
    return (s.size + s.block_size - 1) / s.block_size;
}

/*!
 * Largest number of track slots in a state.
 *
 * Slot indices are masked to this many bits where a kernel obtains them, so
 * the compiler knows the byte offset of any per-slot element fits 32 bits:
 * every SoA access of a kernel then shares one 32-bit offset register with
 * the array base in scalar registers, instead of a 64-bit address in vector
 * registers per array.
 */
inline constexpr size_type max_track_slots = size_type(1) << 28;

CELER_FORCEINLINE_FUNCTION bool
thread_to_slot(ThreadSlots const& ts, size_type tid, size_type size, size_type* slot)
{
    constexpr size_type mask = max_track_slots - 1;
    if (ts.range)
    {
        size_type const begin = ts.range[0];
        size_type const end = ts.range_end ? *ts.range_end : ts.range[1];
        if (tid >= end - begin)
            return false;
        *slot = ts.slots[begin + tid] & mask;
        return true;
    }
    if (tid >= size)
        return false;
    *slot = (ts.slots ? ts.slots[tid] : tid) & mask;
    return true;
}

//! Indices of the optional per-kernel work counters
enum DiagCounter : int
{
    diag_pre_step = 0,  //!< tracks processed by pre-step
    diag_along_charged,  //!< tracks processed by the charged along-step
    diag_along_neutral,  //!< tracks processed by the neutral along-step
    diag_boundary,  //!< boundary crossings
    diag_discrete,  //!< discrete selections
    diag_interaction,  //!< interactor calls
    //! discrete selections made inside the charged along-step kernel
    diag_discrete_charged,
    num_diag_counters = 8
};

//! Copies of each work counter, one per workgroup index modulo this: the
//! waves of a launch add to different addresses instead of contending for
//! one (row c of the counters holds counter c's copies)
inline constexpr unsigned int diag_spread = 256;

//---------------------------------------------------------------------------//
//! Action IDs of the core actions, known to the kernels
struct CoreScalars
{
    std::uint8_t boundary_action{invalid_id8};
    std::uint8_t propagation_limit_action{invalid_id8};
    std::uint8_t tracking_cut_action{invalid_id8};
    std::uint8_t along_step_user_action{invalid_id8};
    std::uint8_t along_step_neutral_action{invalid_id8};
    std::uint8_t initialize_tracks_action{invalid_id8};
    //! 1: the along-step kernels also run the discrete selection of the
    //! tracks whose step ends in it (the discrete-select action then
    //! launches nothing)
    std::uint8_t fuse_discrete_select{0};
    std::uint8_t pad_;
    std::uint32_t max_steps{invalid_id};
    std::uint32_t num_actions{0};
    TrackOrder track_order{TrackOrder::none};
    PrecisionMode precision{PrecisionMode::fp64};
};

/*!
 * Immutable shared data.
 */
struct CoreParamsRef
{
    CoreScalars scalars;
    GeoParamsRef geo;
    std::uint32_t const* volume_to_material{nullptr};
    std::uint32_t const* volume_to_detector{nullptr};  //!< calo tally map
    PhysicsParamsRef phys;
    FieldParamsRef field;  //!< uniform magnetic field (disabled: none)
    XorwowJumpTables const* rng_jump{nullptr};
    std::uint32_t rng_seed{0};
    //! Along-step action the pre-step chooses for each particle type
    std::uint8_t const* particle_along_action{nullptr};
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
