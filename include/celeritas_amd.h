/*----------------------------------*-C-*------------------------------------*
 * celeritas_amd: MI355X-native per-step transport loop — C ABI
 *
 * Plain C entry points (no C++ or torch types) for binding from any
 * language (ctypes, cgo, JNI, N-API). Every function returns 0 on success
 * and a nonzero code on failure; constructors return NULL on failure.
 * celer_amd_last_error() returns the message of the last failure on the
 * calling thread. Handles are opaque.
 *
 * Reference interfaces each group replaces (paths relative to the
 * celeritas source tree):
 *   params    celeritas/global/CoreParams.hh:68-172 (CoreParams(Input))
 *   state     celeritas/global/CoreState.hh:74-182 (CoreState<device>)
 *   actions   corecel/sys/ActionInterface.hh:176-186 (step(params, state))
 *             corecel/sys/ActionRegistry.hh:42-105 (find_action, labels)
 *   stepper   celeritas/global/Stepper.hh:43-201 (Stepper<device>)
 *   offload   accel/LocalTransporter.hh:52-109 (Push/Flush/InitializeEvent)
 *   launch    celeritas/track/{ExtendFromPrimaries,InitializeTracks,ExtendFromSecondaries}Action.cu, celeritas/random/RngReseed.cu:59
 *---------------------------------------------------------------------------*/
#ifndef CELERITAS_AMD_H
#define CELERITAS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/*---------------------------------------------------------------------------*
 * PLAIN DATA (layouts match the device structs)
 *---------------------------------------------------------------------------*/
/* Primary particle: celeritas/phys/Primary.hh:23-31 */
typedef struct CelerPrimary
{
    uint32_t particle_id;
    uint32_t event_id;
    double energy; /* MeV */
    double pos[3]; /* cm */
    double dir[3];
    double time; /* s */
} CelerPrimary;

/* Step counters: celeritas/global/Stepper.hh:61-70 (StepperResult) */
typedef struct CelerStepperResult
{
    uint32_t generated;
    uint32_t queued;
    uint32_t active;
    uint32_t alive;
} CelerStepperResult;

/* Device counters: celeritas/track/CoreStateCounters.hh:24-47 */
typedef struct CelerCounters
{
    uint32_t num_initializers;
    uint32_t num_vacancies;
    uint32_t num_active;
    uint32_t num_secondaries;
    uint32_t num_alive;
    uint32_t num_generated;
    uint32_t num_new_tracks;
    uint32_t init_base;
    uint32_t vacancy_base;
    uint32_t prev_secondaries;
    uint32_t secondary_stack_size;
    uint32_t error_flags;
    uint32_t steps_taken; /* step iterations run on the device */
    uint32_t num_new_neutral; /* neutral tracks in the last charge-partitioned init */
    /* tracks failed because a sampling loop ran out of its iteration budget
     * (error_flags bit 3 is set with it; not fatal to the run) */
    uint32_t num_sampling_failures;
    /* events completed on the device since the serving loop's start */
    uint32_t num_completed;
    /* sum over step iterations of the active tracks */
    unsigned long long track_steps;
} CelerCounters;

/* Track initializer: celeritas/track/TrackInitData.hh:66-80 */
typedef struct CelerTrackInitializer
{
    uint32_t track_id;
    uint32_t parent_id;
    uint32_t event_id;
    uint32_t particle_id;
    double time;
    double pos[3];
    double dir[3];
    double energy;
    uint32_t volume;   /* parent's interior volume, UINT32_MAX: locate */
    uint32_t reserved;
} CelerTrackInitializer;

/* Secondary: celeritas/phys/Secondary.hh (particle_id UINT32_MAX = cut) */
typedef struct CelerSecondary
{
    uint32_t particle_id;
    uint32_t pad_;
    double energy;
    double dir[3];
} CelerSecondary;

/* Per-slot state fields accessible with celer_state_copy_field */
typedef enum CelerStateField
{
    CELER_FIELD_TRACK_ID = 0,        /* u32 */
    CELER_FIELD_PARENT_ID,           /* u32 */
    CELER_FIELD_EVENT_ID,            /* u32 */
    CELER_FIELD_NUM_STEPS,           /* u32 */
    CELER_FIELD_TIME,                /* f64 */
    CELER_FIELD_STATUS,              /* u8  (TrackStatus) */
    CELER_FIELD_STEP_LENGTH,         /* f64 */
    CELER_FIELD_POST_STEP_ACTION,    /* u8  */
    CELER_FIELD_ALONG_STEP_ACTION,   /* u8  */
    CELER_FIELD_PARTICLE_ID,         /* u8  */
    CELER_FIELD_ENERGY,              /* f64 */
    CELER_FIELD_POS_X,               /* f64 */
    CELER_FIELD_POS_Y,
    CELER_FIELD_POS_Z,
    CELER_FIELD_DIR_X,               /* f64 */
    CELER_FIELD_DIR_Y,
    CELER_FIELD_DIR_Z,
    CELER_FIELD_VOLUME,              /* u32 */
    CELER_FIELD_SURFACE,             /* u32 */
    CELER_FIELD_INTERACTION_MFP,     /* f64 */
    CELER_FIELD_SEC_START,           /* u32 */
    CELER_FIELD_SEC_COUNT,           /* u32 */
    CELER_FIELD_ENERGY_DEPOSITION,   /* f64 */
    CELER_FIELD_RNG_XS0,             /* u32 x 6 words, SoA order */
    CELER_FIELD_RNG_XS1,
    CELER_FIELD_RNG_XS2,
    CELER_FIELD_RNG_XS3,
    CELER_FIELD_RNG_XS4,
    CELER_FIELD_RNG_WEYL,
    CELER_FIELD_VACANCIES,           /* u32 [num_slots] */
    CELER_FIELD_SECONDARY_COUNTS,    /* u32 [num_slots + 1] */
    CELER_FIELD_TRACK_COUNTERS,      /* u32 [max_events] */
    CELER_FIELD_INITIALIZERS,        /* CelerTrackInitializer [capacity] */
    CELER_FIELD_SECONDARIES,         /* CelerSecondary [stack capacity] */
    CELER_FIELD_CALO_EDEP,           /* f64 [num_detectors] */
    CELER_FIELD_MACRO_XS,            /* f64 */
    CELER_FIELD_DEDX_RANGE,          /* f64 */
    CELER_FIELD_TRACK_SLOTS,         /* u32 */
    CELER_FIELD_ACTION_OFFSETS,      /* u32 [65]: per-key sorted thread offsets */
    CELER_FIELD_NEXT_STEP,           /* f64 */
    CELER_FIELD_GEO_FAILED,          /* u8 */
    CELER_FIELD_SIZE_
} CelerStateField;

typedef struct CelerParams CelerParams;
typedef struct CelerState CelerState;
typedef struct CelerStepper CelerStepper;
typedef struct CelerTransporter CelerTransporter;
typedef struct CelerServer CelerServer;

/*---------------------------------------------------------------------------*
 * LIBRARY
 *---------------------------------------------------------------------------*/
const char* celer_amd_version(void);
/* 1 if a HIP device is usable, else 0 (initializes nothing on the GPU) */
int celer_amd_device_available(void);
/* Message of the last failure on this thread ("" if none) */
const char* celer_amd_last_error(void);

/*---------------------------------------------------------------------------*
 * PARAMS: CoreParams from a JSON problem description
 * (geometry in ORANGE .org.json form, physics, rng seed, track init)
 *---------------------------------------------------------------------------*/
CelerParams* celer_params_create(const char* json);
void celer_params_destroy(CelerParams* params);
uint32_t celer_params_num_actions(const CelerParams* params);
/* Copy an action label into buf (NUL-terminated, truncated to len) */
int celer_params_action_label(const CelerParams* params,
                              uint32_t action_id,
                              char* buf,
                              size_t len);
/* UINT32_MAX if no action has this label */
uint32_t celer_params_find_action(const CelerParams* params, const char* label);
/* Step-action order (StepActionOrder) or -1 for implicit actions */
int celer_params_action_order(const CelerParams* params, uint32_t action_id);
uint32_t celer_params_find_particle(const CelerParams* params, int pdg);
uint32_t celer_params_num_volumes(const CelerParams* params);
/*
 * Expand a compact physics description (elements, materials, cutoffs,
 * options) into a full "physics" input with cross-section, stopping-power,
 * range, Seltzer-Berger and Livermore tables. Host only (no device needed).
 * Replaces the reference's Geant4 import feeding PhysicsParams::build_xs
 * (celeritas/phys/PhysicsParams.cc:445-636, io/ImportPhysicsTable.hh).
 * With buf == NULL only *needed (bytes incl. NUL) is set.
 */
int celer_generate_physics(const char* spec_json, char* buf, size_t len,
                           size_t* needed);
/*
 * Read one element's G4EMLOW data as JSON (host only). dir is the folder
 * holding the files, or NULL to use $G4LEDATA/brem_SB (Seltzer-Berger) or
 * $G4LEDATA/livermore/phot_epics2014 (Livermore photoelectric). The result
 * is the element's "sb_table" ({x, y, values}) or "livermore_pe" object
 * ({xs_hi, xs_lo, thresh_lo, thresh_hi, shells}); physics specs name the
 * same folders under "g4emlow" to build tables from them.
 * Replaces io/SeltzerBergerReader.hh:28-52 (SeltzerBergerReader::operator())
 * and io/LivermorePEReader.hh:27-52 (LivermorePEReader::operator()).
 * With buf == NULL only *needed (bytes incl. NUL) is set.
 */
int celer_read_seltzer_berger(const char* dir, int z, char* buf, size_t len,
                              size_t* needed);
int celer_read_livermore_pe(const char* dir, int z, char* buf, size_t len,
                            size_t* needed);
/* Serialize host-side tables (physics, geometry) as JSON for checkers */
int celer_params_export_json(const CelerParams* params, char* buf, size_t len,
                             size_t* needed);
/* Attributes of every step kernel launched so far in this process, one
 * entry per kernel and launch shape, as a JSON array: registers, scratch,
 * static and dynamic LDS, resident workgroups per CU, waves per SIMD and
 * occupancy. Replaces the reference's KernelRegistry / KernelAttributes
 * (corecel/sys/KernelAttributes.hh:33-121), written to its output JSON. */
int celer_kernel_attributes_json(char* buf, size_t len, size_t* needed);
/* Time every following launch of the step kernels (enable = 1) with events
 * the runtime sets at the dispatch's own start and end, or stop (0). Not
 * for launches inside a stream capture. The reference times actions with
 * stream synchronizations (ActionSequence.cc:78-138); this times kernels. */
int celer_kernel_timing(int enable);
/* Device time and launch count per kernel and launch shape since the last
 * reset, as a JSON array (waits for the timed launches to end); a call with
 * no buffer (a size query) does not reset */
int celer_kernel_times_json(int reset, char* buf, size_t len, size_t* needed);

/*---------------------------------------------------------------------------*
 * STATE: one CoreState<device> with its own HIP stream
 *---------------------------------------------------------------------------*/
CelerState* celer_state_create(CelerParams* params,
                               uint32_t stream_id,
                               uint32_t num_slots);
void celer_state_destroy(CelerState* state);
uint32_t celer_state_size(const CelerState* state);
/* Buffer primaries for the next extend-from-primaries action */
int celer_state_insert_primaries(CelerState* state,
                                 const CelerPrimary* primaries,
                                 uint32_t count);
/* Run one step action by ID (like StepActionInterface::step) */
int celer_state_execute_action(CelerState* state, uint32_t action_id);
/* Synchronize and copy the device counters */
int celer_state_counters(CelerState* state, CelerCounters* out);
/* Copy a field: direction 0 = device->host, 1 = host->device */
int celer_state_copy_field(CelerState* state,
                           CelerStateField field,
                           int direction,
                           void* host,
                           size_t bytes);
/* Overwrite the device counters (test harness: mock interactions) */
int celer_state_set_counters(CelerState* state, const CelerCounters* in);
/* Per-event reseed: SplitMix64(seed) + subsequence event*N + slot */
int celer_state_reseed(CelerState* state, uint64_t event_id);
/* Seed each slot with its own subsequence and offset */
int celer_state_rng_initialize(CelerState* state,
                               const uint64_t* subsequence,
                               const uint64_t* offset);
/* Draw per slot: kind 0 = u32, 1 = f64 canonical, 2 = f32 canonical */
int celer_state_rng_draw(CelerState* state,
                         uint32_t draws_per_slot,
                         int kind,
                         void* host_out);
int celer_state_reset(CelerState* state);
int celer_state_synchronize(CelerState* state);

/*---------------------------------------------------------------------------*
 * STEPPER: Stepper<device>
 *---------------------------------------------------------------------------*/
CelerStepper* celer_stepper_create(CelerParams* params,
                                   uint32_t stream_id,
                                   uint32_t num_slots,
                                   int action_times);
void celer_stepper_destroy(CelerStepper* stepper);
/* Insert primaries (may be NULL/0) then take one step */
int celer_stepper_step(CelerStepper* stepper,
                       const CelerPrimary* primaries,
                       uint32_t count,
                       CelerStepperResult* result);
/* Take count steps without new primaries, reading the counters back once
 * at the end (kernel launch sizes and ranges are device-resident, so the
 * steps need no host input); result is the last step's. (The reference's
 * Stepper reads its counters after every step, global/Stepper.cc:125-139.) */
int celer_stepper_run_steps(CelerStepper* stepper, uint32_t count,
                            CelerStepperResult* result);
int celer_stepper_warm_up(CelerStepper* stepper);
int celer_stepper_kill_active(CelerStepper* stepper);
int celer_stepper_reseed(CelerStepper* stepper, uint64_t event_id);
/* The stepper's state (owned by the stepper) */
CelerState* celer_stepper_state(CelerStepper* stepper);
/* Labels of step actions in execution order, newline-separated */
int celer_stepper_action_sequence(const CelerStepper* stepper,
                                  char* buf,
                                  size_t len);
/*
 * Accumulated device time [s] and number of timed steps per step action
 * (execution order), for a stepper created with action_times != 0. Device
 * time is measured with events bracketing each action's launches on the
 * state stream. Counterpart of ActionSequence::accum_time
 * (celeritas/global/ActionSequence.cc:78-138) and Stepper::actions().
 * n must equal the number of step actions; reset != 0 zeroes the totals
 * after reading.
 */
int celer_stepper_action_times(CelerStepper* stepper,
                               double* seconds,
                               uint64_t* counts,
                               uint32_t n,
                               int reset);
/*---------------------------------------------------------------------------*
 * OFFLOAD FRONT END: accel/LocalTransporter.hh:52-109 without Geant4
 *---------------------------------------------------------------------------*/
/*
 * SetupOptions (accel/SetupOptions.hh): max_num_tracks slots, auto_flush
 * (0: max_num_tracks), max_steps iterations per flush; with nonfatal_flush
 * an exceeded max_steps kills the active tracks instead of failing
 * (LocalTransporter.cc:60-80, 247-300).
 */
CelerTransporter* celer_transporter_create(CelerParams* params,
                                           uint32_t max_num_tracks,
                                           uint32_t auto_flush,
                                           uint32_t max_steps,
                                           uint32_t stream_id,
                                           int nonfatal_flush,
                                           int action_times);
void celer_transporter_destroy(CelerTransporter* transporter);
/* LocalTransporter::InitializeEvent: reseed track RNG streams */
int celer_transporter_initialize_event(CelerTransporter* transporter, int event_id);
/* LocalTransporter::Push for each track (flushes at auto_flush) */
int celer_transporter_push(CelerTransporter* transporter,
                           const CelerPrimary* tracks,
                           uint32_t count);
/* LocalTransporter::Flush: transport buffered tracks to completion */
int celer_transporter_flush(CelerTransporter* transporter);
/* LocalTransporter::Finalize: fails if tracks remain buffered */
int celer_transporter_finalize(CelerTransporter* transporter);
/* LocalTransporter::GetBufferSize */
uint32_t celer_transporter_buffer_size(const CelerTransporter* transporter);
/* The transporter's state (tallies, fields), owned by the transporter */
CelerState* celer_transporter_state(CelerTransporter* transporter);

/*
 * Copy a device buffer of `bytes` (rounded down to 16 B) `reps` times with
 * a one-element-per-thread kernel; *ms receives the mean kernel time. The
 * reference point for converting hardware memory counters into bytes (no
 * counterpart in the reference). Returns 0 or a negative hipError_t value.
 */
int celer_calibration_copy(size_t bytes, uint32_t reps, float* ms);
/*
 * Per-kernel work counters (tracks processed by pre-step, charged and
 * neutral along-step, boundary crossings, ...; up to 8 values). enable: 1
 * on (zeroed), 0 off, -1 unchanged; out (may be NULL) receives n values,
 * reset != 0 zeroes them after reading. Counterpart of the reference's
 * ActionDiagnostic/StepDiagnostic counters (celeritas/user/
 * ActionDiagnostic.hh): one atomic per wavefront when on, none when off.
 */
int celer_state_diagnostics(CelerState* state, int enable, uint64_t* out,
                            uint32_t n, int reset);

/*---------------------------------------------------------------------------*
 * LOW-LEVEL LAUNCH ABI: enqueue kernels on a caller stream with the POD
 * device views obtained from celer_*_device_ref. hipStream_t is passed as
 * void*. Returns a hipError_t value.
 *---------------------------------------------------------------------------*/
const void* celer_params_device_ref(const CelerParams* params);
const void* celer_state_device_ref(const CelerState* state);
void* celer_state_stream(const CelerState* state);

/* track/ExtendFromPrimariesAction.cu:23-38 */
int celer_hip_extend_from_primaries(const void* state_ref,
                                    const CelerPrimary* device_primaries,
                                    uint32_t count,
                                    void* stream);
/* track/InitializeTracksAction.cu:23-36 */
int celer_hip_initialize_tracks(const void* params_ref,
                                const void* state_ref,
                                void* stream);
/* track/ExtendFromSecondariesAction.cu:45-69 + TrackInitAlgorithms.cu */
int celer_hip_extend_from_secondaries(const void* params_ref,
                                      const void* state_ref,
                                      void* stream);
/* random/RngReseed.cu:59-72 */
int celer_hip_reseed(const void* params_ref,
                     const void* state_ref,
                     uint64_t event_id,
                     void* stream);

/*
 * Per-action kernels (the executors each step action's step(params,
 * CoreState<device>&) launches: ActionLauncher.device.hh:102-135 with
 * ConditionalTrackExecutor). A launch visits the threads of a thread map:
 * with range == NULL every slot (through slots, or in place when slots is
 * NULL); with range pointing at &action_offsets[key] (device memory, as
 * returned by celer_state_action_offsets + key) only that sorted key's
 * threads, a count the host never reads. Kernels skip slots whose status
 * or action does not match, so any map is correct; a sorted one packs the
 * work. The mixed-precision build of a kernel is used when the params ask
 * for it. Returns a hipError_t value.
 */
typedef struct CelerThreadSlots
{
    const uint32_t* range; /* {begin, end} of a sorted key, or NULL */
    const uint32_t* slots; /* thread -> slot (track_slots), or NULL */
} CelerThreadSlots;

/* Device pointers of a state's thread -> slot map and per-key offsets
 * (max 64 keys + 1), written by celer_hip_sort_tracks */
const uint32_t* celer_state_track_slots(const CelerState* state);
const uint32_t* celer_state_action_offsets(const CelerState* state);

/* phys/detail/PreStepExecutor.hh:46-115 (PreStepAction) */
int celer_hip_pre_step_launch(const void* params_ref,
                              const void* state_ref,
                              const CelerThreadSlots* ts,
                              void* stream);
/* global/alongstep/AlongStepNeutralAction.cc / AlongStepGeneralLinearAction.cc
 * / AlongStepUniformMscAction.cc: tracks whose along-step action is
 * action_id; charged != 0 runs the charged (MSC, energy loss, field) body */
int celer_hip_along_step_launch(const void* params_ref,
                                const void* state_ref,
                                uint32_t action_id,
                                int charged,
                                const CelerThreadSlots* ts,
                                void* stream);
/* phys/detail/DiscreteSelectExecutor.hh:38-63 (DiscreteSelectAction) */
int celer_hip_discrete_select_launch(const void* params_ref,
                                     const void* state_ref,
                                     const CelerThreadSlots* ts,
                                     void* stream);
/* em/executor/<Model>Executor.hh through phys/InteractionApplier.hh:106-171:
 * tracks whose post-step action is action_id, sampled by model class
 * 0 klein_nishina, 1 livermore_pe, 2 bethe_heitler, 3 moller_bhabha,
 * 4 seltzer_berger, 5 relativistic_brem, 6 eplusgg, 7 combined_brem */
int celer_hip_interaction_launch(const void* params_ref,
                                 const void* state_ref,
                                 uint32_t action_id,
                                 uint32_t model_class,
                                 const CelerThreadSlots* ts,
                                 void* stream);
/* geo/detail/BoundaryExecutor.hh:42-84 (BoundaryAction) */
int celer_hip_boundary_launch(const void* params_ref,
                              const void* state_ref,
                              const CelerThreadSlots* ts,
                              void* stream);
/* phys/detail/TrackingCutExecutor.hh:49-84 (TrackingCutAction) */
int celer_hip_tracking_cut_launch(const void* params_ref,
                                  const void* state_ref,
                                  const CelerThreadSlots* ts,
                                  void* stream);
/* user/detail/SimpleCaloExecutor.hh:48-66: per-detector deposition of
 * num_detectors bins (volume -> detector map of the params) */
int celer_hip_calo_tally_launch(const void* params_ref,
                                const void* state_ref,
                                uint32_t num_detectors,
                                const CelerThreadSlots* ts,
                                void* stream);
/* track/SortTracksAction.cc:102-131 with detail/TrackSortUtils.cu:145-206:
 * reindex track_slots stably by a per-slot key (0 status, 1 particle type,
 * 2 along-step action, 3 post-step action; inactive slots take key
 * num_keys - 1) and write the per-key offsets to action_offsets[0..num_keys]
 * on the device (1 <= num_keys <= 64) */
int celer_hip_sort_tracks(const void* state_ref,
                          uint32_t key_kind,
                          uint32_t num_keys,
                          void* stream);
/* The along-step sort as the sort-tracks-along-step action runs it, before
 * the pre-step: key = the along-step action the pre-step will choose for
 * the slot (its particle's, by charge; inactive and errored slots take
 * num_keys - 1), which orders the slots as the reference's sort by the
 * along-step action after the pre-step does (SortTracksAction.cc:102-131)
 * and lets the pre-step run over the same permutation */
int celer_hip_sort_tracks_along_before_pre(const void* params_ref,
                                           const void* state_ref,
                                           uint32_t num_keys,
                                           void* stream);
/* The same stable multi-split over a caller's keys (device memory):
 * indices 0..n-1 ordered by keys[i] (values at or above num_keys - 1 take
 * key num_keys - 1), equal keys in increasing index order, into sorted[n],
 * and offsets[0..num_keys] (offsets[num_keys] = n), all on the device.
 * Allocates its histogram scratch (4 * 64 * (ceil(n / 1024) + 1) bytes) on the
 * stream and frees it there. The track-slot form is celer_hip_sort_tracks
 * (detail/TrackSortUtils.cu:145-206 sorts track_slots by per-slot action
 * keys with thrust::sort_by_key). */
int celer_hip_partition_keys(const uint8_t* keys,
                             uint32_t n,
                             uint32_t num_keys,
                             uint32_t* sorted,
                             uint32_t* offsets,
                             void* stream);
/* Allocation-free celer_hip_partition_keys. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Its histogram scratch is a static device array of 1 MiB + 256 B per GPU, held
 * until the process ends, so n is limited to CELER_PARTITION_NOALLOC_MAX
 * (larger n fails with hipErrorInvalidValue and a message from
 * celer_amd_last_error()). The array keeps the last call's contents, every
 * call rewrites what it reads, and calls that overlap on one GPU race.
 * Results are identical to celer_hip_partition_keys. */
#define CELER_PARTITION_NOALLOC_MAX (1u << 22)
int celer_hip_partition_keys_noalloc(const uint8_t* keys,
                                     uint32_t n,
                                     uint32_t num_keys,
                                     uint32_t* sorted,
                                     uint32_t* offsets,
                                     void* stream);

/*---------------------------------------------------------------------------*
 * SERVING: continuous batching of events that arrive over time. No
 * counterpart in the reference, whose offload front end transports one
 * event per Flush (accel/LocalTransporter.cc:247-300); this one admits
 * events in arrival order while the initializer queue (max_queued, 0: the
 * slot count) has room, so events of uneven size overlap on the device,
 * and detects each event's completion from device-resident per-event
 * live-track counts. The server owns a stepper and its state.
 *---------------------------------------------------------------------------*/
/* Workgroup size of the step kernels (64, 128 or 256; default 256), the
 * autotuner's launch parameter. Set it before capturing a graph. */
int celer_state_set_block_size(CelerState* state, uint32_t block_size);

/* Capture the step sequence (steps without new primaries) as a hipGraph;
 * later celer_stepper_step calls without primaries replay it with one
 * launch. Not with per-action timing. (The reference has no graph path; the
 * capture point is the Stepper's action loop, global/Stepper.cc:125-139.) */
int celer_stepper_capture_step(CelerStepper* stepper);

CelerServer* celer_server_create(CelerParams* params,
                                 uint32_t num_track_slots,
                                 uint32_t stream_id);
void celer_server_destroy(CelerServer* server);
/* The server's state (for field access); owned by the server */
CelerState* celer_server_state(CelerServer* server);
/* Step iterations the server launches back to back per counter readback
 * while no arrived event waits for admission (default 4). No counterpart in
 * the reference, whose loop reads its counters after every step
 * (global/Stepper.cc:125-139). */
int celer_server_set_steps_per_sync(CelerServer* server, uint32_t steps);
/* Serve num_events events: event e is primaries[event_offsets[e] ..
 * event_offsets[e+1]) (their event_id is replaced by e) arriving
 * arrivals[e] seconds after the call starts (non-decreasing). Writes per
 * event the admission and completion times [s] and the deposited energy
 * [MeV], and totals = {wall time [s], step iterations, track-steps}.
 * max_steps > 0 aborts with an error after that many steps. */
int celer_server_run(CelerServer* server,
                     const CelerPrimary* primaries,
                     const uint32_t* event_offsets,
                     const double* arrivals,
                     uint32_t num_events,
                     uint32_t max_queued,
                     uint32_t max_steps,
                     double* admit_times,
                     double* done_times,
                     double* edep,
                     double* totals);

/*---------------------------------------------------------------------------*
 * VERIFICATION: run one interactor model (ModelClass) on the device with
 * replayed uniform streams (stride doubles per sample). Writes 17 doubles
 * per sample: action, energy, dir[3], edep, nsec, then (particle, energy,
 * dir[3]) for up to two secondaries. Counterpart of the reference's device
 * interactor tests (test/celeritas/em/ interactor tests and their .test.cu wrappers).
 *---------------------------------------------------------------------------*/
int celer_verify_interactor(const void* params_ref,
                            const void* state_ref,
                            uint32_t model,
                            uint32_t particle,
                            const double* energies,
                            const double* dirs,
                            uint32_t material,
                            uint32_t element,
                            uint32_t n,
                            const double* uniforms,
                            uint32_t stride,
                            double* out,
                            void* stream);

/*---------------------------------------------------------------------------*
 * VERIFICATION: uniform-field propagation (field/FieldPropagator.hh:150-331
 * with FieldDriver and DormandPrinceStepper). Each of ntracks tracks (11
 * doubles: pos[3] cm, dir[3], momentum MeV/c, charge e, field[3] T) runs
 * nsteps requests through one propagator, as the reference's unit tests
 * do (test/celeritas/field/FieldPropagator.test.cc); a request of 0 crosses
 * the boundary and -1 starts a new propagator. Writes 10 doubles per
 * request: distance, boundary, looping, pos[3], dir[3], volume. Driver
 * options are the params' (max_substeps > 0 overrides that option).
 *---------------------------------------------------------------------------*/
int celer_verify_field_propagate(const void* params_ref,
                                 int max_substeps,
                                 const double* tracks,
                                 uint32_t ntracks,
                                 const double* steps,
                                 uint32_t nsteps,
                                 double* out,
                                 void* stream);

/*---------------------------------------------------------------------------*
 * VERIFICATION: point location as in track initialization
 * (orange/univ/SimpleUnitTracker.hh:161-190, candidates from the BIH of
 * orange/detail/BIHTraverser.hh). Each of n points is 6 doubles (pos [cm],
 * dir); writes its volume ID, or 0xffffffff where initialization fails
 * (the point lies on a face, or in no volume and there is no background).
 *---------------------------------------------------------------------------*/
int celer_verify_geo_locate(const void* params_ref,
                            const double* points,
                            uint32_t n,
                            uint32_t* out,
                            void* stream);
/*
 * VERIFICATION: ray tracing as the linear propagator drives the geometry
 * (OrangeTrackView::find_next_step / move_to_boundary / cross_boundary /
 * find_safety, orange/OrangeTrackView.hh:555-775). Each of n rays is 6
 * doubles (pos [cm], dir). Per ray, up to max_segments segments are
 * written at [ray * max_segments + k]: the volume, the distance to its
 * boundary and the safety distance at the segment's midpoint; counts[ray]
 * is the number of segments (tracing stops on leaving the world).
 */
int celer_verify_geo_trace(const void* params_ref,
                           const double* rays,
                           uint32_t n,
                           uint32_t max_segments,
                           uint32_t* counts,
                           uint32_t* volumes,
                           double* distances,
                           double* safeties,
                           void* stream);

/*---------------------------------------------------------------------------*
 * USER ACTIONS: diagnostics and detector step points. Each is created on
 * the params before the stepper (or state) whose steps it observes; the
 * handle shares ownership of its actions with the params' registry.
 *---------------------------------------------------------------------------*/
typedef struct CelerActionDiagnostic CelerActionDiagnostic;
typedef struct CelerStepDiagnostic CelerStepDiagnostic;
typedef struct CelerStepCollector CelerStepCollector;
typedef struct CelerDetectorSteps CelerDetectorSteps;

/* Replaces ActionDiagnostic::make_and_insert(core)
 * (src/celeritas/user/ActionDiagnostic.cc:40-52): counts of post-step
 * actions by particle over valid tracks, at the post-step order. */
int celer_action_diagnostic_create(CelerParams* params, CelerActionDiagnostic** out);
void celer_action_diagnostic_destroy(CelerActionDiagnostic* diag);
/* Shape of the counts: [num_particles][num_actions] (0 before a stepper or
 * state has been set up with it) */
int celer_action_diagnostic_shape(const CelerActionDiagnostic* diag,
                                  uint32_t* num_particles,
                                  uint32_t* num_actions);
/* Replaces ActionDiagnostic::calc_actions (ActionDiagnostic.cc:171-195):
 * counts summed over streams, written row-major into out[size] */
int celer_action_diagnostic_counts(const CelerActionDiagnostic* diag,
                                   uint64_t* out,
                                   size_t size);
int celer_action_diagnostic_clear(CelerActionDiagnostic* diag);

/* Replaces StepDiagnostic::make_and_insert(core, max_step_bin)
 * (src/celeritas/user/StepDiagnostic.cc:30-42): steps per killed track by
 * particle, bins 0..max_step_bin plus overflow. */
int celer_step_diagnostic_create(CelerParams* params,
                                 uint32_t max_step_bin,
                                 CelerStepDiagnostic** out);
void celer_step_diagnostic_destroy(CelerStepDiagnostic* diag);
int celer_step_diagnostic_shape(const CelerStepDiagnostic* diag,
                                uint32_t* num_particles,
                                uint32_t* num_bins);
/* Replaces StepDiagnostic::calc_steps (StepDiagnostic.cc:115-135) */
int celer_step_diagnostic_counts(const CelerStepDiagnostic* diag, uint64_t* out, size_t size);
int celer_step_diagnostic_clear(CelerStepDiagnostic* diag);

/* Track properties gathered at a step point (StepPointSelection,
 * src/celeritas/user/StepData.hh:26-60); nonzero selects */
typedef struct CelerStepPointSelection
{
    uint8_t time;
    uint8_t pos;
    uint8_t dir;
    uint8_t volume_id;
    uint8_t energy;
} CelerStepPointSelection;

/* Track properties gathered every step (StepSelection, StepData.hh:64-120);
 * points[0] is the pre-step point, points[1] the post-step point */
typedef struct CelerStepSelection
{
    CelerStepPointSelection points[2];
    uint8_t event_id;
    uint8_t parent_id;
    uint8_t track_step_count;
    uint8_t action_id;
    uint8_t step_length;
    uint8_t particle;
    uint8_t energy_deposition;
} CelerStepSelection;

/* Replaces StepCollector::make_and_insert(core, interfaces)
 * (src/celeritas/user/StepCollector.cc:33-100) for detector hits: inserts
 * the "step-gather-pre" and "step-gather-post" actions. volume_to_detector
 * (num_volumes entries, 0xffffffff: not a detector) selects the slots whose
 * steps are kept; null keeps every active track with detector 0. With
 * nonzero_energy_deposition, post-step points without deposition are
 * skipped (StepGatherExecutor.hh:96-105). Hits accumulate on the device,
 * hit_capacity per stream, until read. */
int celer_step_collector_create(CelerParams* params,
                                const CelerStepSelection* selection,
                                const uint32_t* volume_to_detector,
                                uint32_t num_volumes,
                                int nonzero_energy_deposition,
                                uint32_t hit_capacity,
                                CelerStepCollector** out);
void celer_step_collector_destroy(CelerStepCollector* coll);
/* Replaces copy_steps<device>(DetectorStepOutput*, state)
 * (src/celeritas/user/DetectorSteps.cu:140-200): copies every hit gathered
 * on the stream since the last read, in step order and within a step in
 * slot order, then empties the device buffer. */
int celer_step_collector_read(CelerStepCollector* coll,
                              uint32_t stream_id,
                              CelerDetectorSteps** out);
void celer_detector_steps_destroy(CelerDetectorSteps* steps);
/* Number of hits; steps dropped because the buffer was full, and their hits */
uint32_t celer_detector_steps_size(const CelerDetectorSteps* steps);
int celer_detector_steps_overflow(const CelerDetectorSteps* steps,
                                  uint32_t* overflow_steps,
                                  uint32_t* dropped_hits);
/* Copy a column (DetectorStepOutput member): "detector", "track_id",
 * "event_id", "parent_id", "track_step_count", "action_id", "particle"
 * (uint32), "step_length", "energy_deposition" (double), and per point
 * "pre." or "post." + "time", "energy" (double), "pos", "dir" (3 doubles
 * per hit), "volume_id" (uint32; 0xffffffff outside the world). Returns an
 * error for an unselected column or a too-small buffer. */
int celer_detector_steps_column(const CelerDetectorSteps* steps,
                                const char* name,
                                void* dst,
                                size_t bytes);

#ifdef __cplusplus
}
#endif

#endif /* CELERITAS_AMD_H */
