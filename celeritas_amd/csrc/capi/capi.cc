//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file capi/capi.cc
//! C ABI over the host C++ classes (include/celeritas_amd.h).
//---------------------------------------------------------------------------//
#include "celeritas_amd.h"

#include <cstring>
#include <string>

#include "accel/LocalTransporter.hh"
#include "accel/ServingLoop.hh"
#include "corecel/KernelRegistry.hh"
#include "global/CoreActions.hh"
#include "global/CoreParams.hh"
#include "global/CoreState.hh"
#include "phys/TableGen.hh"
#include "global/Stepper.hh"
#include "io/G4EmlowReaders.hh"
#include "track/TrackActions.hh"
#include "track/SortKernels.hh"
#include "track/TrackKernels.hh"
#include "user/UserActions.hh"

using namespace celeritas_amd;

static_assert(sizeof(CelerPrimary) == sizeof(Primary), "Primary layout");
static_assert(sizeof(CelerCounters) == sizeof(CoreCounters), "Counters layout");
static_assert(sizeof(CelerTrackInitializer) == sizeof(TrackInitializer),
              "TrackInitializer layout");
static_assert(sizeof(CelerSecondary) == sizeof(Secondary), "Secondary layout");
static_assert(sizeof(CelerStepSelection) == sizeof(StepSelection), "StepSelection layout");

struct CelerParams
{
    std::shared_ptr<CoreParams> params;
};

struct CelerState
{
    std::shared_ptr<CoreParams> params;
    std::shared_ptr<CoreState> state;
};

struct CelerStepper
{
    std::shared_ptr<CoreParams> params;
    std::shared_ptr<Stepper> stepper;
    CelerState state_handle;
};

struct CelerTransporter
{
    std::shared_ptr<CoreParams> params;
    std::unique_ptr<LocalTransporter> transporter;
    CelerState state_handle;
};

struct CelerServer
{
    std::shared_ptr<CoreParams> params;
    std::unique_ptr<ServingLoop> loop;
    CelerState state_handle;
    uint32_t steps_per_sync{4};
};

namespace
{
thread_local std::string g_last_error;

template<class F>
int guarded(F&& f)
{
    try
    {
        g_last_error.clear();
        f();
        return 0;
    }
    catch (std::exception const& e)
    {
        g_last_error = e.what();
    }
    catch (...)
    {
        g_last_error = "unknown C++ exception";
    }
    return 1;
}

//! Copy text into a caller buffer, reporting the size it needs (with NUL)
void copy_text(std::string const& s, char* buf, std::size_t len, std::size_t* needed)
{
    if (needed)
        *needed = s.size() + 1;
    CELER_VALIDATE(buf == nullptr || len > s.size(),
                   << "output buffer too small (" << s.size() + 1 << " bytes needed)");
    if (buf)
        std::memcpy(buf, s.c_str(), s.size() + 1);
}

CoreParamsRef const& params_of(void const* ref)
{
    return *static_cast<CoreParamsRef const*>(ref);
}

CoreStateRef const& state_of(void const* ref)
{
    return *static_cast<CoreStateRef const*>(ref);
}

ThreadSlots slots_of(CelerThreadSlots const* ts)
{
    ThreadSlots r;
    if (ts)
    {
        r.range = ts->range;
        r.slots = ts->slots;
    }
    return r;
}

template<class T, class F>
T* guarded_new(F&& f)
{
    T* result = nullptr;
    guarded([&] { result = f(); });
    return result;
}

//! Device pointer and byte size of a state field
std::pair<void*, std::size_t> field_ptr(CoreStateRef const& s,
                                        CoreParams const& p,
                                        CelerStateField f,
                                        size_type num_detectors)
{
    std::size_t const n = s.size;
    switch (f)
    {
        case CELER_FIELD_TRACK_ID: return {s.sim.track_id, 4 * n};
        case CELER_FIELD_PARENT_ID: return {s.sim.parent_id, 4 * n};
        case CELER_FIELD_EVENT_ID: return {s.sim.event_id, 4 * n};
        case CELER_FIELD_NUM_STEPS: return {s.sim.num_steps, 4 * n};
        case CELER_FIELD_TIME: return {s.sim.time, 8 * n};
        case CELER_FIELD_STATUS: return {s.sim.status, n};
        case CELER_FIELD_STEP_LENGTH: return {s.sim.step_length, 8 * n};
        case CELER_FIELD_POST_STEP_ACTION: return {s.sim.post_step_action, n};
        case CELER_FIELD_ALONG_STEP_ACTION: return {s.sim.along_step_action, n};
        case CELER_FIELD_PARTICLE_ID: return {s.particle.particle_id, n};
        case CELER_FIELD_ENERGY: return {s.particle.energy, 8 * n};
        case CELER_FIELD_POS_X: return {s.geo.pos[0], 8 * n};
        case CELER_FIELD_POS_Y: return {s.geo.pos[1], 8 * n};
        case CELER_FIELD_POS_Z: return {s.geo.pos[2], 8 * n};
        case CELER_FIELD_DIR_X: return {s.geo.dir[0], 8 * n};
        case CELER_FIELD_DIR_Y: return {s.geo.dir[1], 8 * n};
        case CELER_FIELD_DIR_Z: return {s.geo.dir[2], 8 * n};
        case CELER_FIELD_VOLUME: return {s.geo.volume, 4 * n};
        case CELER_FIELD_SURFACE: return {s.geo.surface, 4 * n};
        case CELER_FIELD_INTERACTION_MFP: return {s.phys.interaction_mfp, 8 * n};
        case CELER_FIELD_SEC_START: return {s.phys.sec_start, 4 * n};
        case CELER_FIELD_SEC_COUNT: return {s.phys.sec_count, 4 * n};
        case CELER_FIELD_ENERGY_DEPOSITION:
            return {s.phys.energy_deposition, 8 * n};
        case CELER_FIELD_RNG_XS0:
        case CELER_FIELD_RNG_XS1:
        case CELER_FIELD_RNG_XS2:
        case CELER_FIELD_RNG_XS3:
        case CELER_FIELD_RNG_XS4:
            return {s.rng.xs[f - CELER_FIELD_RNG_XS0], 4 * n};
        case CELER_FIELD_RNG_WEYL: return {s.rng.weyl, 4 * n};
        case CELER_FIELD_VACANCIES: return {s.init.vacancies, 4 * n};
        case CELER_FIELD_SECONDARY_COUNTS:
            return {s.init.secondary_counts, 4 * (n + 1)};
        case CELER_FIELD_TRACK_COUNTERS:
            return {s.init.track_counters, 4 * std::size_t(s.init.max_events)};
        case CELER_FIELD_INITIALIZERS:
            return {s.init.initializers,
                    sizeof(TrackInitializer) * std::size_t(s.init.capacity)};
        case CELER_FIELD_SECONDARIES:
            return {s.secondaries.storage,
                    sizeof(Secondary) * std::size_t(s.secondaries.capacity)};
        case CELER_FIELD_CALO_EDEP:
            return {s.calo_edep, 8 * std::size_t(num_detectors)};
        case CELER_FIELD_MACRO_XS: return {s.phys.macro_xs, 8 * n};
        case CELER_FIELD_DEDX_RANGE: return {s.phys.dedx_range, 8 * n};
        case CELER_FIELD_TRACK_SLOTS: return {s.track_slots, 4 * n};
        case CELER_FIELD_ACTION_OFFSETS:
            return {s.action_offsets,
                    4 * std::size_t(max_sort_keys + 1)};
        case CELER_FIELD_NEXT_STEP: return {s.geo.next_step, 8 * n};
        case CELER_FIELD_GEO_FAILED: return {s.geo.failed, n};
        default: break;
    }
    throw RuntimeError("invalid state field " + std::to_string(int(f)));
}

std::shared_ptr<ExtendFromPrimariesAction const> primaries_action(CoreParams const& p)
{
    auto id = p.action_reg()->find_action(ExtendFromPrimariesAction::label_str());
    CELER_VALIDATE(id != invalid_id, << "no primary action");
    return std::dynamic_pointer_cast<ExtendFromPrimariesAction const>(
        p.action_reg()->action(id));
}
}  // namespace

extern "C" {

//---------------------------------------------------------------------------//
const char* celer_amd_version(void)
{
    return "0.1.0";
}

int celer_amd_device_available(void)
{
    return device_available() ? 1 : 0;
}

const char* celer_amd_last_error(void)
{
    return g_last_error.c_str();
}

//---------------------------------------------------------------------------//
CelerParams* celer_params_create(const char* json)
{
    return guarded_new<CelerParams>([&] {
        require_device();
        return new CelerParams{CoreParams::from_json(json)};
    });
}

void celer_params_destroy(CelerParams* params)
{
    delete params;
}

uint32_t celer_params_num_actions(const CelerParams* params)
{
    return params->params->action_reg()->num_actions();
}

int celer_params_action_label(const CelerParams* params,
                              uint32_t action_id,
                              char* buf,
                              size_t len)
{
    return guarded([&] {
        std::string const& s = params->params->action_reg()->id_to_label(action_id);
        std::snprintf(buf, len, "%s", s.c_str());
    });
}

uint32_t celer_params_find_action(const CelerParams* params, const char* label)
{
    return params->params->action_reg()->find_action(label);
}

int celer_params_action_order(const CelerParams* params, uint32_t action_id)
{
    auto const& a = params->params->action_reg()->action(action_id);
    if (auto sa = std::dynamic_pointer_cast<StepActionInterface const>(a))
        return static_cast<int>(sa->order());
    return -1;
}

uint32_t celer_params_find_particle(const CelerParams* params, int pdg)
{
    return params->params->physics().find_particle(pdg);
}

uint32_t celer_params_num_volumes(const CelerParams* params)
{
    return params->params->geometry().num_volumes();
}

const void* celer_params_device_ref(const CelerParams* params)
{
    return &params->params->device_ref();
}

//---------------------------------------------------------------------------//
CelerState* celer_state_create(CelerParams* params, uint32_t stream_id, uint32_t num_slots)
{
    return guarded_new<CelerState>([&] {
        auto st = std::make_shared<CoreState>(*params->params, stream_id, num_slots);
        return new CelerState{params->params, std::move(st)};
    });
}

void celer_state_destroy(CelerState* state)
{
    delete state;
}

uint32_t celer_state_size(const CelerState* state)
{
    return state->state->size();
}

const void* celer_state_device_ref(const CelerState* state)
{
    return &state->state->ref();
}

void* celer_state_stream(const CelerState* state)
{
    return state->state->stream();
}

int celer_state_insert_primaries(CelerState* state,
                                 const CelerPrimary* primaries,
                                 uint32_t count)
{
    return guarded([&] {
        primaries_action(*state->params)
            ->insert(*state->params, *state->state,
                     reinterpret_cast<Primary const*>(primaries), count);
    });
}

int celer_state_execute_action(CelerState* state, uint32_t action_id)
{
    return guarded([&] {
        auto const& a = state->params->action_reg()->action(action_id);
        auto sa = std::dynamic_pointer_cast<StepActionInterface const>(a);
        CELER_VALIDATE(sa, << "action " << action_id << " is not a step action");
        sa->step(*state->params, *state->state);
    });
}

int celer_state_counters(CelerState* state, CelerCounters* out)
{
    return guarded([&] {
        CoreCounters const& c = state->state->sync_counters();
        std::memcpy(out, &c, sizeof(CoreCounters));
    });
}

int celer_state_set_counters(CelerState* state, const CelerCounters* in)
{
    return guarded([&] {
        CELER_HIP_CALL(hipMemcpyAsync(state->state->ref().counters, in,
                                      sizeof(CoreCounters), hipMemcpyHostToDevice,
                                      state->state->stream()));
        state->state->sync_counters();
    });
}

int celer_state_copy_field(CelerState* state,
                           CelerStateField field,
                           int direction,
                           void* host,
                           size_t bytes)
{
    return guarded([&] {
        auto [ptr, size] = field_ptr(state->state->ref(), *state->params, field,
                                     state->state->num_detectors());
        CELER_VALIDATE(bytes <= size, << "requested " << bytes
                                      << " bytes from a field of " << size);
        hipStream_t s = state->state->stream();
        if (direction == 0)
            CELER_HIP_CALL(hipMemcpyAsync(host, ptr, bytes, hipMemcpyDeviceToHost, s));
        else
            CELER_HIP_CALL(hipMemcpyAsync(ptr, host, bytes, hipMemcpyHostToDevice, s));
        CELER_HIP_CALL(hipStreamSynchronize(s));
    });
}

int celer_state_reseed(CelerState* state, uint64_t event_id)
{
    return guarded([&] {
        auto const& p = *state->params;
        CELER_HIP_CALL(launch_reseed_rng(state->state->ref().rng,
                                         p.device_ref().rng_jump, p.rng().seed(),
                                         event_id, state->state->stream()));
        CELER_HIP_CALL(hipStreamSynchronize(state->state->stream()));
    });
}

int celer_state_rng_initialize(CelerState* state,
                               const uint64_t* subsequence,
                               const uint64_t* offset)
{
    return guarded([&] {
        auto const& p = *state->params;
        std::size_t n = state->state->size();
        DeviceVector<ull_int> sub(n), off(n);
        sub.copy_from_host(reinterpret_cast<ull_int const*>(subsequence), n);
        off.copy_from_host(reinterpret_cast<ull_int const*>(offset), n);
        CELER_HIP_CALL(launch_rng_initialize(state->state->ref().rng,
                                             p.device_ref().rng_jump, p.rng().seed(),
                                             sub.data(), off.data(),
                                             state->state->stream()));
        CELER_HIP_CALL(hipStreamSynchronize(state->state->stream()));
    });
}

int celer_state_rng_draw(CelerState* state, uint32_t draws, int kind, void* host_out)
{
    return guarded([&] {
        CELER_VALIDATE(kind >= 0 && kind <= 2, << "bad draw kind");
        std::size_t n = std::size_t(state->state->size()) * draws;
        std::size_t bytes = n * (kind == 1 ? 8 : 4);
        DeviceVector<char> buf(bytes);
        CELER_HIP_CALL(launch_rng_draw(state->state->ref().rng, draws, kind,
                                       buf.data(), state->state->stream()));
        CELER_HIP_CALL(hipMemcpyAsync(host_out, buf.data(), bytes,
                                      hipMemcpyDeviceToHost, state->state->stream()));
        CELER_HIP_CALL(hipStreamSynchronize(state->state->stream()));
    });
}

int celer_state_reset(CelerState* state)
{
    return guarded([&] { state->state->reset(); });
}

int celer_state_synchronize(CelerState* state)
{
    return guarded([&] { CELER_HIP_CALL(hipStreamSynchronize(state->state->stream())); });
}

//---------------------------------------------------------------------------//
CelerStepper* celer_stepper_create(CelerParams* params,
                                   uint32_t stream_id,
                                   uint32_t num_slots,
                                   int action_times)
{
    return guarded_new<CelerStepper>([&] {
        StepperInput inp;
        inp.params = params->params;
        inp.stream_id = stream_id;
        inp.num_track_slots = num_slots;
        inp.action_times = action_times != 0;
        auto* result = new CelerStepper;
        result->params = params->params;
        result->stepper = std::make_shared<Stepper>(inp);
        result->state_handle.params = params->params;
        result->state_handle.state = result->stepper->sp_state();
        return result;
    });
}

void celer_stepper_destroy(CelerStepper* stepper)
{
    delete stepper;
}

int celer_stepper_step(CelerStepper* stepper,
                       const CelerPrimary* primaries,
                       uint32_t count,
                       CelerStepperResult* result)
{
    return guarded([&] {
        StepperResult r;
        if (primaries && count > 0)
            r = (*stepper->stepper)(reinterpret_cast<Primary const*>(primaries), count);
        else
            r = (*stepper->stepper)();
        if (result)
        {
            result->generated = r.generated;
            result->queued = r.queued;
            result->active = r.active;
            result->alive = r.alive;
        }
    });
}

int celer_stepper_run_steps(CelerStepper* stepper, uint32_t count, CelerStepperResult* result)
{
    return guarded([&] {
        StepperResult r = stepper->stepper->run_steps(count);
        if (result)
        {
            result->generated = r.generated;
            result->queued = r.queued;
            result->active = r.active;
            result->alive = r.alive;
        }
    });
}

int celer_stepper_warm_up(CelerStepper* stepper)
{
    return guarded([&] { stepper->stepper->warm_up(); });
}

//---------------------------------------------------------------------------//
CelerTransporter* celer_transporter_create(CelerParams* params,
                                           uint32_t max_num_tracks,
                                           uint32_t auto_flush,
                                           uint32_t max_steps,
                                           uint32_t stream_id,
                                           int nonfatal_flush,
                                           int action_times)
{
    return guarded_new<CelerTransporter>([&] {
        SetupOptions opts;
        opts.max_num_tracks = max_num_tracks;
        opts.auto_flush = auto_flush;
        opts.max_steps = max_steps;
        opts.stream_id = stream_id;
        opts.nonfatal_flush = nonfatal_flush != 0;
        opts.action_times = action_times != 0;
        auto* result = new CelerTransporter;
        result->params = params->params;
        result->transporter = std::make_unique<LocalTransporter>(opts, params->params);
        result->state_handle.params = params->params;
        result->state_handle.state = result->transporter->stepper().sp_state();
        return result;
    });
}

void celer_transporter_destroy(CelerTransporter* t)
{
    delete t;
}

int celer_transporter_initialize_event(CelerTransporter* t, int event_id)
{
    return guarded([&] { t->transporter->InitializeEvent(event_id); });
}

int celer_transporter_push(CelerTransporter* t, const CelerPrimary* tracks, uint32_t count)
{
    return guarded([&] {
        for (uint32_t i = 0; i < count; ++i)
            t->transporter->Push(reinterpret_cast<Primary const&>(tracks[i]));
    });
}

CelerServer* celer_server_create(CelerParams* params, uint32_t num_track_slots,
                                 uint32_t stream_id)
{
    return guarded_new<CelerServer>([&] {
        auto* result = new CelerServer;
        result->params = params->params;
        result->loop = std::make_unique<ServingLoop>(params->params, num_track_slots,
                                                     stream_id);
        result->state_handle.params = params->params;
        result->state_handle.state = result->loop->stepper().sp_state();
        return result;
    });
}

void celer_server_destroy(CelerServer* s)
{
    delete s;
}

CelerState* celer_server_state(CelerServer* s)
{
    return &s->state_handle;
}

int celer_server_set_steps_per_sync(CelerServer* s, uint32_t steps)
{
    return guarded([&] {
        CELER_VALIDATE(steps > 0, << "steps per readback must be positive");
        s->steps_per_sync = steps;
    });
}

int celer_server_run(CelerServer* s,
                     const CelerPrimary* primaries,
                     const uint32_t* event_offsets,
                     const double* arrivals,
                     uint32_t num_events,
                     uint32_t max_queued,
                     uint32_t max_steps,
                     double* admit_times,
                     double* done_times,
                     double* edep,
                     double* totals)
{
    return guarded([&] {
        std::vector<ServeEvent> events(num_events);
        for (uint32_t e = 0; e < num_events; ++e)
        {
            CELER_VALIDATE(event_offsets[e + 1] >= event_offsets[e],
                           << "event offsets must not decrease");
            events[e].first_primary = event_offsets[e];
            events[e].num_primaries = event_offsets[e + 1] - event_offsets[e];
            events[e].arrival = arrivals[e];
        }
        ServeOptions opts;
        opts.max_queued = max_queued;
        opts.max_steps = max_steps;
        opts.steps_per_sync = s->steps_per_sync;
        ServeResult r = s->loop->run(reinterpret_cast<Primary const*>(primaries),
                                     events.data(), num_events, opts);
        std::copy(r.admit_time.begin(), r.admit_time.end(), admit_times);
        std::copy(r.done_time.begin(), r.done_time.end(), done_times);
        std::copy(r.edep.begin(), r.edep.end(), edep);
        totals[0] = r.wall_time;
        totals[1] = static_cast<double>(r.steps);
        totals[2] = static_cast<double>(r.track_steps);
    });
}

int celer_state_set_block_size(CelerState* s, uint32_t block_size)
{
    return guarded([&] { s->state->block_size(block_size); });
}

int celer_stepper_capture_step(CelerStepper* s)
{
    return guarded([&] { s->stepper->capture_step(); });
}

int celer_transporter_flush(CelerTransporter* t)
{
    return guarded([&] { t->transporter->Flush(); });
}

int celer_transporter_finalize(CelerTransporter* t)
{
    return guarded([&] { t->transporter->Finalize(); });
}

uint32_t celer_transporter_buffer_size(const CelerTransporter* t)
{
    return t->transporter->GetBufferSize();
}

CelerState* celer_transporter_state(CelerTransporter* t)
{
    return &t->state_handle;
}

int celer_state_diagnostics(CelerState* state, int enable, uint64_t* out, uint32_t n,
                            int reset)
{
    return guarded([&] {
        CoreState& s = *state->state;
        if (enable >= 0)
            s.enable_diagnostics(enable != 0);
        if (out)
        {
            auto v = s.read_diagnostics(reset != 0);
            CELER_VALIDATE(!v.empty(), << "diagnostics are not enabled");
            for (uint32_t i = 0; i < n && i < v.size(); ++i)
                out[i] = v[i];
        }
    });
}

int celer_stepper_action_times(CelerStepper* stepper,
                               double* seconds,
                               uint64_t* counts,
                               uint32_t n,
                               int reset)
{
    return guarded([&] {
        auto const& t = stepper->stepper->action_times();
        auto const& c = stepper->stepper->action_counts();
        CELER_VALIDATE(n == t.size(),
                       << "expected " << t.size() << " step actions, got " << n);
        for (uint32_t i = 0; i < n; ++i)
        {
            if (seconds)
                seconds[i] = t[i];
            if (counts)
                counts[i] = c[i];
        }
        if (reset)
            stepper->stepper->reset_action_times();
    });
}

int celer_stepper_kill_active(CelerStepper* stepper)
{
    return guarded([&] { stepper->stepper->kill_active(); });
}

int celer_stepper_reseed(CelerStepper* stepper, uint64_t event_id)
{
    return guarded([&] { stepper->stepper->reseed(event_id); });
}

CelerState* celer_stepper_state(CelerStepper* stepper)
{
    return &stepper->state_handle;
}

int celer_stepper_action_sequence(const CelerStepper* stepper, char* buf, size_t len)
{
    return guarded([&] {
        std::string s;
        for (auto const& a : stepper->stepper->actions())
        {
            s += std::string(a->label());
            s += '\n';
        }
        std::snprintf(buf, len, "%s", s.c_str());
    });
}

//---------------------------------------------------------------------------//
int celer_hip_extend_from_primaries(const void* state_ref,
                                    const CelerPrimary* device_primaries,
                                    uint32_t count,
                                    void* stream)
{
    return launch_extend_from_primaries(
        *static_cast<CoreStateRef const*>(state_ref),
        reinterpret_cast<Primary const*>(device_primaries), nullptr, count,
        static_cast<hipStream_t>(stream));
}

int celer_hip_initialize_tracks(const void* params_ref, const void* state_ref, void* stream)
{
    return launch_initialize_tracks(*static_cast<CoreParamsRef const*>(params_ref),
                                    *static_cast<CoreStateRef const*>(state_ref),
                                    static_cast<hipStream_t>(stream));
}

int celer_hip_extend_from_secondaries(const void* params_ref,
                                      const void* state_ref,
                                      void* stream)
{
    return launch_extend_from_secondaries(
        *static_cast<CoreParamsRef const*>(params_ref),
        *static_cast<CoreStateRef const*>(state_ref), static_cast<hipStream_t>(stream));
}

int celer_hip_reseed(const void* params_ref,
                     const void* state_ref,
                     uint64_t event_id,
                     void* stream)
{
    auto const& p = *static_cast<CoreParamsRef const*>(params_ref);
    auto const& s = *static_cast<CoreStateRef const*>(state_ref);
    return launch_reseed_rng(s.rng, p.rng_jump, p.rng_seed, event_id,
                             static_cast<hipStream_t>(stream));
}

int celer_generate_physics(const char* spec_json, char* buf, size_t len, size_t* needed)
{
    return guarded([&] {
        copy_text(dump_json(generate_physics(parse_json(spec_json))), buf, len, needed);
    });
}

const uint32_t* celer_state_track_slots(const CelerState* state)
{
    return state->state->ref().track_slots;
}

const uint32_t* celer_state_action_offsets(const CelerState* state)
{
    return state->state->ref().action_offsets;
}

int celer_hip_pre_step_launch(const void* params_ref,
                              const void* state_ref,
                              const CelerThreadSlots* ts,
                              void* stream)
{
    auto const& p = params_of(params_ref);
    auto launch = p.scalars.precision == PrecisionMode::mixed ? launch_pre_step_mixed
                                                             : launch_pre_step;
    return launch(p, state_of(state_ref), slots_of(ts), static_cast<hipStream_t>(stream));
}

int celer_hip_along_step_launch(const void* params_ref,
                                const void* state_ref,
                                uint32_t action_id,
                                int charged,
                                const CelerThreadSlots* ts,
                                void* stream)
{
    auto const& p = params_of(params_ref);
    if (action_id >= p.scalars.num_actions)
        return hipErrorInvalidValue;
    auto const action = static_cast<std::uint8_t>(action_id);
    auto const st = static_cast<hipStream_t>(stream);
    if (!charged)
        return launch_along_step_neutral(p, state_of(state_ref), action, slots_of(ts), st);
    auto launch = p.scalars.precision == PrecisionMode::mixed
                      ? launch_along_step_charged_mixed
                      : launch_along_step_charged;
    return launch(p, state_of(state_ref), action, slots_of(ts), st);
}

int celer_hip_discrete_select_launch(const void* params_ref,
                                     const void* state_ref,
                                     const CelerThreadSlots* ts,
                                     void* stream)
{
    auto const& p = params_of(params_ref);
    auto launch = p.scalars.precision == PrecisionMode::mixed ? launch_discrete_select_mixed
                                                             : launch_discrete_select;
    return launch(p, state_of(state_ref), slots_of(ts), static_cast<hipStream_t>(stream));
}

int celer_hip_interaction_launch(const void* params_ref,
                                 const void* state_ref,
                                 uint32_t action_id,
                                 uint32_t model_class,
                                 const CelerThreadSlots* ts,
                                 void* stream)
{
    auto const& p = params_of(params_ref);
    if (action_id >= p.scalars.num_actions)
        return hipErrorInvalidValue;
    auto launch = p.scalars.precision == PrecisionMode::mixed ? launch_interaction_mixed
                                                             : launch_interaction;
    return launch(p, state_of(state_ref), static_cast<std::uint8_t>(action_id), model_class,
                  slots_of(ts), static_cast<hipStream_t>(stream));
}

int celer_hip_boundary_launch(const void* params_ref,
                              const void* state_ref,
                              const CelerThreadSlots* ts,
                              void* stream)
{
    return launch_boundary(params_of(params_ref), state_of(state_ref), slots_of(ts),
                           static_cast<hipStream_t>(stream));
}

int celer_hip_tracking_cut_launch(const void* params_ref,
                                  const void* state_ref,
                                  const CelerThreadSlots* ts,
                                  void* stream)
{
    return launch_tracking_cut(params_of(params_ref), state_of(state_ref), slots_of(ts),
                               static_cast<hipStream_t>(stream));
}

int celer_hip_calo_tally_launch(const void* params_ref,
                                const void* state_ref,
                                uint32_t num_detectors,
                                const CelerThreadSlots* ts,
                                void* stream)
{
    return launch_calo_tally(params_of(params_ref), state_of(state_ref), num_detectors,
                             slots_of(ts), static_cast<hipStream_t>(stream));
}

int celer_hip_sort_tracks_along_before_pre(const void* params_ref,
                                           const void* state_ref,
                                           uint32_t num_keys,
                                           void* stream)
{
    return launch_sort_tracks(state_of(state_ref), SortKey::along_action_of_particle,
                              num_keys, static_cast<hipStream_t>(stream),
                              params_of(params_ref).particle_along_action);
}

int celer_hip_sort_tracks(const void* state_ref,
                          uint32_t key_kind,
                          uint32_t num_keys,
                          void* stream)
{
    if (key_kind > static_cast<uint32_t>(SortKey::post_action))
        return hipErrorInvalidValue;
    return launch_sort_tracks(state_of(state_ref), static_cast<SortKey>(key_kind), num_keys,
                              static_cast<hipStream_t>(stream));
}

int celer_hip_partition_keys(const uint8_t* keys,
                             uint32_t n,
                             uint32_t num_keys,
                             uint32_t* sorted,
                             uint32_t* offsets,
                             void* stream)
{
    if (num_keys < 1 || num_keys > max_sort_keys || (n > 0 && (!keys || !sorted)) || !offsets)
        return hipErrorInvalidValue;
    auto const st = static_cast<hipStream_t>(stream);
    std::uint32_t* counts = nullptr;
    std::size_t const words = sort_counts_size(n);
    if (words > 0)
    {
        hipError_t err = hipMallocAsync(reinterpret_cast<void**>(&counts),
                                        words * sizeof(std::uint32_t), st);
        if (err != hipSuccess)
            return err;
    }
    hipError_t err = launch_partition_keys(keys, n, num_keys, counts, sorted, offsets, st);
    if (counts)
    {
        hipError_t const ferr = hipFreeAsync(counts, st);
        if (err == hipSuccess)
            err = ferr;
    }
    return err;
}

int celer_hip_partition_keys_noalloc(const uint8_t* keys,
                                     uint32_t n,
                                     uint32_t num_keys,
                                     uint32_t* sorted,
                                     uint32_t* offsets,
                                     void* stream)
{
    static_assert(CELER_PARTITION_NOALLOC_MAX == max_partition_noalloc,
                  "allocation-free partition limit");
    g_last_error.clear();
    if (n > max_partition_noalloc)
    {
        g_last_error = "celer_hip_partition_keys_noalloc: " + std::to_string(n)
                       + " keys exceed the allocation-free limit of "
                       + std::to_string(max_partition_noalloc);
        return hipErrorInvalidValue;
    }
    if (num_keys < 1 || num_keys > max_sort_keys || (n > 0 && (!keys || !sorted)) || !offsets)
    {
        g_last_error = "celer_hip_partition_keys_noalloc: invalid arguments";
        return hipErrorInvalidValue;
    }
    std::uint32_t* counts = nullptr;
    hipError_t err = partition_noalloc_scratch(&counts);
    if (err != hipSuccess)
        return err;
    return launch_partition_keys(keys, n, num_keys, counts, sorted, offsets,
                                 static_cast<hipStream_t>(stream));
}

int celer_kernel_attributes_json(char* buf, size_t len, size_t* needed)
{
    return guarded(
        [&] { copy_text(KernelRegistry::instance().to_json(), buf, len, needed); });
}

int celer_kernel_timing(int enable)
{
    return guarded([&] { KernelRegistry::instance().timing(enable != 0); });
}

int celer_kernel_times_json(int reset, char* buf, size_t len, size_t* needed)
{
    // A size query (no buffer) leaves the times in place for the call that
    // reads them
    return guarded([&] {
        copy_text(KernelRegistry::instance().times_json(reset != 0 && buf != nullptr), buf,
                  len, needed);
    });
}

int celer_read_seltzer_berger(const char* dir, int z, char* buf, size_t len, size_t* needed)
{
    return guarded([&] {
        SBTable const t = dir ? SeltzerBergerReader(dir)(z) : SeltzerBergerReader()(z);
        copy_text(dump_json(to_json(t)), buf, len, needed);
    });
}

int celer_read_livermore_pe(const char* dir, int z, char* buf, size_t len, size_t* needed)
{
    return guarded([&] {
        JsonValue const r = dir ? LivermorePEReader(dir)(z) : LivermorePEReader()(z);
        copy_text(dump_json(r), buf, len, needed);
    });
}

int celer_params_export_json(const CelerParams*, char* buf, size_t len, size_t* needed)
{
    return guarded([&] {
        std::string s = "{}";
        if (needed)
            *needed = s.size() + 1;
        if (buf && len > 0)
            std::snprintf(buf, len, "%s", s.c_str());
    });
}

}  // extern "C"

//---------------------------------------------------------------------------//
// USER ACTIONS
//---------------------------------------------------------------------------//
struct CelerActionDiagnostic
{
    std::shared_ptr<ActionDiagnostic> diag;
};

struct CelerStepDiagnostic
{
    std::shared_ptr<StepDiagnostic> diag;
};

struct CelerStepCollector
{
    std::shared_ptr<StepCollector> coll;
};

struct CelerDetectorSteps
{
    DetectorStepOutput out;
};

namespace
{
template<class V>
int copy_column(V const& v, void* dst, size_t bytes, char const* name)
{
    return guarded([&] {
        CELER_VALIDATE(!v.empty() || bytes == 0,
                       << "detector step column '" << name << "' is empty (not selected or no hits)");
        size_t const need = v.size() * sizeof(v[0]);
        CELER_VALIDATE(bytes >= need, << "buffer of " << bytes << " bytes for column '"
                                      << name << "' needs " << need);
        if (need)
            std::memcpy(dst, v.data(), need);
    });
}
}  // namespace

extern "C" {
int celer_action_diagnostic_create(CelerParams* params, CelerActionDiagnostic** out)
{
    return guarded([&] {
        CELER_VALIDATE(params && out, << "null argument");
        auto h = std::make_unique<CelerActionDiagnostic>();
        h->diag = ActionDiagnostic::make_and_insert(*params->params);
        *out = h.release();
    });
}

void celer_action_diagnostic_destroy(CelerActionDiagnostic* diag)
{
    delete diag;
}

int celer_action_diagnostic_shape(const CelerActionDiagnostic* diag,
                                  uint32_t* num_particles,
                                  uint32_t* num_actions)
{
    return guarded([&] {
        CELER_VALIDATE(diag && num_particles && num_actions, << "null argument");
        *num_particles = diag->diag->num_particles();
        *num_actions = diag->diag->num_actions();
    });
}

int celer_action_diagnostic_counts(const CelerActionDiagnostic* diag, uint64_t* out, size_t size)
{
    return guarded([&] {
        CELER_VALIDATE(diag && (out || size == 0), << "null argument");
        auto counts = diag->diag->calc_actions();
        CELER_VALIDATE(size >= counts.size(), << "output needs " << counts.size()
                                              << " entries, got " << size);
        std::copy(counts.begin(), counts.end(), out);
    });
}

int celer_action_diagnostic_clear(CelerActionDiagnostic* diag)
{
    return guarded([&] {
        CELER_VALIDATE(diag, << "null argument");
        diag->diag->clear();
    });
}

int celer_step_diagnostic_create(CelerParams* params,
                                 uint32_t max_step_bin,
                                 CelerStepDiagnostic** out)
{
    return guarded([&] {
        CELER_VALIDATE(params && out, << "null argument");
        auto h = std::make_unique<CelerStepDiagnostic>();
        h->diag = StepDiagnostic::make_and_insert(*params->params, max_step_bin);
        *out = h.release();
    });
}

void celer_step_diagnostic_destroy(CelerStepDiagnostic* diag)
{
    delete diag;
}

int celer_step_diagnostic_shape(const CelerStepDiagnostic* diag,
                                uint32_t* num_particles,
                                uint32_t* num_bins)
{
    return guarded([&] {
        CELER_VALIDATE(diag && num_particles && num_bins, << "null argument");
        *num_particles = diag->diag->num_particles();
        *num_bins = diag->diag->num_bins();
    });
}

int celer_step_diagnostic_counts(const CelerStepDiagnostic* diag, uint64_t* out, size_t size)
{
    return guarded([&] {
        CELER_VALIDATE(diag && (out || size == 0), << "null argument");
        auto counts = diag->diag->calc_steps();
        CELER_VALIDATE(size >= counts.size(), << "output needs " << counts.size()
                                              << " entries, got " << size);
        std::copy(counts.begin(), counts.end(), out);
    });
}

int celer_step_diagnostic_clear(CelerStepDiagnostic* diag)
{
    return guarded([&] {
        CELER_VALIDATE(diag, << "null argument");
        diag->diag->clear();
    });
}

int celer_step_collector_create(CelerParams* params,
                                const CelerStepSelection* selection,
                                const uint32_t* volume_to_detector,
                                uint32_t num_volumes,
                                int nonzero_energy_deposition,
                                uint32_t hit_capacity,
                                CelerStepCollector** out)
{
    return guarded([&] {
        CELER_VALIDATE(params && selection && out, << "null argument");
        StepCollector::Input inp;
        std::memcpy(static_cast<void*>(&inp.selection), selection, sizeof(StepSelection));
        if (volume_to_detector)
            inp.volume_to_detector.assign(volume_to_detector, volume_to_detector + num_volumes);
        inp.nonzero_energy_deposition = nonzero_energy_deposition != 0;
        inp.hit_capacity = hit_capacity;
        auto h = std::make_unique<CelerStepCollector>();
        h->coll = StepCollector::make_and_insert(*params->params, std::move(inp));
        *out = h.release();
    });
}

void celer_step_collector_destroy(CelerStepCollector* coll)
{
    delete coll;
}

int celer_step_collector_read(CelerStepCollector* coll,
                              uint32_t stream_id,
                              CelerDetectorSteps** out)
{
    return guarded([&] {
        CELER_VALIDATE(coll && out, << "null argument");
        auto h = std::make_unique<CelerDetectorSteps>();
        h->out = coll->coll->read(stream_id);
        *out = h.release();
    });
}

void celer_detector_steps_destroy(CelerDetectorSteps* steps)
{
    delete steps;
}

uint32_t celer_detector_steps_size(const CelerDetectorSteps* steps)
{
    return steps ? steps->out.size() : 0;
}

int celer_detector_steps_overflow(const CelerDetectorSteps* steps,
                                  uint32_t* overflow_steps,
                                  uint32_t* dropped_hits)
{
    return guarded([&] {
        CELER_VALIDATE(steps && overflow_steps && dropped_hits, << "null argument");
        *overflow_steps = steps->out.overflow_steps;
        *dropped_hits = steps->out.dropped_hits;
    });
}

int celer_detector_steps_column(const CelerDetectorSteps* steps,
                                const char* name,
                                void* dst,
                                size_t bytes)
{
    if (!steps || !name || (!dst && bytes))
        return guarded([] { CELER_VALIDATE(false, << "null argument"); });
    DetectorStepOutput const& o = steps->out;
    std::string n(name);
    if (n == "detector")
        return copy_column(o.detector, dst, bytes, name);
    if (n == "track_id")
        return copy_column(o.track_id, dst, bytes, name);
    if (n == "event_id")
        return copy_column(o.event_id, dst, bytes, name);
    if (n == "parent_id")
        return copy_column(o.parent_id, dst, bytes, name);
    if (n == "track_step_count")
        return copy_column(o.track_step_count, dst, bytes, name);
    if (n == "action_id")
        return copy_column(o.action_id, dst, bytes, name);
    if (n == "particle")
        return copy_column(o.particle, dst, bytes, name);
    if (n == "step_length")
        return copy_column(o.step_length, dst, bytes, name);
    if (n == "energy_deposition")
        return copy_column(o.energy_deposition, dst, bytes, name);
    for (int p = 0; p < 2; ++p)
    {
        std::string const prefix = p == 0 ? "pre." : "post.";
        if (n.compare(0, prefix.size(), prefix) != 0)
            continue;
        std::string const f = n.substr(prefix.size());
        DetectorStepOutput::Point const& pt = o.points[p];
        if (f == "time")
            return copy_column(pt.time, dst, bytes, name);
        if (f == "pos")
            return copy_column(pt.pos, dst, bytes, name);
        if (f == "dir")
            return copy_column(pt.dir, dst, bytes, name);
        if (f == "volume_id")
            return copy_column(pt.volume_id, dst, bytes, name);
        if (f == "energy")
            return copy_column(pt.energy, dst, bytes, name);
    }
    return guarded([&] { CELER_VALIDATE(false, << "unknown detector step column '" << n << "'"); });
}
}
