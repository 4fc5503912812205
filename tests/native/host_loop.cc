// Host step loop: the product's per-track step bodies (StepExecutors.hh,
// AlongStepExecutor.hh, TrackInitImpl.hh) driven slot by slot on the CPU
// over a whole workload, in the mixed-precision or the FP64 build, with
// the kernels' precision split (mixed: pre-step, charged along-step,
// discrete select and interactions; FP64: neutral along-step, boundary,
// tracking cut and track initialization).
//
// Track management is a plain serial restatement of the device's: new
// tracks pop initializers from the back into vacancies from the back;
// after the post-step actions, surviving secondaries become initializers
// and dead slots vacancies. Events are dealt round-robin to worker threads,
// each with its own state (the reference's event-parallel host mode,
// app/celer-sim/celer-sim.cc:115-137).
//
// Every per-track call is timed in CPU cycles and a watchdog thread ends
// the run if one call stays in progress for longer than a limit, printing
// the call's phase and the track's full input state (including its RNG
// words, so the call can be replayed). The longest call of each phase is
// reported with its inputs.
//
// Usage: host_loop problem.json primaries.bin slots threads mixed|fp64
//                  [stall_seconds] [max_iterations]
//        host_loop problem.json --replay 'snapshot json line' 1 mixed
#include <atomic>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>
#include <x86intrin.h>

#include <hip/hip_runtime_api.h>

#include "global/AlongStepExecutor.hh"
#include "global/CoreParams.hh"
#include "global/StepExecutors.hh"
#include "phys/PhysicsActions.hh"
#include "track/TrackInitImpl.hh"

#include "host_loop.hh"

// Device memory is host memory here (hip_host_stubs.cc): the params'
// tables are host arrays

using namespace celeritas_amd;

namespace
{
//---------------------------------------------------------------------------//
enum Phase : int
{
    ph_init,
    ph_pre,
    ph_along_charged,
    ph_along_neutral,
    ph_discrete,
    ph_interact,
    ph_boundary,
    ph_cut,
    num_phases
};
char const* const phase_names[] = {"init", "pre-step", "along-charged", "along-neutral",
                                   "discrete", "interaction", "boundary", "tracking-cut"};

//! Inputs of one per-track call
struct Snapshot
{
    int phase{-1};
    std::uint32_t slot{0};
    std::uint32_t event{0};
    std::uint32_t track{0};
    std::uint32_t particle{0};
    std::uint32_t volume{0};
    std::uint32_t status{0};
    std::uint32_t post{0};
    std::uint32_t along{0};
    std::uint32_t element{0};
    double energy{0};
    double step{0};
    double mfp{0};
    double macro_xs{0};
    double range{0};
    double pos[3]{};
    double dir[3]{};
    double msc[3]{};
    std::uint32_t rng[6]{};
};

void print_snapshot(FILE* f, char const* what, Snapshot const& s, std::uint64_t cycles)
{
    std::fprintf(f,
                 "{\"what\": \"%s\", \"phase\": \"%s\", \"cycles\": %" PRIu64
                 ", \"slot\": %u, \"event\": %u, \"track\": %u, \"particle\": %u, "
                 "\"volume\": %u, \"status\": %u, \"post\": %u, \"along\": %u, "
                 "\"element\": %u, \"energy\": %.17g, \"step\": %.17g, \"mfp\": %.17g, "
                 "\"macro_xs\": %.17g, \"range\": %.17g, \"pos\": [%.17g, %.17g, %.17g], "
                 "\"dir\": [%.17g, %.17g, %.17g], \"msc\": [%.17g, %.17g, %.17g], "
                 "\"rng\": [%u, %u, %u, %u, %u, %u]}\n",
                 what, s.phase >= 0 ? phase_names[s.phase] : "none", cycles, s.slot,
                 s.event, s.track, s.particle, s.volume, s.status, s.post, s.along,
                 s.element, s.energy, s.step, s.mfp, s.macro_xs, s.range, s.pos[0],
                 s.pos[1], s.pos[2], s.dir[0], s.dir[1], s.dir[2], s.msc[0], s.msc[1],
                 s.msc[2], s.rng[0], s.rng[1], s.rng[2], s.rng[3], s.rng[4], s.rng[5]);
    std::fflush(f);
}

//---------------------------------------------------------------------------//
//! Host arrays of one state
struct HostState
{
    std::vector<std::uint32_t> u32;
    std::vector<double> f64;
    std::vector<std::uint8_t> u8;
    std::vector<Secondary> secondaries;
    CoreCounters counters{};
    CoreStateRef ref;

    HostState(size_type n, std::uint32_t num_proc)
    {
        // Offsets into three pools, filled in the order of the fields below
        size_type const nu32 = 16 + 6, nf64 = 17 + num_proc, nu8 = 12;
        u32.assign(std::size_t(n) * nu32, 0);
        f64.assign(std::size_t(n) * nf64, 0);
        u8.assign(std::size_t(n) * nu8, 0);
        secondaries.resize(std::size_t(3) * n);
        std::uint32_t* pu = u32.data();
        double* pf = f64.data();
        std::uint8_t* pb = u8.data();
        auto U = [&] { auto* r = pu; pu += n; return r; };
        auto F = [&] { auto* r = pf; pf += n; return r; };
        auto B = [&] { auto* r = pb; pb += n; return r; };
        CoreStateRef& s = ref;
        s.size = n;
        s.sim.track_id = U();
        s.sim.parent_id = U();
        s.sim.event_id = U();
        s.sim.num_steps = U();
        s.sim.time = F();
        s.sim.status = B();
        s.sim.step_length = F();
        s.sim.post_step_action = B();
        s.sim.along_step_action = B();
        s.sim.num_looping_steps = U();
        s.particle.particle_id = B();
        s.particle.energy = F();
        for (int k = 0; k < 3; ++k)
        {
            s.geo.pos[k] = F();
            s.geo.dir[k] = F();
        }
        s.geo.volume = U();
        s.geo.surface = U();
        s.geo.sense = B();
        s.geo.boundary = B();
        s.geo.failed = B();
        s.geo.next_step = F();
        s.geo.next_surface = U();
        s.geo.next_sense = B();
        s.phys.interaction_mfp = F();
        s.phys.macro_xs = F();
        s.phys.per_process_xs = pf;
        pf += std::size_t(n) * num_proc;
        s.phys.dedx_range = F();
        s.phys.energy_deposition = F();
        s.phys.sec_start = U();
        s.phys.sec_count = U();
        s.phys.element = B();
        s.phys.msc_range_init = F();
        s.phys.msc_range_factor = F();
        s.phys.msc_limit_min = F();
        s.phys.msc_true_path = F();
        s.phys.msc_geom_path = F();
        s.phys.msc_alpha = F();
        s.phys.msc_is_displaced = B();
        s.phys.msc_enabled = B();
        for (int k = 0; k < 5; ++k)
            s.rng.xs[k] = U();
        s.rng.weyl = U();
        s.rng.size = n;
        s.pre_volume = U();
        s.secondaries.storage = secondaries.data();
        s.secondaries.capacity = static_cast<size_type>(secondaries.size());
        s.counters = &counters;
    }
};

//! SplitMix64 (a distinct, well-mixed XORWOW state per slot)
std::uint64_t splitmix(std::uint64_t& x)
{
    std::uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Options
{
    bool mixed{true};
    double stall_seconds{20};
    std::uint64_t max_iterations{200000};
};

struct Actions
{
    std::uint8_t discrete, boundary, cut, neutral, charged;
    std::vector<int> model_class;  //!< per action id: ModelClass or -1
};

//! Per-worker results and watchdog record
struct Worker
{
    std::atomic<std::uint64_t> seq{0};
    std::atomic<int> in_call{0};
    Snapshot current;
    std::uint64_t max_cycles[num_phases]{};
    Snapshot worst[num_phases];
    std::uint64_t calls[num_phases]{};
    std::uint64_t track_steps{0};
    std::uint64_t iterations{0};
    std::uint64_t errored{0};
    std::uint64_t max_track_steps{0};
    std::uint64_t events{0};
    std::vector<double> edep;
    bool hit_max_iterations{false};
    std::uint64_t sampling_failures{0};
};

void snapshot(CoreStateRef const& s, size_type slot, int phase, Snapshot* out)
{
    out->phase = phase;
    out->slot = slot;
    out->event = s.sim.event_id[slot];
    out->track = s.sim.track_id[slot];
    out->particle = s.particle.particle_id[slot];
    out->volume = s.geo.volume[slot];
    out->status = s.sim.status[slot];
    out->post = s.sim.post_step_action[slot];
    out->along = s.sim.along_step_action[slot];
    out->element = s.phys.element[slot];
    out->energy = s.particle.energy[slot];
    out->step = s.sim.step_length[slot];
    out->mfp = s.phys.interaction_mfp[slot];
    out->macro_xs = s.phys.macro_xs[slot];
    out->range = s.phys.dedx_range[slot];
    for (int k = 0; k < 3; ++k)
    {
        out->pos[k] = s.geo.pos[k][slot];
        out->dir[k] = s.geo.dir[k][slot];
    }
    out->msc[0] = s.phys.msc_range_init[slot];
    out->msc[1] = s.phys.msc_range_factor[slot];
    out->msc[2] = s.phys.msc_limit_min[slot];
    for (int k = 0; k < 5; ++k)
        out->rng[k] = s.rng.xs[k][slot];
    out->rng[5] = s.rng.weyl[slot];
}

template<class F>
inline void timed(Worker& w, CoreStateRef const& s, size_type slot, int phase, F&& f)
{
    snapshot(s, slot, phase, &w.current);
    w.seq.fetch_add(1, std::memory_order_relaxed);
    w.in_call.store(1, std::memory_order_release);
    std::uint64_t const t0 = __rdtsc();
    f();
    std::uint64_t const dt = __rdtsc() - t0;
    w.in_call.store(0, std::memory_order_release);
    ++w.calls[phase];
    if (dt > w.max_cycles[phase])
    {
        w.max_cycles[phase] = dt;
        w.worst[phase] = w.current;
    }
}

//---------------------------------------------------------------------------//
void run_worker(CoreParamsRef const& p,
                Actions const& act,
                std::vector<Primary> const& prims,
                size_type nslots,
                std::uint32_t num_proc,
                std::uint64_t seed,
                std::size_t num_det,
                Options const& opt,
                Worker& w)
{
    HostState hs(nslots, num_proc);
    CoreStateRef const& s = hs.ref;
    std::uint64_t sm = seed;
    for (size_type i = 0; i < nslots; ++i)
    {
        std::uint64_t const a = splitmix(sm), b = splitmix(sm), c = splitmix(sm);
        s.rng.xs[0][i] = std::uint32_t(a);
        s.rng.xs[1][i] = std::uint32_t(a >> 32) | 1u;
        s.rng.xs[2][i] = std::uint32_t(b);
        s.rng.xs[3][i] = std::uint32_t(b >> 32);
        s.rng.xs[4][i] = std::uint32_t(c);
        s.rng.weyl[i] = std::uint32_t(c >> 32);
    }
    w.edep.assign(num_det, 0.0);
    std::vector<TrackInitializer> inits;
    std::vector<std::uint32_t> track_counter;
    for (Primary const& pr : prims)
    {
        if (pr.event_id >= track_counter.size())
            track_counter.resize(pr.event_id + 1, 0);
        TrackInitializer ti{};
        ti.track_id = track_counter[pr.event_id]++;
        ti.parent_id = invalid_id;
        ti.event_id = pr.event_id;
        ti.particle_id = pr.particle_id;
        ti.time = pr.time;
        for (int k = 0; k < 3; ++k)
        {
            ti.pos[k] = pr.pos[k];
            ti.dir[k] = pr.dir[k];
        }
        ti.energy = pr.energy;
        ti.volume = invalid_id;  // located from the position
        inits.push_back(ti);
    }
    // The reference pushes primaries in order and pops from the back
    std::vector<std::uint32_t> vacancies;
    for (size_type i = 0; i < nslots; ++i)
        vacancies.push_back(nslots - 1 - i);
    size_type alive = 0;
    while (!inits.empty() || alive > 0)
    {
        if (w.iterations >= opt.max_iterations)
        {
            w.hit_max_iterations = true;
            break;
        }
        ++w.iterations;
        // Initialize
        while (!inits.empty() && !vacancies.empty())
        {
            TrackInitializer const ti = inits.back();
            inits.pop_back();
            size_type const slot = vacancies.back();
            vacancies.pop_back();
            timed(w, s, slot, ph_init, [&] { init_track(p, s, slot, ti); });
            ++alive;
        }
        // Pre-step
        hs.counters.secondary_stack_size = 0;
        for (size_type slot = 0; slot < nslots; ++slot)
        {
            auto const st = static_cast<TrackStatus>(s.sim.status[slot]);
            if (st == TrackStatus::inactive)
            {
                exec::pre_step_inactive(s, slot);
                continue;
            }
            ++w.track_steps;
            timed(w, s, slot, ph_pre, [&] {
                if (opt.mixed)
                    hl::mixed_pre_step(p, s, slot, st);
                else
                    exec::pre_step_track(p, s, slot, st);
            });
        }
        // Along-step
        for (size_type slot = 0; slot < nslots; ++slot)
        {
            auto const st = static_cast<TrackStatus>(s.sim.status[slot]);
            if (st == TrackStatus::inactive)
                continue;
            std::uint8_t const a = s.sim.along_step_action[slot];
            if (a == act.neutral)
            {
                timed(w, s, slot, ph_along_neutral,
                      [&] { exec::along_step_neutral_track(p, s, slot, st); });
            }
            else if (a == act.charged)
            {
                timed(w, s, slot, ph_along_charged, [&] {
                    if (opt.mixed)
                        hl::mixed_along_charged(p, s, slot, st);
                    else if (p.field.enabled)
                        exec::along_step_charged_track<true>(p, s, slot, st);
                    else
                        exec::along_step_charged_track<false>(p, s, slot, st);
                });
            }
        }
        // Discrete select (pre-post)
        for (size_type slot = 0; slot < nslots; ++slot)
        {
            if (s.sim.post_step_action[slot] != act.discrete
                || s.sim.status[slot] == std::uint8_t(TrackStatus::inactive))
                continue;
            timed(w, s, slot, ph_discrete, [&] {
                if (opt.mixed)
                    hl::mixed_discrete(p, s, slot);
                else
                    exec::discrete_select_track(p, s, slot);
            });
        }
        // Post-step
        for (size_type slot = 0; slot < nslots; ++slot)
        {
            auto const st = static_cast<TrackStatus>(s.sim.status[slot]);
            if (st == TrackStatus::inactive)
                continue;
            std::uint8_t const post = s.sim.post_step_action[slot];
            if (post < act.model_class.size() && act.model_class[post] >= 0)
            {
                auto const m = static_cast<ModelClass>(act.model_class[post]);
                timed(w, s, slot, ph_interact, [&] {
                    if (opt.mixed)
                    {
                        hl::mixed_interaction(p, s, slot, m);
                        return;
                    }
                    switch (m)
                    {
#define HL_MODEL(M)                                          \
    case ModelClass::M:                                      \
        exec::interaction_track<ModelClass::M>(p, s, slot); \
        break;
                        HL_MODEL(klein_nishina)
                        HL_MODEL(livermore_pe)
                        HL_MODEL(bethe_heitler)
                        HL_MODEL(moller_bhabha)
                        HL_MODEL(seltzer_berger)
                        HL_MODEL(relativistic_brem)
                        HL_MODEL(eplusgg)
                        HL_MODEL(combined_brem)
#undef HL_MODEL
                        default:
                            break;
                    }
                });
            }
            else if (post == act.boundary && st == TrackStatus::alive)
            {
                timed(w, s, slot, ph_boundary, [&] { exec::boundary_track(p, s, slot); });
            }
        }
        for (size_type slot = 0; slot < nslots; ++slot)
        {
            auto const st = static_cast<TrackStatus>(s.sim.status[slot]);
            if (st == TrackStatus::inactive || s.sim.post_step_action[slot] != act.cut)
                continue;
            if (st == TrackStatus::errored)
                ++w.errored;
            timed(w, s, slot, ph_cut, [&] { exec::tracking_cut_track(p, s, slot); });
        }
        // Tally, secondaries and vacancies
        for (size_type slot = 0; slot < nslots; ++slot)
        {
            auto const st = static_cast<TrackStatus>(s.sim.status[slot]);
            if (st == TrackStatus::inactive)
                continue;
            double const e = s.phys.energy_deposition[slot];
            if (e != 0)
            {
                std::uint32_t const det = p.volume_to_detector
                                              ? p.volume_to_detector[s.pre_volume[slot]]
                                              : invalid_id;
                if (det != invalid_id && det < num_det)
                    w.edep[det] += e;
            }
            std::uint32_t const start = s.phys.sec_start[slot];
            std::uint32_t const count = s.phys.sec_count[slot];
            for (std::uint32_t k = 0; k < count; ++k)
            {
                Secondary const& sec = s.secondaries.storage[start + k];
                if (sec.particle_id == invalid_id)
                    continue;
                std::uint32_t const ev = s.sim.event_id[slot];
                TrackInitializer ti{};
                ti.track_id = track_counter[ev]++;
                ti.parent_id = s.sim.track_id[slot];
                ti.event_id = ev;
                ti.particle_id = sec.particle_id;
                ti.time = s.sim.time[slot];
                for (int d = 0; d < 3; ++d)
                {
                    ti.pos[d] = s.geo.pos[d][slot];
                    ti.dir[d] = sec.dir[d];
                }
                ti.energy = sec.energy;
                // Made in the parent's volume unless the parent is on a surface
                ti.volume = (s.geo.surface[slot] == invalid_id && !s.geo.failed[slot])
                                ? s.geo.volume[slot]
                                : invalid_id;
                inits.push_back(ti);
            }
            if (st == TrackStatus::killed)
            {
                std::uint64_t const n = s.sim.num_steps[slot];
                if (n > w.max_track_steps)
                    w.max_track_steps = n;
                s.sim.status[slot] = std::uint8_t(TrackStatus::inactive);
                vacancies.push_back(slot);
                --alive;
            }
        }
    }
    w.events = track_counter.size();
    w.sampling_failures = hs.counters.num_sampling_failures;
}
}  // namespace

//---------------------------------------------------------------------------//
/*!
 * Rerun one per-track call from a snapshot line printed by the loop (the
 * track's state, including its RNG words, is restored into slot 0 of a
 * one-slot state) and print the track afterwards.
 */
int replay(CoreParamsRef const& p, Actions const& act, std::uint32_t num_proc,
           std::string const& line, bool mixed)
{
    JsonValue const j = parse_json(line);
    HostState hs(1, num_proc);
    CoreStateRef const& s = hs.ref;
    std::string const phase = j.at("phase").as_string();
    s.sim.event_id[0] = static_cast<std::uint32_t>(j.get_int("event", 0));
    s.sim.track_id[0] = static_cast<std::uint32_t>(j.get_int("track", 0));
    s.particle.particle_id[0] = static_cast<std::uint8_t>(j.get_int("particle", 0));
    s.geo.volume[0] = static_cast<std::uint32_t>(j.get_int("volume", 0));
    s.sim.status[0] = static_cast<std::uint8_t>(j.get_int("status", 0));
    s.sim.post_step_action[0] = static_cast<std::uint8_t>(j.get_int("post", 0));
    s.sim.along_step_action[0] = static_cast<std::uint8_t>(j.get_int("along", 0));
    s.phys.element[0] = static_cast<std::uint8_t>(j.get_int("element", 0));
    s.particle.energy[0] = j.at("energy").as_double();
    s.sim.step_length[0] = j.at("step").as_double();
    s.phys.interaction_mfp[0] = j.at("mfp").as_double();
    s.phys.macro_xs[0] = j.at("macro_xs").as_double();
    s.phys.dedx_range[0] = j.at("range").as_double();
    auto const pos = j.at("pos").as_doubles();
    auto const dir = j.at("dir").as_doubles();
    auto const msc = j.at("msc").as_doubles();
    for (int k = 0; k < 3; ++k)
    {
        s.geo.pos[k][0] = pos[k];
        s.geo.dir[k][0] = dir[k];
    }
    s.geo.surface[0] = invalid_id;
    s.phys.msc_range_init[0] = msc[0];
    s.phys.msc_range_factor[0] = msc[1];
    s.phys.msc_limit_min[0] = msc[2];
    auto const words = j.at("rng").as_doubles();
    for (int k = 0; k < 5; ++k)
        s.rng.xs[k][0] = static_cast<std::uint32_t>(words[k]);
    s.rng.weyl[0] = static_cast<std::uint32_t>(words[5]);
    auto const st = static_cast<TrackStatus>(s.sim.status[0]);
    std::uint8_t const post = s.sim.post_step_action[0];
    std::fprintf(stderr, "replaying %s (post action %u)\n", phase.c_str(), post);
    if (phase == "interaction" && post < act.model_class.size() && act.model_class[post] >= 0)
    {
        auto const m = static_cast<ModelClass>(act.model_class[post]);
        std::fprintf(stderr, "model class %d\n", static_cast<int>(m));
        if (mixed)
            hl::mixed_interaction(p, s, 0, m);
    }
    else if (phase == "along-charged" && mixed)
        hl::mixed_along_charged(p, s, 0, st);
    else if (phase == "pre-step" && mixed)
        hl::mixed_pre_step(p, s, 0, st);
    else if (phase == "discrete" && mixed)
        hl::mixed_discrete(p, s, 0);
    Snapshot out;
    snapshot(s, 0, -1, &out);
    print_snapshot(stdout, "after", out, 0);
    return 0;
}

//---------------------------------------------------------------------------//
int main(int argc, char** argv)
{
    if (argc < 6)
    {
        std::fprintf(stderr,
                     "usage: %s problem.json primaries.bin slots threads mixed|fp64 "
                     "[stall_seconds] [max_iterations]\n",
                     argv[0]);
        return 2;
    }
    std::ifstream f(argv[1]);
    std::stringstream ss;
    ss << f.rdbuf();
    auto params = CoreParams::from_json(ss.str());
    CoreParamsRef const& p = params->device_ref();

    bool const is_replay = std::string(argv[2]) == "--replay";
    std::ifstream pf(is_replay ? "/dev/null" : argv[2], std::ios::binary);
    std::vector<char> raw((std::istreambuf_iterator<char>(pf)), std::istreambuf_iterator<char>());
    std::vector<Primary> prims(raw.size() / sizeof(Primary));
    std::memcpy(prims.data(), raw.data(), prims.size() * sizeof(Primary));
    size_type const nslots = std::atoi(argv[3]);
    int const nthreads = std::atoi(argv[4]);
    Options opt;
    opt.mixed = std::string(argv[5]) == "mixed";
    if (argc > 6)
        opt.stall_seconds = std::atof(argv[6]);
    if (argc > 7)
        opt.max_iterations = std::strtoull(argv[7], nullptr, 10);

    Actions act;
    act.discrete = p.phys.scalars.discrete_action;
    act.boundary = p.scalars.boundary_action;
    act.cut = p.scalars.tracking_cut_action;
    act.neutral = p.scalars.along_step_neutral_action;
    act.charged = p.scalars.along_step_user_action;
    auto const& reg = *params->action_reg();
    act.model_class.assign(reg.num_actions(), -1);
    for (size_type i = 0; i < reg.num_actions(); ++i)
    {
        if (auto const* m = dynamic_cast<ModelAction const*>(reg.action(ActionId{i}).get()))
            act.model_class[i] = static_cast<int>(m->model_class());
    }
    std::uint32_t num_proc = 0;
    for (std::uint32_t i = 0; i < p.phys.num_particles; ++i)
        num_proc = std::max(num_proc, p.phys.particle[i].num_processes);
    std::size_t const num_det = params->num_detectors();
    if (std::string(argv[2]) == "--replay")
        return replay(p, act, num_proc, argv[3], std::string(argv[5]) == "mixed");

    // Events round-robin over the workers
    std::vector<std::vector<Primary>> chunks(nthreads);
    for (Primary const& pr : prims)
        chunks[pr.event_id % nthreads].push_back(pr);
    std::vector<Worker> workers(nthreads);
    std::atomic<bool> done{false};
    std::thread watchdog([&] {
        std::vector<std::uint64_t> last(nthreads, ~0ull);
        std::vector<int> still(nthreads, 0);
        while (!done.load())
        {
            std::this_thread::sleep_for(std::chrono::milliseconds(250));
            for (int t = 0; t < nthreads; ++t)
            {
                std::uint64_t const q = workers[t].seq.load();
                if (q == last[t] && workers[t].in_call.load())
                {
                    if (++still[t] * 0.25 >= opt.stall_seconds)
                    {
                        print_snapshot(stdout, "stall", workers[t].current, 0);
                        std::fflush(stdout);
                        std::_Exit(3);
                    }
                }
                else
                {
                    still[t] = 0;
                }
                last[t] = q;
            }
        }
    });
    auto const t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
    {
        pool.emplace_back([&, t] {
            run_worker(p, act, chunks[t], nslots, num_proc, 0x1234567ull + 977 * t, num_det,
                       opt, workers[t]);
        });
    }
    for (auto& th : pool)
        th.join();
    double const secs
        = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    done = true;
    watchdog.join();

    // Summary
    std::uint64_t steps = 0, iters = 0, errored = 0, max_ts = 0, failures = 0;
    bool hit = false;
    std::vector<double> edep(num_det, 0.0);
    for (auto const& w : workers)
    {
        steps += w.track_steps;
        iters = std::max(iters, w.iterations);
        errored += w.errored;
        failures += w.sampling_failures;
        max_ts = std::max(max_ts, w.max_track_steps);
        hit = hit || w.hit_max_iterations;
        for (std::size_t d = 0; d < num_det; ++d)
            edep[d] += w.edep[d];
    }
    for (int ph = 0; ph < num_phases; ++ph)
    {
        std::uint64_t best = 0;
        int bw = 0;
        std::uint64_t calls = 0;
        for (int t = 0; t < nthreads; ++t)
        {
            calls += workers[t].calls[ph];
            if (workers[t].max_cycles[ph] > best)
            {
                best = workers[t].max_cycles[ph];
                bw = t;
            }
        }
        if (calls)
            print_snapshot(stdout, "longest", workers[bw].worst[ph], best);
    }
    std::printf("{\"summary\": true, \"mode\": \"%s\", \"primaries\": %zu, "
                "\"track_steps\": %" PRIu64 ", \"max_iterations\": %" PRIu64
                ", \"errored\": %" PRIu64 ", \"sampling_failures\": %" PRIu64
                ", \"max_steps_per_track\": %" PRIu64
                ", \"hit_iteration_limit\": %s, \"seconds\": %.3f, \"edep\": [",
                opt.mixed ? "mixed" : "fp64", prims.size(), steps, iters, errored, failures,
                max_ts,
                hit ? "true" : "false", secs);
    for (std::size_t d = 0; d < num_det; ++d)
        std::printf("%s%.17g", d ? ", " : "", edep[d]);
    std::printf("]}\n");
    return hit ? 4 : 0;
}
