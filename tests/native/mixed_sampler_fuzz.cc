// Mixed-precision samplers on the CPU: every interactor of every process
// (at energies where the process can be selected) and the charged
// along-step physics (MSC step limit, path conversions, scattering, energy
// loss with fluctuations), including NaN safety/step/direction inputs, are
// driven by a counting XORWOW engine that throws after 10^6 draws in one
// call (a rejection loop that does not end) under an alarm watchdog (a loop
// that does not draw). Outputs are checked for finite, in-range energies
// and unit directions. The HIP allocation calls are replaced by host
// memory, so the device tables are host arrays here.
// Usage: mixed_sampler_fuzz problem.json samples seed
#define CELER_MIXED_PRECISION 1
#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <unistd.h>
#include <hip/hip_runtime_api.h>
#include "corecel/Precision.hh"
#include "em/Interactors.hh"
#include "phys/AlongStepPhysics.hh"
#include "phys/PhysicsImpl.hh"
#include "phys/PhysicsParams.hh"
#include "global/ActionInterface.hh"
#include "corecel/Json.hh"
extern "C" {
hipError_t hipMalloc(void** p, size_t n) { *p = std::calloc(n ? n : 1, 1); return hipSuccess; }
hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
hipError_t hipMemset(void* d, int v, size_t n) { std::memset(d, v, n); return hipSuccess; }
const char* hipGetErrorString(hipError_t) { return "stub"; }
hipError_t hipGetLastError() { return hipSuccess; }
}
using namespace celeritas_amd;
struct Limit : std::runtime_error { using std::runtime_error::runtime_error; };
struct CountingEngine
{
    XorwowEngine e;
    unsigned long n{0}, limit{1000000};
    bool exhausted{false};  // set by a loop that ran out of its iteration budget
    std::uint32_t operator()() { if (++n > limit) throw Limit("draw limit"); return e(); }
};
static char g_case[512];
static void on_alarm(int) { std::fprintf(stderr, "WATCHDOG (no draws) in: %s\n", g_case); std::_Exit(3); }

int main(int argc, char** argv)
{
    std::ifstream f(argv[1]); std::stringstream ss; ss << f.rdbuf();
    long const nsamp = argc > 2 ? std::atol(argv[2]) : 100000;
    unsigned const seed = argc > 3 ? std::atoi(argv[3]) : 1;
    JsonValue doc = parse_json(ss.str());
    ActionRegistry reg;
    PhysicsParams phys(doc.at("physics"), &reg);
    PhysicsParamsRef const& p = phys.device_ref();
    // Host state for secondaries
    std::vector<Secondary> storage(64);
    CoreCounters counters{};
    CoreStateRef st{};
    st.secondaries.storage = storage.data();
    st.secondaries.capacity = storage.size();
    st.counters = &counters;
    std::mt19937_64 gen(seed);
    std::uniform_real_distribution<double> U(0, 1);
    auto logu = [&](double a, double b) { return a * std::exp(std::log(b / a) * U(gen)); };
    CountingEngine rng;
    rng.e.v[0] = 123456789u ^ seed; rng.e.v[1] = 362436069u; rng.e.v[2] = 521288629u; rng.e.v[3] = 88675123u; rng.e.v[4] = 5783321u; rng.e.d = 6615241u + seed;
    std::signal(SIGALRM, on_alarm);
    long fails = 0, calls = 0, nbad = 0, nfailed = 0;
    std::printf("particles %u materials %u\n", p.num_particles, p.num_materials);
    for (long it = 0; it < nsamp; ++it)
    {
        if (it % 20000 == 0) alarm(20);
        std::uint32_t const pid = gen() % p.num_particles;
        ParticlePhysics const& part = p.particle[pid];
        std::uint32_t const mat = gen() % p.num_materials;
        MaterialRecord const& m = p.materials[mat];
        double dir[3] = {0, 0, 1};
        // Interactors of every model of every process
        for (std::uint32_t pr = 0; pr < part.num_processes; ++pr)
        {
            ProcessRecord const& proc = part.process[pr];
            for (std::uint32_t k = 0; k < proc.num_models; ++k)
            {
                double const lo = std::max(proc.models[k].energy_low, 1e-3);
                double const hi = k + 1 < proc.num_models ? proc.models[k + 1].energy_low : 1e5;
                if (!(hi > lo)) continue;
                double const energy = logu(lo, hi);
                std::uint32_t const el = m.num_elements ? gen() % m.num_elements : 0;
                std::snprintf(g_case, sizeof g_case, "interact model=%u pid=%u mat=%u el=%u E=%.17g",
                              proc.models[k].model, pid, mat, el, energy);
                counters.secondary_stack_size = 0;
                InteractionContext ctx{p, st, 0, mat, el, pid};
                rng.n = 0; ++calls;
                if (proc.models[k].model == 2 && energy < 1.0219978) continue;
                if (!(phys::calc_xs(p, proc, mat, energy) > 0)) continue;
                try {
                    Interaction r = sample_interaction(static_cast<ModelClass>(proc.models[k].model), ctx, energy, dir, rng);
                    auto bad_dir = [](double const* d) { double n = d[0]*d[0]+d[1]*d[1]+d[2]*d[2]; return !(std::fabs(n - 1) < 1e-3); };
                    bool bad = r.action != Interaction::failed && (!std::isfinite(r.energy) || r.energy < 0 || r.energy > energy * (1 + 1e-6)
                               || (r.action == Interaction::scattered && r.energy > 0 && bad_dir(r.dir)) || !std::isfinite(r.energy_deposition));
                    for (std::uint32_t j = 0; r.action != Interaction::failed && j < r.num_secondaries; ++j)
                    {
                        Secondary const& sec = storage[r.sec_start + j];
                        if (sec.particle_id == invalid_id) continue;
                        if (!std::isfinite(sec.energy) || sec.energy < 0 || sec.energy > energy + 1.03 || bad_dir(sec.dir)) bad = true;
                    }
                    if (r.action == Interaction::failed) ++nfailed;
                    if (bad) { ++nbad; if (nbad < 20) { std::printf("BAD OUTPUT: %s -> action %d E %.9g dep %.9g nsec %u\n", g_case, int(r.action), r.energy, r.energy_deposition, r.num_secondaries);
                        for (std::uint32_t j = 0; j < r.num_secondaries; ++j) { Secondary const& q = storage[r.sec_start + j]; std::printf("   sec %u: pid %u E %.9g dir %.9g %.9g %.9g |d|^2-1 %.3g\n", j, q.particle_id, q.energy, q.dir[0], q.dir[1], q.dir[2], q.dir[0]*q.dir[0]+q.dir[1]*q.dir[1]+q.dir[2]*q.dir[2]-1); } } }
                }
                catch (Limit const&) { ++fails; std::printf("LOOP: %s\n", g_case); }
                if (rng.exhausted) { ++fails; rng.exhausted = false; std::printf("BUDGET: %s\n", g_case); }
            }
        }
        // Along-step physics of charged particles
        if (part.eloss_process < 0) continue;
        double const energy = logu(1e-3, 1e5);
        double const range = phys::calc_range(p.grids[part.range_grid + mat], p.grid_reals, energy);
        bool const lepton = pid == p.electron_id || pid == p.positron_id;
        double step = range * logu(1e-6, 1.0);
        double safety = U(gen) < 0.2 ? 0.0 : logu(1e-7, 5.0);
        double const nanmode = U(gen);
        if (nanmode < 0.1) safety = NAN;
        else if (nanmode < 0.2) step = NAN;
        else if (nanmode < 0.3) { dir[0] = dir[1] = dir[2] = NAN; }
        std::snprintf(g_case, sizeof g_case, "along pid=%u mat=%u E=%.17g range=%.17g step=%.17g safety=%.17g",
                      pid, mat, energy, range, step, safety);
        rng.n = 0; ++calls;
        try
        {
            double true_path = step, geom_path = step, alpha = 0;
            bool msc = lepton && p.msc.enabled && energy > p.msc.low_energy_limit && energy < p.msc.high_energy_limit;
            alongstep::MscTrack mt{p, part, mat, pid == p.electron_id ? 0u : 1u, energy, range, 0};
            alongstep::MscRange mr{0, 0, 0};
            if (msc)
            {
                mt.lambda = mt.calc_msc_mfp(energy);
                true_path = alongstep::step_limit(mt, mr, safety == 0, safety, step, rng);
                geom_path = alongstep::to_geo(mt, true_path, &alpha);
                if (geom_path > mt.lambda) geom_path = mt.lambda;
                if (U(gen) < 0.3)
                {
                    double g2 = geom_path * U(gen);
                    true_path = alongstep::from_geo(mt, true_path, alpha, g2);
                    geom_path = g2;
                }
                if (geom_path > 0)
                {
                    double sf = U(gen) < 0.5 ? safety : 0;
                    auto res = alongstep::msc_scatter(mt, m, pid == p.positron_id, dir, sf, sf > 0, geom_path,
                                           true_path, mr.limit_min, rng);
                    double n = res.dir[0]*res.dir[0]+res.dir[1]*res.dir[1]+res.dir[2]*res.dir[2];
                    bool bad = !(std::fabs(n - 1) < 1e-3) || !std::isfinite(res.displacement[0] + res.displacement[1] + res.displacement[2]);
                    if (!(true_path >= 0 && true_path <= step * (1 + 1e-9)) || !(geom_path >= 0)) bad = true;
                    if (nanmode < 0.3) bad = false;
                    if (bad) { ++nbad; if (nbad < 20) std::printf("BAD MSC: %s -> true %.9g geom %.9g |dir|^2 %.9g action %d\n", g_case, true_path, geom_path, n, res.action); }
                }
            }
            double dep = alongstep::calc_eloss(p, part, pid, mat, energy, range, true_path, U(gen) < 0.5, rng);
            if (nanmode >= 0.3 && !(dep >= 0 && dep <= energy)) { ++nbad; if (nbad < 20) std::printf("BAD ELOSS: %s -> %.9g\n", g_case, dep); }
        }
        catch (Limit const&) { ++fails; std::printf("LOOP: %s\n", g_case); }
        if (rng.exhausted) { ++fails; rng.exhausted = false; std::printf("BUDGET: %s\n", g_case); }
    }
    std::printf("calls %ld loops %ld bad %ld failed %ld\n", calls, fails, nbad, nfailed);
    return fails != 0;
}
