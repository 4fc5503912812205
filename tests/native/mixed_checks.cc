// CPU checks of the mixed-precision building blocks that once stalled the
// full-scale transport: the FP32 uniform draw stays strictly inside (0, 1);
// the discrete selection never falls through to a process without a cross
// section; a sampler given an input outside its domain (pair production
// below threshold, whose Tsai-Urban loop can never accept) gives up after
// its iteration budget instead of spinning.
//
// Prints one "ok <name>" line per check; exits non-zero on the first
// failure.
#define CELER_MIXED_PRECISION 1
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "corecel/Precision.hh"
#include "em/StandardEm.hh"
#include "global/StepExecutors.hh"

using namespace celeritas_amd;

namespace
{
int failures = 0;

void check(bool ok, char const* name)
{
    std::printf("%s %s\n", ok ? "ok" : "FAIL", name);
    if (!ok)
        ++failures;
}

//! Engine returning a fixed word
struct FixedEngine
{
    std::uint32_t word;
    bool exhausted{false};
    std::uint32_t operator()() { return word; }
};
}  // namespace

int main()
{
    // (a) Draw range: the extreme words give 2^-24 and 1 - 2^-24 exactly
    {
        FixedEngine lo{0u}, hi{0xffffffffu};
        float const u0 = draw_canonical(lo);
        float const u1 = draw_canonical(hi);
        check(u0 > 0 && u0 == std::ldexp(1.0f, -24), "draw_lowest_word");
        check(u1 < 1 && u1 == 1 - std::ldexp(1.0f, -24), "draw_highest_word");
        // Every 23-bit value is below 1 (not only the extremes)
        bool all = true;
        for (std::uint32_t k = (1u << 23) - 4096; k < (1u << 23); ++k)
        {
            FixedEngine e{k << 9};
            float const u = draw_canonical(e);
            all = all && u > 0 && u < 1;
        }
        check(all, "draw_top_words_below_one");
    }

    // (b) Discrete selection with a sum of cross sections that cannot reach
    // the draw (a state where the first processes add to less than
    // macro_xs u): the fall-through skips a last process with no cross
    // section
    {
        ProcessRecord procs[3]{};
        for (auto& pr : procs)
        {
            pr.num_models = 1;
            pr.models[0].action = 0;
            pr.models[0].elem_cdf = invalid_id;
            pr.xs_grid = invalid_id;
        }
        procs[0].models[0].action = 10;
        procs[1].models[0].action = 11;
        procs[2].models[0].action = 12;
        ParticlePhysics part{};
        part.num_processes = 3;
        for (int i = 0; i < 3; ++i)
            part.process[i] = procs[i];
        std::uint32_t vol_to_mat[1] = {0};
        CoreParamsRef p{};
        p.phys.particle = &part;
        p.phys.num_particles = 1;
        p.volume_to_material = vol_to_mat;
        CoreStateRef s{};
        s.size = 1;
        double macro = 1000, xs[3] = {1.0, 2.0, 0.0}, mfp = 0;
        std::uint8_t pid = 0, post = 0, element = 0;
        std::uint32_t volume = 0;
        double energy = 0.04;
        std::uint32_t rngw[6] = {123456789u, 362436069u, 521288629u, 88675123u, 5783321u, 6615241u};
        s.phys.macro_xs = &macro;
        s.phys.per_process_xs = xs;
        s.phys.interaction_mfp = &mfp;
        s.phys.element = &element;
        s.particle.particle_id = &pid;
        s.particle.energy = &energy;
        s.geo.volume = &volume;
        s.sim.post_step_action = &post;
        for (int k = 0; k < 5; ++k)
            s.rng.xs[k] = rngw + k;
        s.rng.weyl = rngw + 5;
        exec::discrete_select_track(p, s, 0);
        check(post == 11, "discrete_fall_through_skips_zero_xs");
    }

    // (c) Tsai-Urban polar angle for a negative lepton kinetic energy (what
    // a pair-production sample below threshold produces): no u satisfies
    // u <= umax < 0, so the loop ends only through its budget
    {
        XorwowEngine rng{};
        rng.v[0] = 123456789u;
        rng.v[1] = 362436069u;
        rng.v[2] = 521288629u;
        rng.v[3] = 88675123u;
        rng.v[4] = 5783321u;
        rng.d = 6615241u;
        double const mu = sample_tsai_urban(-0.98, 0.51099895, rng);
        check(rng.exhausted, "tsai_urban_below_threshold_exhausts_budget");
        (void)mu;
        // A physical input accepts at once and leaves the engine usable
        XorwowEngine ok = rng;
        ok.exhausted = false;
        double const mu2 = sample_tsai_urban(10.0, 0.51099895, ok);
        check(!ok.exhausted && mu2 >= -1 && mu2 <= 1, "tsai_urban_physical_input");
    }
    return failures ? 1 : 0;
}
