//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/capi.cc
//! C entry points of liboracle.so for the Python test harness.
//---------------------------------------------------------------------------//
#include <algorithm>
#include <cstring>
#include <string>

#include "core.hh"
#include "physics.hh"
#include "user.hh"
#include "problem.hh"
#include "rng.hh"
#include "loop.hh"
#include <stdexcept>

using namespace orc;

namespace
{
thread_local std::string g_err;

template<class F>
int guard(F&& f)
{
    try
    {
        g_err.clear();
        f();
        return 0;
    }
    catch (std::exception const& e)
    {
        g_err = e.what();
    }
    return 1;
}

struct OrcProblem
{
    std::shared_ptr<Problem> p;
};
struct OrcState
{
    std::unique_ptr<State> s;
};

void load(XorwowState& s, std::uint32_t const* w)
{
    for (int i = 0; i < 5; ++i)
        s.x[i] = w[i];
    s.d = w[5];
}
void save(XorwowState const& s, std::uint32_t* w)
{
    for (int i = 0; i < 5; ++i)
        w[i] = s.x[i];
    w[5] = s.d;
}
}  // namespace

extern "C" {

const char* orc_last_error()
{
    return g_err.c_str();
}

//---------------------------------------------------------------------------//
// RNG
void orc_rng_seeded_states(std::uint32_t seed, std::uint32_t stream,
                           std::uint32_t n, std::uint32_t* out)
{
    auto st = seeded_states(seed, stream, n);
    for (std::uint32_t i = 0; i < n; ++i)
        save(st[i], out + 6 * i);
}

void orc_rng_initialize(std::uint32_t seed, unsigned long long subsequence,
                        unsigned long long offset, std::uint32_t* out6)
{
    XorwowState s;
    xorwow_initialize(s, seed, subsequence, offset);
    save(s, out6);
}

void orc_rng_discard(std::uint32_t* state6, unsigned long long count)
{
    XorwowState s;
    load(s, state6);
    xorwow_discard(s, count);
    save(s, state6);
}

//! kind 0 = u32, 1 = double, 2 = float
void orc_rng_draw(std::uint32_t* state6, std::uint32_t n, int kind, void* out)
{
    XorwowState s;
    load(s, state6);
    for (std::uint32_t i = 0; i < n; ++i)
    {
        if (kind == 0)
            static_cast<std::uint32_t*>(out)[i] = xorwow_next(s);
        else if (kind == 1)
            static_cast<double*>(out)[i] = canonical_double(s);
        else
            static_cast<float*>(out)[i] = canonical_float(s);
    }
    save(s, state6);
}

//---------------------------------------------------------------------------//
// PROBLEM AND STATE
void* orc_problem_create(const char* json)
{
    OrcProblem* r = nullptr;
    guard([&] { r = new OrcProblem{Problem::from_json(json)}; });
    return r;
}

void orc_problem_destroy(void* p)
{
    delete static_cast<OrcProblem*>(p);
}

int orc_problem_action_label(void* p, std::uint32_t id, char* buf, std::size_t len)
{
    return guard([&] {
        auto const& labels = static_cast<OrcProblem*>(p)->p->actions.labels;
        std::snprintf(buf, len, "%s", labels.at(id).c_str());
    });
}

std::uint32_t orc_problem_num_actions(void* p)
{
    return static_cast<std::uint32_t>(static_cast<OrcProblem*>(p)->p->actions.labels.size());
}

void* orc_state_create(void* problem, std::uint32_t num_slots, std::uint32_t stream)
{
    OrcState* r = nullptr;
    guard([&] { r = new OrcState{make_state(static_cast<OrcProblem*>(problem)->p,
                                            num_slots, stream)}; });
    return r;
}

void orc_state_destroy(void* s)
{
    delete static_cast<OrcState*>(s);
}

//! Primaries in the C ABI layout (particle, event, energy, pos, dir, time)
struct CPrimary
{
    std::uint32_t particle;
    std::uint32_t event;
    double energy;
    double pos[3];
    double dir[3];
    double time;
};

int orc_state_insert_primaries(void* s, CPrimary const* prim, std::uint32_t n)
{
    return guard([&] {
        std::vector<PrimaryP> v;
        for (std::uint32_t i = 0; i < n; ++i)
        {
            PrimaryP p;
            p.particle = prim[i].particle;
            p.event = prim[i].event;
            p.energy = prim[i].energy;
            for (int k = 0; k < 3; ++k)
            {
                p.pos[k] = prim[i].pos[k];
                p.dir[k] = prim[i].dir[k];
            }
            p.time = prim[i].time;
            v.push_back(p);
        }
        insert_primaries(*static_cast<OrcState*>(s)->s, v);
    });
}

//! Run an action by label
int orc_state_action(void* s, const char* label)
{
    return guard([&] { run_action(*static_cast<OrcState*>(s)->s, label); });
}

int orc_state_mock_interact(void* s, std::uint32_t const* nsec, std::uint8_t const* alive)
{
    return guard([&] {
        State& st = *static_cast<OrcState*>(s)->s;
        std::size_t n = st.tracks.size();
        mock_interact(st, std::vector<std::uint32_t>(nsec, nsec + n),
                      std::vector<char>(alive, alive + n));
    });
}

//! Counters: initializers, vacancies, active, secondaries, alive, generated
void orc_state_counters(void* s, std::uint32_t* out)
{
    Counters const& c = static_cast<OrcState*>(s)->s->counters;
    out[0] = c.num_initializers;
    out[1] = c.num_vacancies;
    out[2] = c.num_active;
    out[3] = c.num_secondaries;
    out[4] = c.num_alive;
    out[5] = c.num_generated;
}

//! Integer track fields by name; returns number written
std::uint32_t orc_state_get_u32(void* s, const char* name, std::uint32_t* out,
                                std::uint32_t max)
{
    State const& st = *static_cast<OrcState*>(s)->s;
    std::string f(name);
    std::uint32_t n = 0;
    auto put = [&](std::uint32_t v) {
        if (n < max)
            out[n] = v;
        ++n;
    };
    if (f == "vacancies")
        for (std::uint32_t i = 0; i < st.counters.num_vacancies; ++i)
            put(st.vacancies[i]);
    else if (f == "init_ids")
        for (std::uint32_t i = 0; i < st.counters.num_initializers; ++i)
            put(st.initializers[i].track_id);
    else
        for (auto const& t : st.tracks)
        {
            if (f == "track_id") put(t.track_id);
            else if (f == "parent_id") put(t.parent_id);
            else if (f == "event_id") put(t.event_id);
            else if (f == "status") put(t.status);
            else if (f == "particle") put(t.particle);
            else if (f == "volume") put(t.geo.vol);
            else if (f == "post_action") put(t.post_action);
            else if (f == "along_action") put(t.along_action);
            else if (f == "num_steps") put(t.num_steps);
            else if (f == "rng")
                for (int k = 0; k < 6; ++k)
                    put(k < 5 ? t.rng.x[k] : t.rng.d);
        }
    return n;
}

//! Real track fields by name; returns number written
std::uint32_t orc_state_get_f64(void* s, const char* name, double* out,
                                std::uint32_t max)
{
    State const& st = *static_cast<OrcState*>(s)->s;
    std::string f(name);
    std::uint32_t n = 0;
    auto put = [&](double v) {
        if (n < max)
            out[n] = v;
        ++n;
    };
    if (f == "calo_edep")
    {
        for (double v : st.calo_edep)
            put(v);
        return n;
    }
    if (f == "total_edep")
    {
        put(st.total_edep);
        return n;
    }
    for (auto const& t : st.tracks)
    {
        if (f == "energy") put(t.energy);
        else if (f == "pos") for (double v : t.geo.pos) put(v);
        else if (f == "dir") for (double v : t.geo.dir) put(v);
        else if (f == "time") put(t.time);
        else if (f == "step_length") put(t.step_length);
        else if (f == "interaction_mfp") put(t.interaction_mfp);
        else if (f == "edep") put(t.edep);
    }
    return n;
}

//! Set RNG words of all slots (6 per slot)
void orc_state_set_rng(void* s, std::uint32_t const* words)
{
    State& st = *static_cast<OrcState*>(s)->s;
    for (std::size_t i = 0; i < st.tracks.size(); ++i)
        load(st.tracks[i].rng, words + 6 * i);
}

//---------------------------------------------------------------------------//
// GEOMETRY
//! Trace a ray: writes volume IDs and distances until leaving the world
std::uint32_t orc_geo_trace(void* problem, double const* pos, double const* dir,
                            std::uint32_t max, std::uint32_t* vols, double* dists,
                            double* safeties)
{
    Geometry const& g = static_cast<OrcProblem*>(problem)->p->geo;
    GeoState st;
    geo_initialize(g, st, pos, dir);
    std::uint32_t n = 0;
    while (n < max && !geo_is_outside(st) && !st.failed)
    {
        bool bnd;
        double d = geo_find_next_step(g, st, 0, &bnd);
        vols[n] = st.vol;
        dists[n] = d;
        if (safeties)
        {
            // Safety at the midpoint of the segment
            GeoState mid = st;
            geo_move_internal(mid, d / 2);
            safeties[n] = geo_find_safety(g, mid);
        }
        ++n;
        if (!bnd)
            break;
        geo_move_to_boundary(st);
        geo_cross_boundary(g, st);
    }
    return n;
}

//! Volume of each point (pos[3], dir[3]) by exhaustive search, or
//! 0xffffffff where initialization fails
void orc_geo_locate(void* problem, double const* points, std::uint32_t n,
                    std::uint32_t* out)
{
    Geometry const& g = static_cast<OrcProblem*>(problem)->p->geo;
    for (std::uint32_t i = 0; i < n; ++i)
    {
        GeoState st;
        geo_initialize(g, st, points + 6 * i, points + 6 * i + 3);
        out[i] = st.failed ? 0xffffffffu : st.vol;
    }
}

}  // extern "C"

extern "C" {
//! GenerateCanonical32 applied to a fixed word sequence (SequenceEngine)
void orc_canonical_from_words(std::uint32_t const* words, std::uint32_t n_out,
                              int kind, void* out)
{
    std::uint32_t pos = 0;
    for (std::uint32_t i = 0; i < n_out; ++i)
    {
        if (kind == 1)
        {
            unsigned long long upper = words[pos++];
            unsigned long long lower = words[pos++];
            static_cast<double*>(out)[i]
                = static_cast<double>((upper << 21) ^ lower) * 0x1p-53;
        }
        else
        {
            static_cast<float*>(out)[i] = 0x1p-32f * static_cast<float>(words[pos++]);
        }
    }
}
}

extern "C" {
//! Draw from many independent states (n x 6 words, updated in place)
void orc_rng_draw_batch(std::uint32_t* states, std::uint32_t n, std::uint32_t draws,
                        int kind, void* out)
{
#pragma omp parallel for schedule(static)
    for (std::uint32_t i = 0; i < n; ++i)
    {
        std::size_t base = std::size_t(i) * draws;
        if (kind == 0)
            orc_rng_draw(states + 6 * std::size_t(i), draws, 0,
                         static_cast<std::uint32_t*>(out) + base);
        else if (kind == 1)
            orc_rng_draw(states + 6 * std::size_t(i), draws, 1,
                         static_cast<double*>(out) + base);
        else
            orc_rng_draw(states + 6 * std::size_t(i), draws, 2,
                         static_cast<float*>(out) + base);
    }
}

//! Initialize many states with (subsequence, offset) pairs
void orc_rng_initialize_batch(std::uint32_t seed, std::uint32_t n,
                              unsigned long long const* sub,
                              unsigned long long const* off, std::uint32_t* out)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (std::uint32_t i = 0; i < n; ++i)
        orc_rng_initialize(seed, sub[i], off[i], out + 6 * std::size_t(i));
}

//! Initializer records (same layout as the product's C ABI struct)
std::uint32_t orc_state_initializers(void* s, void* out, std::uint32_t max)
{
    struct Rec
    {
        std::uint32_t track_id, parent_id, event_id, particle;
        double time, pos[3], dir[3], energy;
    };
    State const& st = *static_cast<OrcState*>(s)->s;
    std::uint32_t n = st.counters.num_initializers;
    Rec* r = static_cast<Rec*>(out);
    for (std::uint32_t i = 0; i < n && i < max; ++i)
    {
        Initializer const& ti = st.initializers[i];
        r[i].track_id = ti.track_id;
        r[i].parent_id = ti.parent_id;
        r[i].event_id = ti.event_id;
        r[i].particle = ti.particle;
        r[i].time = ti.time;
        r[i].energy = ti.energy;
        for (int k = 0; k < 3; ++k)
        {
            r[i].pos[k] = ti.pos[k];
            r[i].dir[k] = ti.dir[k];
        }
    }
    return n;
}
}

extern "C" {
//! Run step iterations like StepperTestBase::run: first step with the
//! primaries, then until empty or max_iters. Writes active/queued per step.
std::uint32_t orc_run(void* state, void const* prim, std::uint32_t nprim,
                      std::uint32_t max_iters, std::uint32_t* active,
                      std::uint32_t* queued)
{
    std::uint32_t n = 0;
    int rc = guard([&] {
        State& st = *static_cast<OrcState*>(state)->s;
        orc_state_insert_primaries(state, static_cast<CPrimary const*>(prim), nprim);
        if (!g_err.empty())
            throw std::runtime_error(g_err);
        StepResult r = step_once(st);
        active[n] = r.active;
        queued[n] = r.queued;
        ++n;
        while ((r.queued > 0 || r.alive > 0) && n < max_iters)
        {
            r = step_once(st);
            active[n] = r.active;
            queued[n] = r.queued;
            ++n;
        }
    });
    return rc == 0 ? n : 0;
}
}

extern "C" {
//! One step iteration (pending primaries are consumed)
int orc_step_once(void* state, std::uint32_t* out4)
{
    return guard([&] {
        StepResult r = step_once(*static_cast<OrcState*>(state)->s);
        out4[0] = r.generated;
        out4[1] = r.queued;
        out4[2] = r.active;
        out4[3] = r.alive;
    });
}
}

extern "C" {
//! Per-event track counters
std::uint32_t orc_state_track_counters(void* s, std::uint32_t* out, std::uint32_t max)
{
    State const& st = *static_cast<OrcState*>(s)->s;
    std::uint32_t n = 0;
    for (auto v : st.track_counters)
    {
        if (n < max)
            out[n] = v;
        ++n;
    }
    return n;
}
}

#include "alongstep.hh"
#include "em.hh"

extern "C" {
//! Sample interactions with one model. Per sample, 17 doubles: action,
//! energy, dir[3], edep, nsec, then (particle, energy, dir[3]) x 2.
//! rng_kind 0: default-seeded std::mt19937 through std::generate_canonical
//! (returns the 32-bit draw count in *count); 1: replay 'uniforms' with
//! 'stride' values per sample.
int orc_interactor_sample(void* problem, const char* model, std::uint32_t particle,
                          double energy, double const* dir, std::uint32_t material,
                          std::uint32_t element, std::uint32_t n, std::uint32_t stack,
                          int rng_kind, double const* uniforms, std::uint32_t stride,
                          double* out, unsigned long long* count)
{
    return guard([&] {
        auto prob = static_cast<OrcProblem*>(problem)->p;
        auto st = make_state(prob, 1, 0);
        st->stack.assign(stack, SecondaryP{});
        InteractorInput in;
        in.pd = prob->physics.get();
        in.particle = particle;
        in.energy = energy;
        for (int i = 0; i < 3; ++i)
            in.dir[i] = dir[i];
        in.material = material;
        in.element = element;
        std::mt19937 mt;
        MtUniform mtu{mt};
        for (std::uint32_t k = 0; k < n; ++k)
        {
            Interaction r;
            if (rng_kind == 0)
            {
                r = sample_model(model, in, *st, mtu);
            }
            else
            {
                ReplayUniform rep{uniforms + std::size_t(k) * stride, 0, stride};
                r = sample_model(model, in, *st, rep);
            }
            double* o = out + 17 * std::size_t(k);
            o[0] = r.action;
            o[1] = r.energy;
            for (int i = 0; i < 3; ++i)
                o[2 + i] = r.dir[i];
            o[5] = r.energy_deposition;
            o[6] = r.num_secondaries;
            for (std::uint32_t j = 0; j < 2; ++j)
            {
                double* so = o + 7 + 5 * j;
                if (j < r.num_secondaries && r.action != Interaction::failed)
                {
                    SecondaryP const& sec = st->stack[r.sec_start + j];
                    so[0] = sec.particle == none ? -1.0 : double(sec.particle);
                    so[1] = sec.energy;
                    for (int i = 0; i < 3; ++i)
                        so[2 + i] = sec.dir[i];
                }
                else
                {
                    for (int i = 0; i < 5; ++i)
                        so[i] = 0;
                }
            }
        }
        if (count)
            *count = mtu.count;
    });
}

//! Uniform doubles from a seeded std::mt19937 (std::generate_canonical)
void orc_mt_uniforms(std::uint32_t seed, std::uint64_t n, double* out)
{
    std::mt19937 mt(seed);
    MtUniform u{mt};
    for (std::uint64_t i = 0; i < n; ++i)
        out[i] = u();
}

//---------------------------------------------------------------------------//
// Bremsstrahlung pieces (SeltzerBerger.test.cc, RelativisticBrem.test.cc)
//! A default-seeded std::mt19937 that persists across calls
void* orc_mt_create() { return new std::mt19937; }
void orc_mt_destroy(void* mt) { delete static_cast<std::mt19937*>(mt); }

int orc_sb_argmax(void* problem, std::uint32_t element, std::uint32_t* out,
                  std::uint32_t n)
{
    return guard([&] {
        auto const& t = static_cast<OrcProblem*>(problem)->p->physics->elements.at(element).sb;
        if (t.argmax.size() != n)
            throw std::runtime_error("argmax size mismatch");
        std::copy(t.argmax.begin(), t.argmax.end(), out);
    });
}

double orc_sb_positron_scale(int z, double mass, double kc, double inc, double exit)
{
    return sb_positron_scale(z, mass, kc, inc, exit);
}

//! Sample n SB photon energies on a persistent engine: out = {max_xs, mean
//! exit fraction, mean 32-bit draws per sample}
int orc_sb_sample_many(void* problem, void* mt, std::uint32_t material,
                       std::uint32_t element, double energy, double mass,
                       double kc, int corrector, std::uint32_t n, double* out)
{
    return guard([&] {
        auto const& pd = *static_cast<OrcProblem*>(problem)->p->physics;
        MaterialDef const& mat = pd.materials.at(material);
        ElementDef const& el = pd.elements.at(mat.elements.at(element).first);
        MtUniform u{*static_cast<std::mt19937*>(mt)};
        double total = 0;
        double max_xs = 0;
        for (std::uint32_t i = 0; i < n; ++i)
            total += sample_sb_energy(el, mat.electron_density, energy, mass, kc,
                                      corrector, u, &max_xs);
        out[0] = max_xs;
        out[1] = total / (n * energy);
        out[2] = double(u.count) / n;
    });
}

//! Run n interactions on a persistent engine with a stack of n secondaries
//! per call (the reference's stress tests); returns the 32-bit draw count
int orc_interactor_stress(void* problem, void* mt, const char* model,
                          std::uint32_t particle, double energy, double const* dir,
                          std::uint32_t material, std::uint32_t n,
                          unsigned long long* count, double* sums)
{
    return guard([&] {
        auto prob = static_cast<OrcProblem*>(problem)->p;
        auto st = make_state(prob, 1, 0);
        st->stack.assign(std::max<std::uint32_t>(n, 1) * 2, SecondaryP{});
        InteractorInput in;
        in.pd = prob->physics.get();
        in.particle = particle;
        in.energy = energy;
        for (int i = 0; i < 3; ++i)
            in.dir[i] = dir[i];
        make_unit(in.dir);
        in.material = material;
        MtUniform u{*static_cast<std::mt19937*>(mt)};
        // sums: secondary count, energy, cosine to the (unnormalized) dir
        for (std::uint32_t k = 0; k < n; ++k)
        {
            Interaction r = sample_model(model, in, *st, u);
            if (r.action == Interaction::failed)
                throw std::runtime_error("interaction failed in stress loop");
            for (std::uint32_t j = 0; sums && j < r.num_secondaries; ++j)
            {
                SecondaryP const& sec = st->stack[r.sec_start + j];
                sums[0] += 1;
                sums[1] += sec.energy;
                sums[2] += dir[0] * sec.dir[0] + dir[1] * sec.dir[1] + dir[2] * sec.dir[2];
            }
        }
        *count = u.count;
    });
}

//! EnergyLossHelper + EnergyLossUrbanDistribution for an electron: n
//! samples on a default-seeded mt19937; out = {max_energy, beta_sq, mean
//! loss, 32-bit draws}
int orc_urban_fluct(void* problem, std::uint32_t material, double energy, double mean_loss,
                    std::uint32_t n, double* out)
{
    return guard([&] {
        auto const& pd = *static_cast<OrcProblem*>(problem)->p->physics;
        if (!pd.fluct)
            throw std::runtime_error("fluctuations are off");
        double const me = pd.particles.at(pd.electron).mass;
        double const gamma = 1 + energy / me;
        double const ig = me / (energy + me);
        double const beta_sq = 1 - ig * ig;
        double const two_mebsgs = 2 * me * beta_sq * (gamma * gamma);
        double const max_energy
            = std::min(pd.cutoff_energy[pd.electron][material], 0.5 * energy);
        std::mt19937 mt;
        MtUniform u{mt};
        double sum = 0;
        for (std::uint32_t i = 0; i < n; ++i)
            sum += urban_fluct_sample(pd.fluct_params.at(material), pd.materials.at(material),
                                      mean_loss, max_energy, two_mebsgs, beta_sq, u);
        out[0] = max_energy;
        out[1] = beta_sq;
        out[2] = sum / n;
        out[3] = double(u.count);
    });
}

double orc_pe_macro_xs(void* problem, std::uint32_t material, double energy)
{
    auto const& pd = *static_cast<OrcProblem*>(problem)->p->physics;
    return livermore_pe_macro_xs(pd, material, energy);
}

double orc_rb_dxsec(void* problem, std::uint32_t element, double total_energy,
                    double gamma_energy)
{
    auto const& pd = *static_cast<OrcProblem*>(problem)->p->physics;
    return rb_dxsec(pd.elements.at(element), total_energy, gamma_energy);
}

//! Relativistic DCS with the LPM switch of RBDiffXsCalculator (the element
//! is a component of the material)
double orc_rb_dxsec_lpm(void* problem, std::uint32_t material, std::uint32_t component,
                        int enable_lpm, double total_energy, double gamma_energy)
{
    auto const& pd = *static_cast<OrcProblem*>(problem)->p->physics;
    MaterialDef const& mat = pd.materials.at(material);
    ElementDef const& el = pd.elements.at(mat.elements.at(component).first);
    return RbDxsec(el, mat, enable_lpm != 0, total_energy)(gamma_energy);
}

//! Plant a deliberate mistake in the LPM functions (0: none)
void orc_lpm_set_mutation(int m)
{
    orc::g_lpm_mutation = m;
}

//! Material radiation length [cm] as the physics data holds it
double orc_material_rad_length(void* problem, std::uint32_t material)
{
    auto const& pd = *static_cast<OrcProblem*>(problem)->p->physics;
    return pd.materials.at(material).rad_length;
}

//! LPM functions (xi, G, phi) for a material component
void orc_lpm_functions(void* problem, std::uint32_t material, std::uint32_t component,
                       int dielectric, double gamma_energy, double epsilon, double* out)
{
    auto const& pd = *static_cast<OrcProblem*>(problem)->p->physics;
    MaterialDef const& mat = pd.materials.at(material);
    ElementDef const& el = pd.elements.at(mat.elements.at(component).first);
    LpmFunctions const f = lpm_functions(el, mat, dielectric != 0, gamma_energy, epsilon);
    out[0] = f.xi;
    out[1] = f.g;
    out[2] = f.phi;
}
}

extern "C" {
//! Sample a sequence of items from ONE default-seeded std::mt19937 stream.
//! Item i: model name (NUL-separated list), particle, energy, dir[3],
//! material, element, repeat count, stack reset capacity (0: keep stack).
//! Writes 17 doubles per sample; returns the total 32-bit draw count.
int orc_interactor_sequence(void* problem, const char* models, std::uint32_t nitems,
                            std::uint32_t const* particle, double const* energy,
                            double const* dir, std::uint32_t const* material,
                            std::uint32_t const* element, std::uint32_t const* repeat,
                            std::uint32_t const* stack_reset, double* out,
                            unsigned long long* count)
{
    return guard([&] {
        auto prob = static_cast<OrcProblem*>(problem)->p;
        auto st = make_state(prob, 1, 0);
        std::mt19937 mt;
        MtUniform mtu{mt};
        char const* name = models;
        std::size_t k = 0;
        for (std::uint32_t i = 0; i < nitems; ++i)
        {
            std::string model(name);
            name += model.size() + 1;
            if (stack_reset[i] > 0)
            {
                st->stack.assign(stack_reset[i], SecondaryP{});
                st->stack_size = 0;
            }
            InteractorInput in;
            in.pd = prob->physics.get();
            in.particle = particle[i];
            in.energy = energy[i];
            for (int d = 0; d < 3; ++d)
                in.dir[d] = dir[3 * i + d];
            in.material = material[i];
            in.element = element[i];
            for (std::uint32_t r = 0; r < repeat[i]; ++r, ++k)
            {
                Interaction res = sample_model(model, in, *st, mtu);
                double* o = out + 17 * k;
                o[0] = res.action;
                o[1] = res.energy;
                for (int d = 0; d < 3; ++d)
                    o[2 + d] = res.dir[d];
                o[5] = res.energy_deposition;
                o[6] = res.num_secondaries;
                for (std::uint32_t j = 0; j < 2; ++j)
                {
                    double* so = o + 7 + 5 * j;
                    bool ok = j < res.num_secondaries && res.action != Interaction::failed;
                    SecondaryP const* sec = ok ? &st->stack[res.sec_start + j] : nullptr;
                    so[0] = (!sec || sec->particle == none) ? -1.0 : double(sec->particle);
                    so[1] = sec ? sec->energy : 0;
                    for (int d = 0; d < 3; ++d)
                        so[2 + d] = sec ? sec->dir[d] : 0;
                }
            }
        }
        *count = mtu.count;
    });
}
}

extern "C" {
//! Replay interactors sample by sample with per-sample energies/directions
int orc_interactor_replay_batch(void* problem, const char* model, std::uint32_t particle,
                                double const* energies, double const* dirs,
                                std::uint32_t material, std::uint32_t element,
                                std::uint32_t n, double const* uniforms,
                                std::uint32_t stride, double* out)
{
    return guard([&] {
        auto prob = static_cast<OrcProblem*>(problem)->p;
        auto st = make_state(prob, 1, 0);
        st->stack.assign(std::size_t(2) * n + 2, SecondaryP{});
        InteractorInput in;
        in.pd = prob->physics.get();
        in.particle = particle;
        in.material = material;
        in.element = element;
        for (std::uint32_t k = 0; k < n; ++k)
        {
            in.energy = energies[k];
            for (int i = 0; i < 3; ++i)
                in.dir[i] = dirs[3 * k + i];
            ReplayUniform rep{uniforms + std::size_t(k) * stride, 0, stride};
            Interaction r = sample_model(model, in, *st, rep);
            double* o = out + 17 * std::size_t(k);
            o[0] = r.action;
            o[1] = r.energy;
            for (int i = 0; i < 3; ++i)
                o[2 + i] = r.dir[i];
            o[5] = r.energy_deposition;
            o[6] = r.num_secondaries;
            for (std::uint32_t j = 0; j < 2; ++j)
            {
                double* so = o + 7 + 5 * j;
                bool ok = j < r.num_secondaries && r.action != Interaction::failed;
                SecondaryP const* sec = ok ? &st->stack[r.sec_start + j] : nullptr;
                so[0] = (!sec || sec->particle == none) ? -1.0 : double(sec->particle);
                so[1] = sec ? sec->energy : 0;
                for (int i = 0; i < 3; ++i)
                    so[2 + i] = sec ? sec->dir[i] : 0;
            }
        }
    });
}
}

//---------------------------------------------------------------------------//
// FIELD PROPAGATION
//---------------------------------------------------------------------------//
#include "field.hh"

extern "C" {
/*!
 * Propagate one charged track through the problem geometry in a uniform
 * field with successive step requests through one propagator object, as
 * the reference's unit tests do; a request of 0 crosses the boundary the
 * track is on and -1 only starts a new propagator. Per
 * request, out holds 11 values: distance, boundary, looping, position (3),
 * direction (3), volume, and the cumulative number of Dormand-Prince steps.
 */
int orc_field_propagate(void* problem, double const* pos, double const* dir,
                        double momentum, double charge, double const* field,
                        int max_substeps, double const* steps, std::uint32_t nsteps,
                        double* out)
{
    return guard([&] {
        Geometry const& g = static_cast<OrcProblem*>(problem)->p->geo;
        GeoState s;
        geo_initialize(g, s, pos, dir);
        UniformFieldStepper stepper;
        for (int i = 0; i < 3; ++i)
            stepper.field[i] = field[i];
        stepper.coeff = charge * lorentz_per_cm_tesla;
        FieldOptions opts;
        if (max_substeps > 0)
            opts.max_substeps = max_substeps;
        FieldPropagatorState ps = make_field_state(s, momentum);
        for (std::uint32_t k = 0; k < nsteps; ++k)
        {
            FieldPropagation r;
            if (steps[k] > 0)
            {
                r = field_propagate(g, s, stepper, opts, ps, steps[k]);
            }
            else
            {
                // New propagator (after crossing the boundary for 0)
                if (steps[k] == 0)
                    geo_cross_boundary(g, s);
                ps = make_field_state(s, momentum);
            }
            double* o = out + 11 * k;
            o[0] = r.distance;
            o[1] = r.boundary;
            o[2] = r.looping;
            for (int i = 0; i < 3; ++i)
            {
                o[3 + i] = s.pos[i];
                o[6 + i] = s.dir[i];
            }
            o[9] = s.vol;
            o[10] = static_cast<double>(stepper.count);
        }
    });
}
}

//---------------------------------------------------------------------------//
// URBAN MSC PIECES (msc.hh)
//---------------------------------------------------------------------------//
#include "msc.hh"

extern "C" {
//! Material coefficients for an effective Z: out = {stepmin[2], theta[2],
//! tail[3], tail_corr, scaled_zeff[2], d_over_r}
void orc_msc_material_coeffs(double zeff, double* out)
{
    MscDataO::Material const d = msc_material_coeffs(zeff);
    double const v[] = {d.stepmin[0], d.stepmin[1], d.theta[0], d.theta[1],
                        d.tail[0],    d.tail[1],    d.tail[2],  d.tail_corr,
                        d.scaled_zeff[0], d.scaled_zeff[1], d.d_over_r};
    std::copy(std::begin(v), std::end(v), out);
}

/*!
 * The step-limit sampling of UrbanMsc.test.cc:296-371 for an electron in a
 * material of effective Z: per energy i (with its range and MSC MFP), a
 * fresh MSC range, then nsamples limits of a physical step equal to the
 * range, on a boundary or not, with zero safety; the engine (from
 * orc_mt_create) continues across calls. out per energy: mean step,
 * range_init, range_factor, limit_min.
 */
int orc_msc_step_limit_sample(void* mt, int alg, int on_boundary, double zeff,
                              double range_factor, double const* energy,
                              double const* range, double const* lambda, std::uint32_t n,
                              std::uint32_t nsamples, double* out)
{
    return guard([&] {
        if (alg < 0 || alg > 2)
            throw std::runtime_error("unknown MSC step limit algorithm");
        MscDataO::Material const md = msc_material_coeffs(zeff);
        MtUniform u{*static_cast<std::mt19937*>(mt)};
        for (std::uint32_t i = 0; i < n; ++i)
        {
            MscLimitInput in;
            in.alg = static_cast<MscLimitAlg>(alg);
            in.energy = energy[i];
            in.range = range[i];
            in.lambda = lambda[i];
            in.md = &md;
            in.par = 0;
            in.range_factor = range_factor;
            in.on_boundary = on_boundary != 0;
            in.safety = 0;
            in.phys_step = range[i];
            MscRangeO r;
            double sum = 0;
            for (std::uint32_t k = 0; k < nsamples; ++k)
                sum += msc_step_limit(in, r, u);
            out[4 * i] = sum / nsamples;
            out[4 * i + 1] = r.range_init;
            out[4 * i + 2] = r.range_factor;
            out[4 * i + 3] = r.limit_min;
        }
    });
}

double orc_urban_positron_correction(double zeff, double y)
{
    return urban_positron_correction(zeff, y);
}

//! MSC helper quantities and step conversions of an e- (particle 0) or e+
//! (1) at an energy in a material: out_info = {range, lambda, geometric
//! path of tstep, alpha, end energy of tstep}; true_out[i] = true path of
//! the geometric path gsteps[i] (true path limited by tstep)
int orc_msc_step_conversion(void* problem, std::uint32_t material, int par, double energy,
                            double tstep, double const* gsteps, std::uint32_t n,
                            double* out_info, double* true_out)
{
    return guard([&] {
        auto const& pd = *static_cast<OrcProblem*>(problem)->p->physics;
        if (!pd.msc.enabled)
            throw std::runtime_error("MSC is off");
        std::uint32_t const pid = par == 0 ? pd.electron : pd.positron;
        int const ep = pd.eloss_process.at(pid);
        if (ep < 0)
            throw std::runtime_error("no energy-loss process");
        ProcessDef const& eloss = pd.processes.at(pd.particle_processes.at(pid).at(ep));
        MscCtx ctx{pd, eloss, material, par, energy, eloss.range.at(material).calc(energy)};
        ctx.lambda = ctx.mfp(energy);
        double alpha = 0;
        double const me = pd.particles.at(pd.electron).mass;
        double const gp = msc_step_to_geo(ctx, me, tstep, &alpha);
        out_info[0] = ctx.range;
        out_info[1] = ctx.lambda;
        out_info[2] = gp;
        out_info[3] = alpha;
        out_info[4] = tstep < ctx.range ? ctx.end_energy(tstep) : 0;
        for (std::uint32_t i = 0; i < n; ++i)
            true_out[i] = msc_step_from_geo(ctx, tstep, alpha, gsteps[i]);
    });
}

//! n samples of simple_scattering on a default-seeded mt19937
void orc_msc_simple_scattering(double xmean, double x2mean, std::uint32_t n, double* out)
{
    std::mt19937 mt;
    MtUniform u{mt};
    for (std::uint32_t i = 0; i < n; ++i)
        out[i] = msc_simple_scattering(xmean, x2mean, u);
}

//! n samples of sample_cos_theta on a default-seeded mt19937, for the
//! coefficients of an effective Z; in = {tau, theta0, true_path,
//! limit_min, lambda, rad_length}
void orc_msc_sample_cos_theta(double zeff, double const* in, std::uint32_t n, double* out)
{
    MscDataO::Material const md = msc_material_coeffs(zeff);
    double const tau = in[0];
    MscAngleInput const a{std::exp(-tau), (1 + 2 * std::exp(-2.5 * tau)) / 3, tau, in[1],
                          in[2], in[3], in[4], in[5], &md};
    std::mt19937 mt;
    MtUniform u{mt};
    for (std::uint32_t i = 0; i < n; ++i)
        out[i] = msc_sample_cos_theta(a, u);
}

double orc_msc_theta0(double zeff, double y, double e_in, double e_end, double me)
{
    MscDataO::Material const md = msc_material_coeffs(zeff);
    return msc_theta0(md, y, e_in, e_end, me);
}
}

extern "C" {
//! Parameters of the tail model for the same inputs as
//! orc_msc_sample_cos_theta: out = {xsi, x, c, d, prob, qprob, xmean_1,
//! xmean_2, ea}
void orc_msc_tail_parts(double zeff, double const* in, double* out)
{
    MscDataO::Material const md = msc_material_coeffs(zeff);
    double const tau = in[0], theta0 = in[1], true_path = in[2], limit_min = in[3];
    double const lambda = in[4], rad_length = in[5];
    double const xmean = std::exp(-tau);
    double const maxtau = true_path < limit_min ? limit_min / lambda : tau;
    double const uu = detail::fastpow(maxtau, 1.0 / 6);
    double const radlen_mfp = true_path / (tau * rad_length);
    double xsi = std::fma(uu, std::fma(uu, md.tail[2], md.tail[1]), md.tail[0])
                 + md.tail_corr * std::log(radlen_mfp);
    xsi = std::max(xsi, 1.9);
    double const ea = std::exp(-xsi);
    double const sh = 2 * std::sin(0.5 * theta0);
    double const x = sh * sh;
    double const xmean_1 = 1 - x * (1 + (xsi * ea) / (1 - ea));
    double c = xsi;
    if (std::fabs(xsi - 3) < 0.001)
        c = 3.001;
    else if (std::fabs(xsi - 2) < 0.001)
        c = 2.001;
    double const b1 = 2 + (c - xsi) * x;
    double const d = detail::fastpow(c * x / b1, c - 1);
    double const x0 = 1 - xsi * x;
    double const xmean_2 = (x0 + d - (c * x - b1 * d) / (c - 2)) / (1 - d);
    double const f2x0 = (c - 1) / (c * (1 - d));
    double const prob = f2x0 / (ea / (1 - ea) + f2x0);
    double const qprob = xmean / (prob * xmean_1 + (1 - prob) * xmean_2);
    double const v[] = {xsi, x, c, d, prob, qprob, xmean_1, xmean_2, ea};
    std::copy(std::begin(v), std::end(v), out);
}
}

extern "C" {
//! Plant a deliberate mistake in the grid calculators (0: none)
void orc_grid_set_mutation(int m)
{
    orc::g_grid_mutation = m;
}

//! Evaluate a log-uniform grid built from (emin, emax, size, prime index,
//! values): kind 0 = cross section (XsCalculator), 1 = range
//! (RangeCalculator), 2 = inverse range (InverseRangeCalculator, x = range)
void orc_grid_eval(double emin, double emax, std::uint32_t size, std::uint32_t prime,
                   double const* values, int kind, double const* x, std::uint32_t n,
                   double* out)
{
    XsGridO g;
    g.front = std::log(emin);
    g.back = std::log(emax);
    g.size = size;
    g.delta = (g.back - g.front) / (size - 1);
    g.prime = prime;
    g.values.assign(values, values + size);
    for (std::uint32_t i = 0; i < n; ++i)
    {
        out[i] = kind == 0   ? g.calc(x[i])
                 : kind == 1 ? calc_range(g, x[i])
                             : calc_inverse_range(g, x[i]);
    }
}

//! Range and the inverse range of that range for a particle's energy-loss
//! tables in a material
int orc_range_roundtrip(void* problem, std::uint32_t particle, std::uint32_t material,
                        double const* energies, std::uint32_t n, double* range_out,
                        double* energy_out)
{
    return guard([&] {
        auto const& pd = *static_cast<OrcProblem*>(problem)->p->physics;
        int const ep = pd.eloss_process.at(particle);
        if (ep < 0)
            throw std::runtime_error("no energy-loss process");
        ProcessDef const& eloss = pd.processes.at(pd.particle_processes.at(particle).at(ep));
        XsGridO const& g = eloss.range.at(material);
        for (std::uint32_t i = 0; i < n; ++i)
        {
            range_out[i] = calc_range(g, energies[i]);
            energy_out[i] = calc_inverse_range(g, range_out[i]);
        }
    });
}
}

//---------------------------------------------------------------------------//
// EVENT-PARALLEL RUN (celer-sim's OpenMP event mode, celer-sim.cc:115-137)
//---------------------------------------------------------------------------//
#include <chrono>
#include <omp.h>

extern "C" {
//! Transport primaries with one state of num_slots per OpenMP thread; events
//! are dealt round-robin by event ID. out = {track-steps, step iterations
//! summed over threads, wall seconds, threads used}; calo (optional) gets
//! the per-detector deposition summed over threads.
int orc_run_events_parallel(void* problem, void const* prim, std::uint32_t nprim,
                            std::uint32_t num_slots, int num_threads, double* out,
                            double* calo, std::uint32_t num_calo)
{
    return guard([&] {
        auto const& prob = static_cast<OrcProblem*>(problem)->p;
        auto const* p = static_cast<CPrimary const*>(prim);
        int const nt = num_threads > 0 ? num_threads : omp_get_max_threads();
        std::vector<std::vector<PrimaryP>> shares(nt);
        for (std::uint32_t i = 0; i < nprim; ++i)
        {
            PrimaryP q;
            q.particle = p[i].particle;
            q.event = p[i].event;
            q.energy = p[i].energy;
            for (int k = 0; k < 3; ++k)
            {
                q.pos[k] = p[i].pos[k];
                q.dir[k] = p[i].dir[k];
            }
            q.time = p[i].time;
            shares[q.event % nt].push_back(q);
        }
        std::vector<double> steps(nt, 0), iters(nt, 0);
        std::vector<std::vector<double>> tallies(nt);
        std::vector<std::string> errors(nt);
        auto const t0 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(nt)
        {
            int const t = omp_get_thread_num();
            try
            {
                auto st = make_state(prob, num_slots, static_cast<std::uint32_t>(t));
                if (!shares[t].empty())
                {
                    insert_primaries(*st, shares[t]);
                    StepResult r = step_once(*st);
                    steps[t] += r.active;
                    iters[t] += 1;
                    while (r.queued > 0 || r.alive > 0)
                    {
                        r = step_once(*st);
                        steps[t] += r.active;
                        iters[t] += 1;
                    }
                }
                tallies[t] = st->calo_edep;
            }
            catch (std::exception const& e)
            {
                errors[t] = e.what();
            }
        }
        double const secs
            = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        for (auto const& e : errors)
        {
            if (!e.empty())
                throw std::runtime_error(e);
        }
        out[0] = out[1] = 0;
        for (int t = 0; t < nt; ++t)
        {
            out[0] += steps[t];
            out[1] += iters[t];
        }
        out[2] = secs;
        out[3] = nt;
        if (calo)
        {
            for (std::uint32_t d = 0; d < num_calo; ++d)
            {
                calo[d] = 0;
                for (int t = 0; t < nt; ++t)
                    calo[d] += d < tallies[t].size() ? tallies[t][d] : 0.0;
            }
        }
    });
}
}

//---------------------------------------------------------------------------//
// USER OBSERVERS
//---------------------------------------------------------------------------//
extern "C" {
//! Attach observers to a state: action diagnostic over num_actions columns
//! (0: off), step diagnostic with step_bins bins (0: off), step collector
//! (collect != 0) with an optional volume -> detector map
int orc_state_enable_user(void* s, std::uint32_t num_actions, std::uint32_t step_bins,
                          int collect, std::uint32_t const* detector_map,
                          std::uint32_t num_volumes, int nonzero_edep)
{
    return guard([&] {
        State& st = *static_cast<OrcState*>(s)->s;
        auto u = std::make_shared<UserObservers>();
        std::size_t const np = st.problem->physics->particles.size();
        if (num_actions)
        {
            u->action_diag = true;
            u->num_actions = num_actions;
            u->action_counts.assign(np * num_actions, 0);
        }
        if (step_bins)
        {
            u->step_diag = true;
            u->step_bins = step_bins;
            u->step_counts.assign(np * step_bins, 0);
        }
        u->collect = collect != 0;
        if (detector_map)
            u->detector_map.assign(detector_map, detector_map + num_volumes);
        u->nonzero_edep = nonzero_edep != 0;
        st.user = std::move(u);
    });
}

//! Copy the action (which 0) or step (which 1) diagnostic counts
int orc_state_user_counts(void* s, int which, std::uint64_t* out, std::uint32_t size)
{
    return guard([&] {
        State& st = *static_cast<OrcState*>(s)->s;
        if (!st.user)
            throw std::runtime_error("oracle: no user observers");
        auto const& v = which == 0 ? st.user->action_counts : st.user->step_counts;
        if (size < v.size())
            throw std::runtime_error("oracle: counts buffer too small");
        std::copy(v.begin(), v.end(), out);
    });
}

//! Plant a deliberate mistake in the user observers (0: none)
void orc_user_set_mutation(int m)
{
    orc::g_user_mutation = m;
}

//! Number of collected hits
std::uint32_t orc_state_num_hits(void* s)
{
    State& st = *static_cast<OrcState*>(s)->s;
    return st.user ? static_cast<std::uint32_t>(st.user->hits.size()) : 0;
}

//! Copy and clear the hits: 9 u32 per hit (detector, track, event, parent,
//! step count, action, particle, pre volume, post volume) and 18 doubles
//! (step length, edep, then per point time, pos[3], dir[3], energy)
int orc_state_take_hits(void* s, std::uint32_t* u32, double* f64, std::uint32_t size)
{
    return guard([&] {
        State& st = *static_cast<OrcState*>(s)->s;
        if (!st.user)
            throw std::runtime_error("oracle: no user observers");
        auto& hits = st.user->hits;
        if (size < hits.size())
            throw std::runtime_error("oracle: hit buffer too small");
        for (std::size_t i = 0; i < hits.size(); ++i)
        {
            StepHit const& h = hits[i];
            std::uint32_t* u = u32 + 9 * i;
            u[0] = h.detector;
            u[1] = h.track_id;
            u[2] = h.event_id;
            u[3] = h.parent_id;
            u[4] = h.track_step_count;
            u[5] = h.action_id;
            u[6] = h.particle;
            u[7] = h.volume[0];
            u[8] = h.volume[1];
            double* f = f64 + 18 * i;
            f[0] = h.step_length;
            f[1] = h.edep;
            for (int p = 0; p < 2; ++p)
            {
                double* q = f + 2 + 8 * p;
                q[0] = h.time[p];
                for (int j = 0; j < 3; ++j)
                {
                    q[1 + j] = h.pos[p][j];
                    q[4 + j] = h.dir[p][j];
                }
                q[7] = h.energy[p];
            }
        }
        hits.clear();
    });
}
}
