//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see physics.hh).
//---------------------------------------------------------------------------//
#include "physics.hh"
#include "msc.hh"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>

#include "alongstep.hh"
#include "em.hh"
#include "problem.hh"

namespace orc
{
namespace
{
//---------------------------------------------------------------------------//
constexpr double inf = std::numeric_limits<double>::infinity();
constexpr double c_light = 29979245800.0;  // [cm/s]

//! Build an XS grid from lambda (unscaled) and lambda_prim (scaled by E)
//! vectors (grid/ValueGridBuilder.cc:85-172)
XsGridO grid_from_vectors(Json const* lo, Json const* hi)
{
    XsGridO g;
    std::vector<double> values;
    double emin, eprime, emax;
    if (lo && hi)
    {
        auto lx = (*lo)("x").nums(), ly = (*lo)("y").nums();
        auto hx = (*hi)("x").nums(), hy = (*hi)("y").nums();
        values.assign(ly.begin(), ly.end() - 1);
        values.insert(values.end(), hy.begin(), hy.end());
        emin = lx.front();
        eprime = hx.front();
        emax = hx.back();
    }
    else if (hi)
    {
        auto hx = (*hi)("x").nums();
        values = (*hi)("y").nums();
        emin = eprime = hx.front();
        emax = hx.back();
    }
    else
    {
        auto lx = (*lo)("x").nums();
        values = (*lo)("y").nums();
        emin = lx.front();
        emax = lx.back();
        eprime = 0;
    }
    g.front = std::log(emin);
    g.back = std::log(emax);
    g.size = static_cast<std::uint32_t>(values.size());
    g.delta = (g.back - g.front) / (g.size - 1);
    g.values = values;
    if (eprime > 0)
    {
        double le = std::log(eprime);
        auto p = static_cast<std::uint32_t>((le - g.front) / g.delta);
        double gp1 = g.front + g.delta * (p + 1);
        if (std::fabs(gp1 - le) <= 1e-12 * std::max(1.0, std::fabs(le)))
            ++p;
        g.prime = p;
    }
    return g;
}

//! Grid given directly as {emin, emax, prime, values}
XsGridO grid_from_json(Json const& j)
{
    XsGridO g;
    g.values = j("values").nums();
    g.size = static_cast<std::uint32_t>(g.values.size());
    g.front = std::log(j.num("emin", 0));
    g.back = std::log(j.num("emax", 0));
    g.delta = (g.back - g.front) / (g.size - 1);
    double p = j.num("prime", -1);
    g.prime = p < 0 ? 0xffffffffu : static_cast<std::uint32_t>(p);
    return g;
}

std::vector<XsGridO> grids_of(Json const& proc, char const* direct,
                              char const* lo_key, char const* hi_key,
                              std::size_t num_mats)
{
    std::vector<XsGridO> result(num_mats);
    if (auto d = proc.find(direct))
    {
        for (std::size_t m = 0; m < num_mats && m < d->a.size(); ++m)
            if (d->a[m].kind == Json::obj_k)
                result[m] = grid_from_json(d->a[m]);
        return result;
    }
    Json const* lo = lo_key ? proc.find(lo_key) : nullptr;
    Json const* hi = hi_key ? proc.find(hi_key) : nullptr;
    if (!lo && !hi)
        return result;
    for (std::size_t m = 0; m < num_mats; ++m)
    {
        Json const* l = lo ? &lo->a.at(m) : nullptr;
        Json const* h = hi ? &hi->a.at(m) : nullptr;
        result[m] = grid_from_vectors(l, h);
    }
    return result;
}

//---------------------------------------------------------------------------//
// Helpers mirroring the track views
//---------------------------------------------------------------------------//
PhysicsData const& phys(State const& s)
{
    return *s.problem->physics;
}

std::uint32_t material_of(State const& s, Track const& t)
{
    return s.problem->volume_to_material[t.geo.vol];
}

double calc_xs(State const& s, ProcessDef const& p, std::uint32_t mat, double e)
{
    XsGridO const& g = p.xs[mat];
    if (!g)
        return hardwired_xs(s, p, mat, e);
    return g.calc(e);
}

//! PhysicsTrackView::calc_max_xs (PhysicsTrackView.hh:446-478)
double calc_max_xs(State const& s, ProcessDef const& p, std::uint32_t mat, double e)
{
    PhysicsData const& pd = phys(s);
    double emax = p.energy_max_xs[mat];
    double exi = e * pd.min_eprime_over_e;
    if (emax >= exi && emax < e)
        return calc_xs(s, p, mat, emax);
    return std::max(calc_xs(s, p, mat, e), calc_xs(s, p, mat, exi));
}

//! PhysicsTrackView::range_to_step (PhysicsTrackView.hh:624-640)
double range_to_step(PhysicsData const& pd, double range)
{
    double rho = pd.min_range;
    if (range < rho * (1 + std::sqrt(std::numeric_limits<double>::epsilon())))
        return range;
    double alpha = pd.max_step_over_range;
    return alpha * range + rho * (1 - alpha) * (2 - rho / range);
}

//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
// XS GRID
//---------------------------------------------------------------------------//
//! Planted mistakes for the grid-calculator pins (0: none). 1: the upper
//! point of the bracket that is the prime index is not unscaled; 2: the
//! bracket is interpolated in log(E) instead of E; 3: the range below the
//! grid scales linearly instead of with sqrt(E); 4: the inverse range below
//! the grid scales linearly instead of quadratically.
int g_grid_mutation = 0;

double XsGridO::energy(std::uint32_t i) const
{
    return std::exp(front + delta * i);
}

double XsGridO::at(std::uint32_t i) const
{
    double r = values[i];
    if (i >= prime)
        r /= this->energy(i);
    return r;
}

double XsGridO::calc(double e) const
{
    double loge = std::log(e);
    auto extrapolated = [&](std::uint32_t i) {
        double r = values[i];
        if (i >= prime)
            r /= e;
        return r;
    };
    if (loge <= front)
        return extrapolated(0);
    if (loge >= back)
        return extrapolated(size - 1);
    auto lower = static_cast<std::uint32_t>((loge - front) / delta);
    // A log a few ulps below the stored back can round to the last point
    lower = std::min(lower, size - 2);
    double upper_e = std::exp(front + delta * (lower + 1));
    double upper_xs = values[lower + 1];
    if (lower + 1 == prime && g_grid_mutation != 1)
        upper_xs /= upper_e;
    double lower_e = std::exp(front + delta * lower);
    double lower_xs = values[lower];
    // Interpolator<linear, linear>: fma(slope, x - x_l, y_l)
    double slope = ((-lower_xs) + upper_xs) / ((-lower_e) + upper_e);
    double result = std::fma(slope, (-lower_e) + e, lower_xs);
    if (g_grid_mutation == 2)
    {
        double const t = (loge - (front + delta * lower)) / delta;
        result = lower_xs + t * (upper_xs - lower_xs);
    }
    if (lower >= prime)
        result /= e;
    return result;
}

//---------------------------------------------------------------------------//
// SETUP
//---------------------------------------------------------------------------//
void build_physics(Json const& inp, PhysicsData& out)
{
    if (auto parts = inp.find("particles"))
    {
        for (auto const& p : parts->a)
        {
            ParticleDef d;
            d.label = p.str("label", "");
            d.pdg = static_cast<int>(p.num("pdg", 0));
            d.mass = p.num("mass", 0);
            d.charge = p.num("charge", 0);
            out.particles.push_back(d);
        }
    }
    for (std::uint32_t i = 0; i < out.particles.size(); ++i)
    {
        if (out.particles[i].pdg == 22)
            out.gamma = i;
        else if (out.particles[i].pdg == 11)
            out.electron = i;
        else if (out.particles[i].pdg == -11)
            out.positron = i;
    }
    if (auto els = inp.find("elements"))
    {
        for (auto const& e : els->a)
        {
            ElementDef d;
            d.z = static_cast<int>(e.num("z", 0));
            d.atomic_mass = e.num("atomic_mass", 0);
            d.cbrt_z = std::cbrt(static_cast<double>(d.z));
            d.log_z = std::log(static_cast<double>(d.z));
            d.coulomb_correction = coulomb_correction(d.z);
            d.mass_radiation_coeff = mass_radiation_coeff(d.z, d.atomic_mass);
            if (auto sb = e.find("sb_table"))
                d.sb = make_sb_table(*sb);
            if (auto pe = e.find("livermore_pe"))
                d.pe = make_livermore(*pe);
            if (out.electron != none)
                set_rb_element_data(out.particles[out.electron].mass, d);
            out.elements.push_back(d);
        }
    }
    if (auto mats = inp.find("materials"))
    {
        for (auto const& m : mats->a)
        {
            MaterialDef d;
            d.label = m.str("label", "");
            d.number_density = m.num("number_density", 0);
            d.electron_density = m.num("electron_density", 0);
            d.mean_exc_energy = m.num("mean_exc_energy", 0);
            if (d.mean_exc_energy > 0)
            {
                // MaterialParams keeps I as exp(log I)
                d.log_mean_exc_energy = std::log(d.mean_exc_energy);
                d.mean_exc_energy = std::exp(d.log_mean_exc_energy);
            }
            d.rad_length = m.num("rad_length", 0);
            d.zeff = m.num("zeff", 0);
            if (auto comps = m.find("elements"))
                for (auto const& c : comps->a)
                    d.elements.emplace_back(
                        static_cast<std::uint32_t>(c.num("element", 0)),
                        c.num("fraction", 1));
            if (d.rad_length == 0 && d.number_density > 0 && !d.elements.empty())
            {
                // Derived as MaterialParams does (MaterialParams.cc:496-519):
                // 1 / (sum_i w_i (1/X0)_i rho), rho = n sum_i w_i A_i amu
                constexpr double amu_gram = 1.66053906660e-24;
                double avg_amu = 0, rad_coeff = 0;
                for (auto const& c : d.elements)
                {
                    ElementDef const& el = out.elements.at(c.first);
                    avg_amu += c.second * el.atomic_mass;
                    rad_coeff += c.second * el.mass_radiation_coeff;
                }
                double const density = d.number_density * avg_amu * amu_gram;
                d.rad_length = 1 / (rad_coeff * density);
            }
            out.materials.push_back(d);
        }
    }
    std::size_t const num_mats = std::max<std::size_t>(out.materials.size(), 1);
    if (auto opts = inp.find("options"))
    {
        out.secondary_stack_factor = opts->num("secondary_stack_factor", 3);
        out.min_range = opts->num("min_range", out.min_range);
        out.max_step_over_range = opts->num("max_step_over_range", out.max_step_over_range);
        out.min_eprime_over_e = opts->num("min_eprime_over_e", out.min_eprime_over_e);
        out.linear_loss_limit = opts->num("linear_loss_limit", out.linear_loss_limit);
        out.lowest_electron_energy
            = opts->num("lowest_electron_energy", out.lowest_electron_energy);
        out.fixed_step_limiter = opts->num("fixed_step_limiter", 0);
        out.apply_post_interaction_cut = opts->flag("apply_cuts", false);
        out.fluct = opts->flag("eloss_fluctuation", false);
        out.enable_lpm = opts->flag("lpm", false);
    }
    if (out.fluct)
    {
        // FluctuationParams.cc: two oscillators per material
        for (auto const& m : out.materials)
        {
            FluctO f;
            double const avg_z = m.electron_density / m.number_density;
            f.strength[1] = avg_z > 2 ? 2 / avg_z : 0;
            f.strength[0] = 1 - f.strength[1];
            f.binding[1] = 1e-5 * (avg_z * avg_z);
            f.binding[0] = std::pow(m.mean_exc_energy / std::pow(f.binding[1], f.strength[1]),
                                    1 / f.strength[0]);
            f.log_binding[1] = std::log(f.binding[1]);
            f.log_binding[0] = std::log(f.binding[0]);
            out.fluct_params.push_back(f);
        }
    }
    if (auto mj = inp.find("msc"))
    {
        // UrbanMscParams.cc and its calc_material_data
        MscDataO& msc = out.msc;
        msc.enabled = true;
        msc.range_factor = mj->num("range_factor", 0.04);
        msc.safety_factor = mj->num("safety_factor", 0.6);
        msc.lambda_limit = mj->num("lambda_limit", 0.1);
        std::string const alg = mj->str("step_limit_algorithm", "safety");
        if (alg == "minimal")
            msc.limit_alg = MscLimitAlg::minimal;
        else if (alg == "safety_plus")
            msc.limit_alg = MscLimitAlg::safety_plus;
        else if (alg == "safety")
            msc.limit_alg = MscLimitAlg::safety;
        else
            throw std::runtime_error("unknown MSC step limit algorithm '" + alg + "'");
        msc.xs[0] = grids_of(*mj, "electron", nullptr, nullptr, num_mats);
        msc.xs[1] = grids_of(*mj, "positron", nullptr, nullptr, num_mats);
        msc.low = std::exp(msc.xs[0].at(0).front);
        msc.high = std::exp(msc.xs[0].at(0).back);
        for (auto const& m : out.materials)
        {
            MscDataO::Material const d = msc_material_coeffs(m.zeff);
            msc.mat.push_back(d);
        }
    }
    // Cutoffs: {"particle label": [{"energy": E}, ...per material]}
    out.cutoff_energy.assign(out.particles.size(), std::vector<double>(num_mats, 0));
    if (auto cuts = inp.find("cutoffs"))
    {
        for (auto const& kv : cuts->o)
        {
            for (std::uint32_t p = 0; p < out.particles.size(); ++p)
            {
                if (out.particles[p].label != kv.first)
                    continue;
                for (std::size_t m = 0; m < num_mats && m < kv.second.a.size(); ++m)
                    out.cutoff_energy[p][m] = kv.second.a[m].num("energy", 0);
            }
        }
    }
    out.particle_processes.assign(out.particles.size(), {});
    out.eloss_process.assign(out.particles.size(), -1);
    if (auto procs = inp.find("processes"))
    {
        for (auto const& pj : procs->a)
        {
            ProcessDef p;
            p.label = pj.str("label", "");
            p.cls = pj.str("class", "");
            std::string plabel = pj.str("particle", "");
            for (std::uint32_t i = 0; i < out.particles.size(); ++i)
                if (out.particles[i].label == plabel)
                    p.particle = i;
            if (p.particle == none)
                throw std::runtime_error("oracle: unknown process particle " + plabel);
            p.integral = pj.flag("integral", false);
            p.xs = grids_of(pj, "xs", "lambda", "lambda_prim", num_mats);
            p.dedx = grids_of(pj, "dedx_grid", "dedx", nullptr, num_mats);
            p.range = grids_of(pj, "range_grid", "range", nullptr, num_mats);
            for (auto const& mj : pj("models").a)
            {
                ModelDef m;
                m.label = mj.str("label", "");
                m.model = mj.str("model", "");
                m.low = mj.num("low", 0);
                m.high = mj.num("high", 0);
                p.models.push_back(m);
                // Element-selection CDFs (PhysicsParams::build_model_xs)
                std::vector<std::vector<XsGridO>> cdf(num_mats);
                if (Json const* cj = mj.find("element_cdf"))
                {
                    for (std::size_t mat = 0; mat < num_mats && mat < cj->a.size(); ++mat)
                    {
                        Json const& jm = cj->a[mat];
                        if (!jm.find("cdf"))
                            continue;
                        for (auto const& row : jm("cdf").a)
                        {
                            XsGridO g;
                            g.values = row.nums();
                            g.size = static_cast<std::uint32_t>(g.values.size());
                            g.front = std::log(jm.num("emin", 0));
                            g.back = std::log(jm.num("emax", 0));
                            g.delta = (g.back - g.front) / (g.size - 1);
                            cdf[mat].push_back(std::move(g));
                        }
                    }
                }
                p.elem_cdf.push_back(std::move(cdf));
                // A model label names one action shared by all its particles
                if (std::find(out.model_labels.begin(), out.model_labels.end(), m.label)
                    == out.model_labels.end())
                    out.model_labels.push_back(m.label);
            }
            if (p.integral)
            {
                p.energy_max_xs.assign(num_mats, 0);
                for (std::size_t m = 0; m < num_mats; ++m)
                {
                    XsGridO const& g = p.xs[m];
                    if (p.cls == "annihilation" || !g)
                        continue;
                    double xs_max = 0, e_max = 0;
                    for (std::uint32_t i = 0; i < g.size; ++i)
                    {
                        double xs = g.at(i);
                        if (xs > xs_max)
                        {
                            xs_max = xs;
                            e_max = g.energy(i);
                        }
                    }
                    p.energy_max_xs[m] = e_max;
                }
            }
            auto idx = static_cast<std::uint32_t>(out.processes.size());
            auto& list = out.particle_processes[p.particle];
            if (!p.dedx.empty() && p.dedx[0])
                out.eloss_process[p.particle] = static_cast<int>(list.size());
            list.push_back(idx);
            out.processes.push_back(std::move(p));
        }
    }
}

void PhysicsData::finalize_actions(ActionIds const& ids)
{
    discrete_action = ids.discrete_select;
    msc_action = ids.msc_range;
    range_action = ids.eloss_range;
    integral_rejection_action = ids.integral_rejected;
    failure_action = ids.failure;
    fixed_step_action = ids.eloss_range;
    for (auto& p : processes)
        for (auto& m : p.models)
        {
            auto k = std::find(model_labels.begin(), model_labels.end(), m.label)
                     - model_labels.begin();
            m.action = ids.model.at(static_cast<std::size_t>(k));
        }
}

//---------------------------------------------------------------------------//
// PRE-STEP (PreStepExecutor.hh:46-115, PhysicsStepUtils.hh:37-120)
//---------------------------------------------------------------------------//
void pre_step(State& s)
{
    s.stack_size = 0;
    PhysicsData const& pd = phys(s);
    ActionIds const& ids = s.problem->actions;
    for (auto& t : s.tracks)
    {
        if (t.status == inactive)
        {
            t.step_length = inf;
            t.post_action = none;
            t.along_action = none;
            continue;
        }
        t.edep = 0;
        t.sec_start = 0;
        t.sec_count = 0;
        t.element = none;
        t.pre_volume = t.geo.vol;
        if (t.status == errored)
            continue;
        t.status = alive;
        if (pd.particles.empty())
            continue;

        if (!(t.interaction_mfp > 0))
        {
            // ExponentialDistribution(1): -log(u)
            t.interaction_mfp = -std::log(canonical_double(t.rng));
        }
        std::uint32_t const mat = material_of(s, t);
        auto const& procs = pd.particle_processes[t.particle];
        double total = 0;
        for (std::size_t i = 0; i < procs.size(); ++i)
        {
            ProcessDef const& p = pd.processes[procs[i]];
            double xs = p.integral ? calc_max_xs(s, p, mat, t.energy)
                                   : calc_xs(s, p, mat, t.energy);
            total += xs;
            t.per_process_xs[i] = xs;
        }
        t.macro_xs = total;
        double step;
        std::uint32_t action = pd.discrete_action;
        if (t.energy == 0)
        {
            step = 0;
        }
        else
        {
            step = t.interaction_mfp / total;
            int ep = pd.eloss_process[t.particle];
            if (ep >= 0)
            {
                ProcessDef const& p = pd.processes[procs[ep]];
                double range = calc_range(p.range[mat], t.energy);
                t.dedx_range = range;
                double eloss_step = range_to_step(pd, range);
                if (eloss_step <= step)
                {
                    step = eloss_step;
                    action = pd.range_action;
                }
                if (pd.fixed_step_limiter > 0 && pd.fixed_step_limiter < step)
                {
                    step = pd.fixed_step_limiter;
                    action = pd.fixed_step_action;
                }
            }
            else if (procs.empty())
            {
                action = none;
            }
        }
        t.step_length = step;
        t.post_action = action;
        double const charge = pd.particles[t.particle].charge;
        t.along_action = (charge == 0) ? ids.along_neutral : ids.along_user;
    }
}

//---------------------------------------------------------------------------//
// ALONG-STEP (NEUTRAL): linear propagation, time, MFP update
//---------------------------------------------------------------------------//
namespace
{
//! PropagationApplier with a LinearPropagator (PropagationApplier.hh:95-195)
void propagate_linear(State& s, Track& t)
{
    ActionIds const& ids = s.problem->actions;
    if (t.step_length == 0)
        return;
    bool boundary_hit;
    double const orig[3] = {t.geo.pos[0], t.geo.pos[1], t.geo.pos[2]};
    double dist = geo_find_next_step(s.problem->geo, t.geo, t.step_length, &boundary_hit);
    if (boundary_hit)
        geo_move_to_boundary(t.geo);
    else
        geo_move_internal(t.geo, t.step_length);
    if (t.geo.pos[0] == orig[0] && t.geo.pos[1] == orig[1] && t.geo.pos[2] == orig[2])
    {
        // CoreTrackView::apply_errored
        t.status = errored;
        t.along_action = none;
        t.post_action = ids.tracking_cut;
        return;
    }
    if (boundary_hit)
    {
        t.step_length = dist;
        t.post_action = ids.boundary;
    }
    else if (dist < t.step_length)
    {
        t.step_length = dist;
        t.post_action = ids.propagation_limit;
    }
}

//! TimeUpdater (TimeUpdater.hh:28-45)
void update_time(State& s, Track& t)
{
    if (t.status == errored)
        return;
    double const mass = phys(s).particles[t.particle].mass;
    double inv_gamma = mass / (t.energy + mass);
    double beta_sq = 1 - inv_gamma * inv_gamma;
    double speed = std::sqrt(beta_sq) * c_light;
    if (speed > 0)
        t.time += t.step_length / speed;
}

//! TrackUpdater (TrackUpdater.hh:32-66)
void update_track(State& s, Track& t)
{
    if (t.status == errored)
        return;
    if (t.status == alive && t.post_action != phys(s).discrete_action)
    {
        t.interaction_mfp -= t.step_length * t.macro_xs;
    }
    ++t.num_steps;
}
}  // namespace

void along_step_neutral(State& s)
{
    std::uint32_t const aid = s.problem->actions.along_neutral;
    for (auto& t : s.tracks)
    {
        if (t.status == inactive || t.along_action != aid)
            continue;
        propagate_linear(s, t);
        update_time(s, t);
        update_track(s, t);
    }
}

//---------------------------------------------------------------------------//
// DISCRETE SELECT (DiscreteSelectExecutor.hh:38-63, PhysicsStepUtils.hh:253)
//---------------------------------------------------------------------------//
void discrete_select(State& s)
{
    PhysicsData const& pd = phys(s);
    for (auto& t : s.tracks)
    {
        if (t.post_action != pd.discrete_action || t.status == inactive)
            continue;
        t.interaction_mfp = 0;
        std::uint32_t const mat = material_of(s, t);
        auto const& procs = pd.particle_processes[t.particle];
        // Selector: accumulate until positive (random/Selector.hh)
        double accum = -t.macro_xs * canonical_double(t.rng);
        std::size_t pp = procs.size() - 1;
        for (std::size_t i = 0; i + 1 < procs.size(); ++i)
        {
            accum += t.per_process_xs[i];
            if (accum > 0)
            {
                pp = i;
                break;
            }
        }
        ProcessDef const& p = pd.processes[procs[pp]];
        if (p.integral)
        {
            double xs = calc_xs(s, p, mat, t.energy);
            if (canonical_double(t.rng) * t.per_process_xs[pp] > xs)
            {
                t.post_action = pd.integral_rejection_action;
                continue;
            }
        }
        // Model finder: last model whose lower bound is <= energy
        std::size_t mi = 0;
        for (std::size_t k = 1; k < p.models.size(); ++k)
            if (t.energy >= p.models[k].low)
                mi = k;
        t.element = sample_element(s, t, p, mi, mat);
        t.post_action = p.models[mi].action;
    }
}

//---------------------------------------------------------------------------//
// INTERACTION (InteractionApplier.hh:106-171)
//---------------------------------------------------------------------------//
void interact(State& s, std::uint32_t model_action)
{
    PhysicsData const& pd = phys(s);
    ModelDef const* model = nullptr;
    for (auto const& p : pd.processes)
        for (auto const& m : p.models)
            if (m.action == model_action)
                model = &m;
    if (!model)
        throw std::runtime_error("oracle: no model for action");
    for (auto& t : s.tracks)
    {
        if (t.post_action != model_action || t.status == inactive)
            continue;
        Interaction r = sample_interaction(s, t, *model);
        if (r.action == Interaction::failed)
        {
            // step_limit({0, failure}): only applies if limiting
            if (0 < t.step_length)
            {
                t.step_length = 0;
                t.post_action = pd.failure_action;
            }
            continue;
        }
        if (r.action == Interaction::unchanged)
            continue;
        t.energy = r.energy;
        if (r.action != Interaction::absorbed)
            geo_set_dir(s.problem->geo, t.geo, r.dir);
        else
            t.status = killed;
        double deposition = r.energy_deposition;
        std::uint32_t const mat = material_of(s, t);
        for (std::uint32_t k = 0; k < r.num_secondaries; ++k)
        {
            SecondaryP& sec = s.stack[r.sec_start + k];
            if (pd.apply_post_interaction_cut && sec.particle != none
                && sec.energy < pd.cutoff_energy[sec.particle][mat])
            {
                deposition += sec.energy;
                if (sec.particle == pd.positron)
                    deposition += 2 * pd.particles[sec.particle].mass;
                sec = SecondaryP{};
            }
        }
        t.edep += deposition;
        t.sec_start = r.sec_start;
        t.sec_count = r.num_secondaries;
    }
}

//---------------------------------------------------------------------------//
// BOUNDARY (BoundaryExecutor.hh:42-84) AND TRACKING CUT (:49-84)
//---------------------------------------------------------------------------//
void boundary(State& s)
{
    ActionIds const& ids = s.problem->actions;
    for (auto& t : s.tracks)
    {
        if (t.post_action != ids.boundary || t.status != alive)
            continue;
        geo_cross_boundary(s.problem->geo, t.geo);
        if (t.geo.failed)
        {
            t.status = errored;
            t.along_action = none;
            t.post_action = ids.tracking_cut;
        }
        else if (!geo_is_outside(t.geo))
        {
            if (material_of(s, t) == none)
            {
                t.status = errored;
                t.along_action = none;
                t.post_action = ids.tracking_cut;
            }
        }
        else
        {
            t.status = killed;
        }
    }
}

void tracking_cut(State& s)
{
    PhysicsData const& pd = phys(s);
    ActionIds const& ids = s.problem->actions;
    for (auto& t : s.tracks)
    {
        if (t.post_action != ids.tracking_cut || t.status == inactive)
            continue;
        double dep = t.energy;
        if (t.particle == pd.positron)
            dep += 2 * pd.particles[t.particle].mass;
        t.edep += dep;
        t.energy = 0;
        t.status = killed;
    }
}

//---------------------------------------------------------------------------//
// ALONG-STEP (CHARGED, NO FIELD): AlongStepGeneralLinearAction.cc
//---------------------------------------------------------------------------//
double calc_range(XsGridO const& g, double energy)
{
    // RangeCalculator::operator() (grid/RangeCalculator.hh:80-110)
    double const loge = std::log(energy);
    if (loge <= g.front)
        return g.values[0] * std::exp((g_grid_mutation == 3 ? 1.0 : 0.5) * (loge - g.front));
    if (loge >= g.back)
        return g.values[g.size - 1];
    auto const i = std::min(static_cast<std::uint32_t>((loge - g.front) / g.delta), g.size - 2);
    double const xl = std::exp(g.front + g.delta * i);
    double const xr = std::exp(g.front + g.delta * (i + 1));
    double const slope = ((-g.values[i]) + g.values[i + 1]) / ((-xl) + xr);
    return std::fma(slope, (-xl) + energy, g.values[i]);
}

double calc_inverse_range(XsGridO const& g, double range)
{
    // InverseRangeCalculator::operator() (InverseRangeCalculator.hh:87-117)
    auto const& r = g.values;
    if (range < r.front())
    {
        double const f = range / r.front();
        return std::exp(g.front) * (g_grid_mutation == 4 ? f : f * f);
    }
    if (range >= r.back())
        return std::exp(g.back);
    auto it = std::lower_bound(r.begin(), r.end(), range);
    auto i = static_cast<std::uint32_t>(it - r.begin());
    if (*it != range)
        --i;
    double const el = std::exp(g.front + g.delta * i);
    double const er = std::exp(g.front + g.delta * (i + 1));
    double const slope = ((-el) + er) / ((-r[i]) + r[i + 1]);
    return std::fma(slope, (-r[i]) + range, el);
}

void along_step_charged(State& s)
{
    // Urban MSC, linear propagation, fluctuations: oracle/alongstep.cc
    along_step_charged_impl(s);
}

void calo_tally(State& s)
{
    Problem const& prob = *s.problem;
    for (auto const& t : s.tracks)
    {
        if (t.status == inactive || t.edep == 0)
            continue;
        // Every volume (the per-event tally of the serving loop)
        s.total_edep += t.edep;
        if (prob.num_detectors == 0)
            continue;
        std::uint32_t det = prob.volume_to_detector[t.pre_volume];
        if (det != none)
            s.calo_edep[det] += t.edep;
    }
}

//---------------------------------------------------------------------------//
}  // namespace orc
