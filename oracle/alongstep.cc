//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/alongstep.cc
//! Along-step action for charged tracks without a field, restated from
//! global/alongstep/AlongStepGeneralLinearAction.cc and its appliers:
//! - MSC: em/msc/UrbanMsc.hh (is_applicable, limit_step, apply_step),
//!   em/msc/detail/{UrbanMscHelper,UrbanMscSafetyStepLimit,MscStepToGeo,
//!   MscStepFromGeo,UrbanMscScatter,UrbanPositronCorrector}.hh
//! - propagation: alongstep/detail/PropagationApplier.hh:95-195
//! - energy loss: alongstep/detail/{FluctELoss,MeanELoss,ElossApplier}.hh,
//!   em/distribution/{EnergyLossHelper,EnergyLossUrbanDistribution,
//!   EnergyLossGaussianDistribution}.hh, random/distribution/
//!   {Normal,Poisson,UniformReal,Bernoulli}Distribution.hh
//! - TimeUpdater.hh, TrackUpdater.hh
//!
//! Pinned by: the Urban fluctuation golden values of
//! EnergyLossHelper.test.cc:238-285 (mean loss and the mt19937 draw count of
//! 10,000 samples, through orc_urban_fluct_sample). The MSC pieces are in
//! msc.cc, with their pins listed in msc.hh.
//---------------------------------------------------------------------------//
#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>

#include "alongstep.hh"
#include "em.hh"
#include "msc.hh"
#include "physics.hh"
#include "problem.hh"

namespace orc
{
namespace
{
constexpr double pi = 3.14159265358979323846;
constexpr double c_light = 29979245800.0;  // [cm/s]

//---------------------------------------------------------------------------//
// Distributions

template<class U>
unsigned int poisson(double lambda, U& u)
{
    if (lambda <= 16)
    {
        int k = 0;
        double p = std::exp(lambda);
        do
        {
            ++k;
            p *= u();
        } while (p > 1);
        return static_cast<unsigned int>(k - 1);
    }
    NormalDist<U> n{lambda, std::sqrt(lambda)};
    double const x = n(u) + 0.5;
    return x > 0 ? static_cast<unsigned int>(x) : 0u;
}

template<class U>
double uniform_real(double a, double b, U& u)
{
    return std::fma(b - a, u(), a);
}


//---------------------------------------------------------------------------//
// Urban fluctuations
template<class U>
double fast_urban(double mean, double stddev, U& u)
{
    if (stddev <= 4 * mean)
    {
        NormalDist<U> n{mean, stddev};
        double const max_loss = 2 * mean;
        double r;
        do
        {
            r = n(u);
        } while (r <= 0 || r > max_loss);
        return r;
    }
    return uniform_real(0.0, 2 * mean, u);
}

}  // namespace

//---------------------------------------------------------------------------//
template<class U>
double urban_fluct_sample(FluctO const& par, MaterialDef const& mat, double unscaled_mean,
                          double max_energy, double two_mebsgs, double beta_sq, U& u)
{
    constexpr double e0 = 1e-5;
    constexpr double rate = 0.56;
    constexpr double max_coll = 8;
    double const scaling_loss = 0.5 * std::min(1e-3 / max_energy, 1.0) + 1;
    double const mean_loss = unscaled_mean / scaling_loss;
    double b[2] = {par.binding[0], par.binding[1]};
    double xs_exc[2] = {0, 0};
    if (max_energy > mat.mean_exc_energy)
    {
        double const w = std::log(two_mebsgs) - beta_sq;
        double const w0 = mat.log_mean_exc_energy;
        if (w > w0)
        {
            if (w > par.log_binding[1])
            {
                double const c = mean_loss * (1 - rate) / (w - w0);
                for (int i = 0; i < 2; ++i)
                    xs_exc[i] = c * par.strength[i] * (w - par.log_binding[i]) / par.binding[i];
            }
            else
            {
                xs_exc[0] = mean_loss * (1 - rate) / par.binding[0];
            }
            double scaling = 4;
            if (xs_exc[0] < 42)
                scaling = 0.5 + (scaling - 0.5) * std::sqrt(xs_exc[0] / 42);
            b[0] *= scaling;
            xs_exc[0] /= scaling;
        }
    }
    double xs_ion = mean_loss * (max_energy - e0) / (max_energy * e0 * std::log(max_energy / e0));
    if (xs_exc[0] + xs_exc[1] > 0)
        xs_ion *= rate;

    double exc = 0;
    {
        double mean = 0, var = 0;
        for (int i = 0; i < 2; ++i)
        {
            if (xs_exc[i] > max_coll)
            {
                mean += xs_exc[i] * b[i];
                var += xs_exc[i] * (b[i] * b[i]);
            }
            else if (xs_exc[i] > 0)
            {
                unsigned int n = poisson(xs_exc[i], u);
                if (n > 0)
                    exc += uniform_real(double(n - 1), double(n + 1), u) * b[i];
            }
        }
        if (var > 0)
            exc += fast_urban(mean, std::sqrt(var), u);
    }
    double ion = 0;
    {
        double const ratio = max_energy / e0;
        double alpha = 1, mean_num = 0;
        if (xs_ion > max_coll)
        {
            alpha = (xs_ion + max_coll) * ratio / (max_coll * ratio + xs_ion);
            double const mean_loss_coll = alpha * std::log(alpha) / (alpha - 1);
            mean_num = xs_ion * ratio * (alpha - 1) / ((ratio - 1) * alpha);
            double const mean = mean_num * mean_loss_coll * e0;
            double const sd = e0 * std::sqrt(xs_ion * (alpha - mean_loss_coll * mean_loss_coll));
            ion += fast_urban(mean, sd, u);
        }
        if (xs_ion > 0 && ratio > alpha)
        {
            unsigned int n = poisson(xs_ion - mean_num, u);
            for (; n > 0; --n)
                ion += alpha * e0 / uniform_real(alpha / ratio, 1.0, u);
        }
    }
    return scaling_loss * (exc + ion);
}

template double urban_fluct_sample<MtUniform>(FluctO const&, MaterialDef const&, double,
                                              double, double, double, MtUniform&);

//---------------------------------------------------------------------------//
void along_step_charged_impl(State& s)
{
    Problem const& prob = *s.problem;
    PhysicsData const& pd = *prob.physics;
    ActionIds const& ids = prob.actions;
    std::uint32_t const aid = ids.along_user;
    for (auto& t : s.tracks)
    {
        if (t.status == inactive || t.along_action != aid)
            continue;
        XorwowUniform u{t.rng};
        std::uint32_t const mat = prob.volume_to_material[t.geo.vol];
        auto const& procs = pd.particle_processes[t.particle];
        int const ep = pd.eloss_process[t.particle];
        ProcessDef const* eloss = ep >= 0 ? &pd.processes[procs[ep]] : nullptr;

        // --- MscStepLimitApplier / UrbanMsc::limit_step
        bool const applicable = pd.msc.enabled && t.step_length > msc_geom_limit
                                && t.status == alive
                                && (t.particle == pd.electron || t.particle == pd.positron)
                                && t.energy > pd.msc.low && t.energy < pd.msc.high;
        double true_path = 0, geom_path = 0, alpha = 0;
        bool displaced = false;
        MscCtx ctx{pd, *(eloss ? eloss : &pd.processes.front()), mat,
                   t.particle == pd.electron ? 0 : 1, t.energy, t.dedx_range};
        if (applicable)
        {
            ctx.lambda = ctx.mfp(t.energy);
            double const phys_step = t.step_length;
            bool have_path = false;
            if (phys_step <= msc_limit_min_fix)
            {
                true_path = phys_step;
                have_path = true;
            }
            double safety = 0;
            bool const on_boundary = t.geo.surf != none;
            if (!have_path && !on_boundary)
            {
                double const max_step = ctx.range * pd.msc.mat[mat].d_over_r;
                safety = geo_find_safety(prob.geo, t.geo);
                if (safety >= max_step)
                {
                    true_path = phys_step;
                    have_path = true;
                }
            }
            if (!have_path)
            {
                displaced = true;
                MscLimitInput in;
                in.alg = pd.msc.limit_alg;
                in.energy = t.energy;
                in.range = ctx.range;
                in.lambda = ctx.lambda;
                in.md = &pd.msc.mat[mat];
                in.par = ctx.par;
                in.range_factor = pd.msc.range_factor;
                in.safety_factor = pd.msc.safety_factor;
                in.lambda_limit = pd.msc.lambda_limit;
                in.on_boundary = on_boundary;
                in.safety = safety;
                in.phys_step = phys_step;
                MscRangeO r{t.msc_range_init, t.msc_range_factor, t.msc_limit_min};
                true_path = msc_step_limit(in, r, u);
                t.msc_range_init = r.range_init;
                t.msc_range_factor = r.range_factor;
                t.msc_limit_min = r.limit_min;
            }
            bool limited = true_path < phys_step;
            // MscStepToGeo
            geom_path = msc_step_to_geo(ctx, pd.particles[pd.electron].mass, true_path,
                                        &alpha);
            if (geom_path > ctx.lambda)
            {
                geom_path = ctx.lambda;
                limited = true;
            }
            t.step_length = geom_path;
            if (limited)
                t.post_action = pd.msc_action;
        }

        // --- PropagationApplier (linear, or uniform field with looping
        // tracks: PropagationApplier.hh:95-195, SimTrackView.hh:189-219)
        if (t.step_length != 0)
        {
            double const orig[3] = {t.geo.pos[0], t.geo.pos[1], t.geo.pos[2]};
            bool boundary;
            bool looping = false;
            double dist;
            if (prob.field_enabled)
            {
                double const mass = pd.particles[t.particle].mass;
                double const momentum = std::sqrt(t.energy * t.energy + 2 * mass * t.energy);
                UniformFieldStepper stepper;
                for (int i = 0; i < 3; ++i)
                    stepper.field[i] = prob.field[i];
                stepper.coeff = pd.particles[t.particle].charge * lorentz_per_cm_tesla;
                FieldPropagation fp = field_propagate(prob.geo, t.geo, stepper, momentum,
                                                      prob.field_options, t.step_length);
                dist = fp.distance;
                boundary = fp.boundary;
                looping = fp.looping;
            }
            else
            {
                dist = geo_find_next_step(prob.geo, t.geo, t.step_length, &boundary);
                if (boundary)
                    geo_move_to_boundary(t.geo);
                else
                    geo_move_internal(t.geo, t.step_length);
            }
            if (t.geo.pos[0] == orig[0] && t.geo.pos[1] == orig[1] && t.geo.pos[2] == orig[2])
            {
                t.status = errored;
                t.along_action = none;
                t.post_action = ids.tracking_cut;
                continue;
            }
            if (prob.field_enabled)
                t.num_looping_steps = looping ? t.num_looping_steps + 1 : 0;
            if (looping)
            {
                t.step_length = dist;
                std::uint32_t const limit = t.energy < prob.looping_threshold_energy
                                                ? prob.looping_max_subthreshold_steps
                                                : prob.looping_max_steps;
                t.post_action = t.num_looping_steps >= limit ? ids.tracking_cut
                                                             : ids.propagation_limit;
            }
            else if (boundary)
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

        // --- MscApplier / UrbanMsc::apply_step
        if (applicable && geom_path > 0 && t.status == alive)
        {
            if (t.post_action == ids.boundary || t.post_action == ids.propagation_limit)
            {
                // MscStepFromGeo
                geom_path = t.step_length;
                double const tstep = msc_step_from_geo(ctx, true_path, alpha, geom_path);
                true_path = tstep;
                displaced = false;
            }
            t.step_length = true_path;
            double safety = 0;
            if (displaced)
            {
                double displ = 0.73 * std::sqrt((true_path - geom_path) * (true_path + geom_path));
                displ = std::max(displ * (1 + 2 * 0.01), msc_geom_limit);
                safety = std::min(geo_find_safety(prob.geo, t.geo), displ);
                if (safety == 0)
                    displaced = false;
            }
            // UrbanMscScatter
            MaterialDef const& m = pd.materials[mat];
            MscDataO::Material const& md = pd.msc.mat[mat];
            double const e_in = t.energy;
            double e_end = 0, tau = 0, xmean = 0, x2mean = 0, theta0 = 0;
            double lmin = t.msc_limit_min;
            bool skip = false;
            if (true_path == ctx.range || true_path < msc_geom_limit)
            {
                skip = true;
            }
            else
            {
                e_end = ctx.end_energy(true_path);
                if (e_end < 1e-6 || true_path <= ctx.lambda * 1e-16)
                    skip = true;
            }
            if (!skip)
            {
                double const lambda_end = ctx.mfp(e_end);
                double const denom = std::fabs(ctx.lambda - lambda_end) < ctx.lambda * 0.01
                                         ? ctx.lambda
                                         : (ctx.lambda - lambda_end)
                                               / std::log(ctx.lambda / lambda_end);
                tau = true_path / denom;
                if (tau < 8)
                {
                    xmean = std::exp(-tau);
                    x2mean = (1 + 2 * std::exp(-2.5 * tau)) / 3;
                    if (lmin == 0)
                        lmin = 1e-8;
                    lmin = std::min(lmin, pd.msc.lambda_limit);
                    double const me = pd.particles[pd.electron].mass;
                    double y = std::max(lmin, true_path) / m.rad_length;
                    if (t.particle == pd.positron)
                    {
                        // UrbanPositronCorrector
                        y *= urban_positron_correction(m.zeff, std::sqrt(e_in * e_end) / me);
                    }
                    theta0 = msc_theta0(md, y, e_in, e_end, me);
                    if (true_path < lmin)
                        theta0 *= std::sqrt(true_path / lmin);
                    theta0 = std::max(theta0, 0.0);
                    if (theta0 < 1e-8)
                    {
                        if (!displaced)
                            skip = true;
                        else
                            theta0 = 0;
                    }
                }
            }
            if (!skip)
            {
                double cost;
                if (theta0 <= 0)
                {
                    cost = 1;
                }
                else if (tau >= 8)
                {
                    cost = uniform_real(-1.0, 1.0, u);
                }
                else if (2 * e_end < e_in || theta0 > pi / 6)
                {
                    cost = msc_simple_scattering(xmean, x2mean, u);
                }
                else
                {
                    MscAngleInput const in{xmean, x2mean, tau, theta0, true_path,
                                           lmin, ctx.lambda, m.rad_length, &md};
                    cost = msc_sample_cos_theta(in, u);
                }
                double phi = uniform_real(0.0, 2 * pi, u);
                bool moved = false;
                double disp[3] = {0, 0, 0};
                if (displaced)
                {
                    double length = 0.73 * std::sqrt((true_path - geom_path) * (true_path + geom_path));
                    length = std::min(length, (1 - 0.01) * safety);
                    if (length >= msc_geom_limit)
                    {
                        double const psi = -std::log(1 - u() * 0.9988703417569197) / 2.160;
                        double const dphi = u() < 0.5 ? phi + psi : phi - psi;
                        double const local[3] = {std::cos(dphi), std::sin(dphi), 0};
                        rotate(local, t.geo.dir, disp);
                        for (auto& v : disp)
                            v *= length;
                        moved = true;
                    }
                }
                double local[3], newdir[3];
                from_spherical(cost, phi, local);
                rotate(local, t.geo.dir, newdir);
                geo_set_dir(prob.geo, t.geo, newdir);
                if (moved)
                {
                    double const x[3] = {t.geo.pos[0] + disp[0], t.geo.pos[1] + disp[1],
                                         t.geo.pos[2] + disp[2]};
                    geo_move_internal_to(t.geo, x);
                }
            }
        }

        // --- TimeUpdater
        {
            double const mass = pd.particles[t.particle].mass;
            double const inv_gamma = mass / (t.energy + mass);
            double const beta_sq = 1 - inv_gamma * inv_gamma;
            double const speed = std::sqrt(beta_sq) * c_light;
            if (speed > 0)
                t.time += t.step_length / speed;
        }

        // --- ElossApplier with FluctELoss / MeanELoss
        if (eloss && t.energy > 0)
        {
            bool const apply_cut = t.post_action != ids.boundary;
            double const lowest = pd.lowest_electron_energy;
            double const e = t.energy;
            double dep;
            if (apply_cut && e < lowest)
            {
                dep = e;
            }
            else
            {
                double ml = t.step_length * eloss->dedx[mat].calc(e);
                if (ml >= e * pd.linear_loss_limit)
                {
                    ml = (t.step_length == t.dedx_range)
                             ? e
                             : e - calc_inverse_range(eloss->range[mat], t.dedx_range - t.step_length);
                }
                double loss = ml;
                if (pd.fluct && loss < e)
                {
                    if (ml >= 1e-5)
                    {
                        double const me = pd.particles[pd.electron].mass;
                        double const mass = pd.particles[t.particle].mass;
                        double const gamma = 1 + e / mass;
                        double const ig = mass / (e + mass);
                        double const beta_sq = 1 - ig * ig;
                        double const two_mebsgs = 2 * me * beta_sq * (gamma * gamma);
                        double max_transfer;
                        if (t.particle == pd.electron)
                        {
                            max_transfer = 0.5 * e;
                        }
                        else
                        {
                            double const ratio = me / mass;
                            max_transfer = two_mebsgs / (1 + ratio * (2 * gamma + ratio));
                        }
                        double const max_energy
                            = std::min(pd.cutoff_energy[pd.electron][mat], max_transfer);
                        if (max_energy > 1e-5)
                        {
                            loss = urban_fluct_sample(pd.fluct_params[mat], pd.materials[mat], ml,
                                                      max_energy, two_mebsgs, beta_sq, u);
                        }
                    }
                    if (loss >= e)
                        loss = apply_cut ? e : ml;
                }
                dep = (apply_cut && e - loss <= lowest) ? e : loss;
            }
            if (dep > 0)
            {
                t.edep += dep;
                t.energy -= dep;
            }
            if (t.energy == 0)
            {
                bool has_at_rest = false;
                for (auto pi_ : procs)
                    has_at_rest = has_at_rest || pd.processes[pi_].cls == "annihilation";
                if (!has_at_rest)
                {
                    t.status = killed;
                    t.post_action = pd.range_action;
                }
                else
                {
                    t.post_action = pd.discrete_action;
                }
            }
        }

        // --- TrackUpdater
        if (t.status == alive && t.post_action != pd.discrete_action)
            t.interaction_mfp -= t.step_length * t.macro_xs;
        ++t.num_steps;
    }
}

//---------------------------------------------------------------------------//
}  // namespace orc
