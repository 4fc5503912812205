//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/AlongStepPhysics.hh
//! Along-step physics of charged tracks: Urban energy-loss fluctuations and
//! Urban multiple scattering (step limit, path conversions, scattering).
//!
//! Restated from global/alongstep/detail/FluctELoss.hh,
//! em/distribution/{EnergyLossHelper,EnergyLossUrbanDistribution,
//! EnergyLossGaussianDistribution}.hh, random/distribution/{Normal,Poisson}
//! Distribution.hh, em/msc/UrbanMsc.hh and em/msc/detail/{UrbanMscHelper,
//! UrbanMscSafetyStepLimit,MscStepToGeo,MscStepFromGeo,UrbanMscScatter,
//! UrbanPositronCorrector}.hh, keeping the order of operations and random
//! draws so that a replayed uniform stream reproduces the results.
//---------------------------------------------------------------------------//
#pragma once

#include <cmath>

#include "corecel/Precision.hh"
#include "em/InteractorCommon.hh"
#include "PhysicsImpl.hh"
#include "random/XorwowEngine.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
namespace alongstep
{
//---------------------------------------------------------------------------//
// DISTRIBUTIONS
//---------------------------------------------------------------------------//
//! Box-Muller normal deviate that keeps its spare (NormalDistribution)
struct Normal
{
    double mean;
    double stddev;
    double spare{0};
    bool has_spare{false};

    template<class Engine>
    CELER_FUNCTION double operator()(Engine& rng)
    {
        if (has_spare)
        {
            has_spare = false;
            return std::fma(spare, stddev, mean);
        }
        double const theta = (2 * phys::pi) * draw_canonical(rng);
        double const r = mth::sqrt(-2 * mth::log(draw_canonical(rng)));
        spare = r * mth::cos(theta);
        has_spare = true;
        return std::fma(r * mth::sin(theta), stddev, mean);
    }
};

//! Poisson deviate: direct product below lambda = 16, else normal
template<class Engine>
CELER_FUNCTION unsigned int sample_poisson(double lambda, Engine& rng)
{
    if (lambda <= 16)
    {
        int k = 0;
        double p = mth::exp(lambda);
        unsigned int iter = 0;
        do
        {
            ++k;
            p *= draw_canonical(rng);
        } while (p > 1 && within_budget(iter, rng));
        return static_cast<unsigned int>(k - 1);
    }
    Normal n{lambda, mth::sqrt(lambda)};
    double const x = n(rng) + 0.5;
    return x > 0 ? static_cast<unsigned int>(x) : 0u;
}

//! Uniform on [a, a + delta)
template<class Engine>
CELER_FORCEINLINE_FUNCTION double uniform(double a, double b, Engine& rng)
{
    return std::fma(b - a, draw_canonical(rng), a);
}

//---------------------------------------------------------------------------//
// URBAN ENERGY-LOSS FLUCTUATIONS
//---------------------------------------------------------------------------//
inline constexpr double fluct_ionization_energy = 1e-5;  //!< [MeV]
inline constexpr double fluct_rate = 0.56;
inline constexpr double fluct_max_collisions = 8;
inline constexpr double fluct_exc_thresh = 42;
inline constexpr double fluct_fwhm_min_energy = 1e-3;  //!< [MeV]

//! Gaussian or flat fast sampling (EnergyLossUrbanDistribution::sample_fast_urban)
template<class Engine>
CELER_FUNCTION double sample_fast_urban(double mean, double stddev, Engine& rng)
{
    if (stddev <= 4 * mean)
    {
        // EnergyLossGaussianDistribution: truncated to (0, 2 mean]
        Normal n{mean, stddev};
        double const max_loss = 2 * mean;
        double r;
        unsigned int iter = 0;
        do
        {
            r = n(rng);
        } while ((r <= 0 || r > max_loss) && within_budget(iter, rng));
        return r;
    }
    return uniform(0.0, 2 * mean, rng);
}

/*!
 * Urban model sample of the energy loss for a mean loss over a step.
 */
template<class Engine>
CELER_FUNCTION double sample_urban_fluct(UrbanFluctuation const& params,
                                         MaterialRecord const& mat,
                                         double unscaled_mean_loss,
                                         double max_energy,
                                         double two_mebsgs,
                                         double beta_sq,
                                         Engine& rng)
{
    double const e0 = fluct_ionization_energy;
    double const loss_scaling
        = 0.5 * (fluct_fwhm_min_energy / max_energy < 1 ? fluct_fwhm_min_energy / max_energy
                                                        : 1.0)
          + 1;
    double const mean_loss = unscaled_mean_loss / loss_scaling;
    double binding[2] = {params.binding_energy[0], params.binding_energy[1]};
    double xs_exc[2] = {0, 0};
    if (max_energy > mat.mean_exc_energy)
    {
        double const w = mth::log(two_mebsgs) - beta_sq;
        double const w0 = mat.log_mean_exc_energy;
        if (w > w0)
        {
            if (w > params.log_binding_energy[1])
            {
                double const c = mean_loss * (1 - fluct_rate) / (w - w0);
                for (int i = 0; i < 2; ++i)
                {
                    xs_exc[i] = c * params.oscillator_strength[i]
                                * (w - params.log_binding_energy[i]) / params.binding_energy[i];
                }
            }
            else
            {
                xs_exc[0] = mean_loss * (1 - fluct_rate) / params.binding_energy[0];
            }
            double scaling = 4;
            if (xs_exc[0] < fluct_exc_thresh)
                scaling = 0.5 + (scaling - 0.5) * mth::sqrt(xs_exc[0] / fluct_exc_thresh);
            binding[0] *= scaling;
            xs_exc[0] /= scaling;
        }
    }
    double xs_ion = mean_loss * (max_energy - e0) / (max_energy * e0 * mth::log(max_energy / e0));
    if (xs_exc[0] + xs_exc[1] > 0)
        xs_ion *= fluct_rate;

    // Excitation
    double result = 0;
    {
        double mean = 0;
        double variance = 0;
        double exc = 0;
        for (int i = 0; i < 2; ++i)
        {
            if (xs_exc[i] > fluct_max_collisions)
            {
                mean += xs_exc[i] * binding[i];
                variance += xs_exc[i] * (binding[i] * binding[i]);
            }
            else if (xs_exc[i] > 0)
            {
                unsigned int const n = sample_poisson(xs_exc[i], rng);
                if (n > 0)
                    exc += uniform(double(n - 1), double(n + 1), rng) * binding[i];
            }
        }
        if (variance > 0)
            exc += sample_fast_urban(mean, mth::sqrt(variance), rng);
        result += exc;
    }
    // Ionisation
    {
        double ion = 0;
        double const energy_ratio = max_energy / e0;
        double alpha = 1;
        double mean_num_coll = 0;
        if (xs_ion > fluct_max_collisions)
        {
            alpha = (xs_ion + fluct_max_collisions) * energy_ratio
                    / (fluct_max_collisions * energy_ratio + xs_ion);
            double const mean_loss_coll = alpha * mth::log(alpha) / (alpha - 1);
            mean_num_coll = xs_ion * energy_ratio * (alpha - 1) / ((energy_ratio - 1) * alpha);
            double const mean = mean_num_coll * mean_loss_coll * e0;
            double const stddev
                = e0 * mth::sqrt(xs_ion * (alpha - mean_loss_coll * mean_loss_coll));
            ion += sample_fast_urban(mean, stddev, rng);
        }
        if (xs_ion > 0 && energy_ratio > alpha)
        {
            unsigned int n = sample_poisson(xs_ion - mean_num_coll, rng);
            double const a = alpha / energy_ratio;
            unsigned int iter = 0;
            for (; n > 0 && within_budget(iter, rng); --n)
                ion += alpha * e0 / uniform(a, 1.0, rng);
        }
        result += ion;
    }
    return loss_scaling * result;
}

/*!
 * Continuous energy loss with fluctuations (FluctELoss::calc_eloss), or the
 * mean loss when fluctuations are off (MeanELoss::calc_eloss).
 */
template<class Engine>
CELER_FUNCTION double calc_eloss(PhysicsParamsRef const& p,
                                 ParticlePhysics const& part,
                                 std::uint32_t pid,
                                 std::uint32_t mat,
                                 double energy,
                                 double range,
                                 double step,
                                 bool apply_cut,
                                 Engine& rng)
{
    double const lowest = p.scalars.lowest_electron_energy;
    if (apply_cut && energy < lowest)
        return energy;
    double eloss = phys::calc_mean_eloss(p, part, mat, energy, range, step);
    if (p.fluct_enabled && eloss < energy)
    {
        // EnergyLossHelper: e+- always use the Urban model above 10 eV
        double const mean_loss = eloss;
        if (mean_loss >= fluct_ionization_energy)
        {
            double const me = p.model.electron_mass;
            double const mass = p.mass[pid];
            double const gamma = 1 + energy / mass;
            double const inv_gamma = mass / (energy + mass);
            double const beta_sq = 1 - inv_gamma * inv_gamma;
            double const two_mebsgs = 2 * me * beta_sq * (gamma * gamma);
            double max_transfer;
            if (pid == p.electron_id)
            {
                max_transfer = 0.5 * energy;
            }
            else
            {
                double const mass_ratio = me / mass;
                max_transfer = two_mebsgs / (1 + mass_ratio * (2 * gamma + mass_ratio));
            }
            double const ecut = p.cutoff_energy[p.electron_id * p.num_materials + mat];
            double const max_energy = ecut < max_transfer ? ecut : max_transfer;
            if (max_energy > fluct_ionization_energy)
            {
                eloss = sample_urban_fluct(p.fluct[mat], p.materials[mat], mean_loss,
                                           max_energy, two_mebsgs, beta_sq, rng);
            }
        }
        if (eloss >= energy)
            eloss = apply_cut ? energy : mean_loss;
    }
    if (apply_cut && energy - eloss <= lowest)
        return energy;
    return eloss;
}

//---------------------------------------------------------------------------//
// URBAN MULTIPLE SCATTERING
//---------------------------------------------------------------------------//
inline constexpr double msc_tau_small = 1e-16;
inline constexpr double msc_tau_big = 8;
inline constexpr double msc_safety_tol = 0.01;
inline constexpr double msc_geom_limit = 5e-9;  //!< 5e-8 mm [cm]
inline constexpr double msc_dtrl = 0.05;
inline constexpr double msc_limit_min_fix = 1e-9;  //!< 0.01 nm [cm]
inline constexpr double msc_limit_min = 1e-8;  //!< 10 limit_min_fix [cm]
inline constexpr double msc_min_step = 1e-7;  //!< 1 nm [cm]
inline constexpr double msc_min_sampling_energy = 1e-6;  //!< [MeV]
inline constexpr double msc_min_scaling_energy = 5e-3;  //!< [MeV]
inline constexpr double msc_safety_plus_rho = 1e-3;  //!< min_range [cm]
inline constexpr double msc_safety_plus_alpha = 0.35;  //!< max_step_over_range

CELER_FORCEINLINE_FUNCTION double fastpow(double a, double b)
{
    return mth::exp(b * mth::log(a));
}

//! Per-track inputs of the MSC helpers (UrbanMscHelper)
struct MscTrack
{
    PhysicsParamsRef const& p;
    ParticlePhysics const& part;
    std::uint32_t mat;
    std::uint32_t par;  //!< 0 = e-, 1 = e+
    double energy;
    double range;  //!< physics range at the pre-step energy
    double lambda;  //!< transport MFP at the pre-step energy

    CELER_FUNCTION double calc_msc_mfp(double e) const
    {
        XsGrid const& g = p.grids[p.msc.xs_grid + 2 * mat + par];
        double const xsec = phys::calc_grid(g, p.grid_reals, e) / (e * e);
        return 1 / xsec;
    }
    CELER_FUNCTION double calc_inverse_range(double r) const
    {
        return phys::calc_inverse_range(p.grids[part.range_grid + mat], p.grid_reals, r);
    }
    CELER_FUNCTION UrbanMscParMat const& pmdata() const
    {
        return p.msc_par_mat[2 * mat + par];
    }
    CELER_FUNCTION double max_step() const { return range * pmdata().d_over_r; }
    CELER_FUNCTION double calc_end_energy(double step) const
    {
        if (step <= range * msc_dtrl)
        {
            double const dedx = phys::calc_grid(p.grids[part.dedx_grid + mat], p.grid_reals, energy);
            return energy - step * dedx;
        }
        return calc_inverse_range(range - step);
    }
};

//! Geometric path of a true path (MscStepToGeo); alpha is returned too
CELER_FUNCTION double to_geo(MscTrack const& t, double tstep, double* alpha_out)
{
    double alpha = 0;
    double step;
    if (tstep < msc_min_step)
    {
        step = tstep;
    }
    else if (tstep < t.range * msc_dtrl)
    {
        step = -t.lambda * mth::expm1(-tstep / t.lambda);
    }
    else
    {
        double mfp_slope;
        if (t.energy < t.p.model.electron_mass || tstep == t.range)
        {
            alpha = 1 / t.range;
            double const s = 1 - alpha * tstep;
            mfp_slope = s > 0 ? s : 0;
        }
        else
        {
            double const rfinal = t.range - tstep;
            double const e1 = t.calc_inverse_range(rfinal);
            double const lambda1 = t.calc_msc_mfp(e1);
            alpha = (t.lambda - lambda1) / (t.lambda * tstep);
            mfp_slope = lambda1 / t.lambda;
        }
        double const w = 1 + 1 / (alpha * t.lambda);
        step = (1 - fastpow(mfp_slope, w)) / (alpha * w);
    }
    *alpha_out = alpha;
    return step < tstep ? step : tstep;
}

//! True path of a geometric path (MscStepFromGeo)
CELER_FUNCTION double
from_geo(MscTrack const& t, double true_step, double alpha, double gstep)
{
    if (gstep < msc_min_step)
        return gstep;
    double tstep;
    if (alpha == 0)
    {
        tstep = -t.lambda * mth::log1p(-gstep / t.lambda);
        if (tstep < msc_min_step)
            tstep = gstep;
    }
    else
    {
        double const w = 1 + 1 / (alpha * t.lambda);
        double x = alpha * w * gstep;
        x = x < 1 ? x : 1.0;
        double const temp = 1 - fastpow(1 - x, 1 / w);
        tstep = temp / alpha;
        tstep = tstep < t.range ? tstep : t.range;
    }
    // clamp(tstep, gstep, true_step)
    return tstep < gstep ? gstep : (tstep > true_step ? true_step : tstep);
}

//! Persistent MSC range quantities of a slot (MscRange)
struct MscRange
{
    double range_init;
    double range_factor;
    double limit_min;
    CELER_FUNCTION explicit operator bool() const
    {
        return range_init > 0 && range_factor > 0 && limit_min > 0;
    }
};

/*!
 * True-path limit of the "minimal" (UrbanMscMinimalStepLimit.hh:53-126) or
 * "safety" / "safety_plus" (UrbanMscSafetyStepLimit.hh:102-213) algorithm;
 * updates the slot's MscRange when needed. Draws only when sampling.
 */
template<class Engine>
CELER_FUNCTION double step_limit(MscTrack const& t,
                                 MscRange& msc_range,
                                 bool on_boundary,
                                 double safety,
                                 double phys_step,
                                 Engine& rng)
{
    PhysicsParamsRef const& p = t.p;
    double const range = t.range;
    double max_step = phys_step;
    double limit_min, limit;
    if (p.msc.step_limit_algorithm == 0)
    {
        if (!msc_range)
        {
            msc_range.range_init = HUGE_VAL;
            msc_range.range_factor = p.msc.range_factor;
            msc_range.limit_min = 10 * msc_limit_min_fix;
        }
        limit_min = msc_range.limit_min;
        if (on_boundary)
        {
            double const init
                = msc_range.range_factor * (range > t.lambda ? range : t.lambda);
            msc_range.range_init = init > limit_min ? init : limit_min;
        }
        limit = msc_range.range_init;
    }
    else
    {
        bool const plus = p.msc.step_limit_algorithm == 2;
        if (!msc_range || on_boundary)
        {
            MscRange r;
            r.range_factor = p.msc.range_factor;
            r.range_init = plus ? range : (range > t.lambda ? range : t.lambda);
            if (t.lambda > p.msc.lambda_limit)
            {
                double const c = plus ? 0.84 : 0.75;
                r.range_factor *= c + (1 - c) * t.lambda / p.msc.lambda_limit;
            }
            // calc_limit_min
            UrbanMscMaterial const& md = p.msc_material[t.mat];
            double const min_mfp = std::fma(
                t.energy, std::fma(t.energy, md.stepmin_coeff[1], md.stepmin_coeff[0]), 2.0);
            double xm = t.lambda / min_mfp;
            xm *= t.pmdata().scaled_zeff;
            if (t.energy < msc_min_scaling_energy)
                xm *= 0.5 + 0.5 * t.energy / msc_min_scaling_energy;
            r.limit_min = xm > msc_limit_min_fix ? xm : msc_limit_min_fix;
            msc_range = r;
        }
        limit_min = msc_range.limit_min;
        limit = range;
        if (safety < range)
        {
            double const a = msc_range.range_factor * msc_range.range_init;
            double const b = p.msc.safety_factor * safety;
            limit = a > b ? a : b;
        }
        limit = limit > limit_min ? limit : limit_min;
        if (plus && range > msc_safety_plus_rho)
        {
            // Scaled step range alpha r + rho (1 - alpha)(2 - rho / r)
            double const limit_step
                = msc_safety_plus_alpha * range
                  + msc_safety_plus_rho * (1 - msc_safety_plus_alpha)
                        * (2 - msc_safety_plus_rho / range);
            max_step = max_step < limit_step ? max_step : limit_step;
        }
    }
    if (max_step <= limit)
        return max_step;
    if (limit == limit_min)
        return limit_min;
    Normal n{limit, 0.1 * (limit - limit_min)};
    double const sampled = n(rng);
    return sampled < limit_min ? limit_min : (sampled > max_step ? max_step : sampled);
}

//! Lateral displacement length (UrbanMscScatter::calc_displacement)
CELER_FORCEINLINE_FUNCTION double calc_displacement(double geom_path, double true_path)
{
    double const rmax2 = (true_path - geom_path) * (true_path + geom_path);
    return 0.73 * mth::sqrt(rmax2);
}

//! Positron correction to the Highland width (UrbanPositronCorrector)
CELER_FUNCTION double positron_correction(double zeff, double y)
{
    double const a = std::fma(zeff, -4.08e-3, 0.994);
    double const iz = 1 / zeff;
    double const b = std::fma(iz, std::fma(iz, 365.0, 52.6), 7.16);
    double const c = std::fma(zeff, -4.47e-3, 1.0);
    double const d = 1.21e-3 * zeff;
    double const mult = std::fma(zeff, std::fma(zeff, 1.84035e-4, -1.86427e-2), 1.41125);
    double const xl = 0.6, xh = 0.9, e = 113;
    double const x = mth::sqrt(y * (y + 2) / ((y + 1) * (y + 1)));
    double corr;
    if (x < xl)
    {
        corr = a * (1 - mth::exp(-b * x));
    }
    else if (x > xh)
    {
        corr = c + d * mth::exp(e * (x - 1));
    }
    else
    {
        double const yl = a * (1 - mth::exp(-b * xl));
        double const yh = c + d * mth::exp(e * (xh - 1));
        double const y0 = (yh - yl) / (xh - xl);
        double const y1 = yl - y0 * xl;
        corr = y0 * x + y1;
    }
    return corr * mult;
}

//! Result of the MSC scattering of one step
struct MscResult
{
    int action;  //!< 0 unchanged, 1 scattered, 2 displaced
    double dir[3];
    double displacement[3];
};

/*!
 * Scatter the direction and displace the position at the end of a step
 * (UrbanMscScatter).
 */
template<class Engine>
CELER_FUNCTION MscResult msc_scatter(MscTrack const& t,
                                     MaterialRecord const& material,
                                     bool is_positron,
                                     double const* inc_dir,
                                     double safety,
                                     bool is_displaced,
                                     double geom_path,
                                     double true_path,
                                     double limit_min_in,
                                     Engine& rng)
{
    PhysicsParamsRef const& p = t.p;
    UrbanMscMaterial const& msc = p.msc_material[t.mat];
    MscResult result;
    result.action = 0;
    result.dir[0] = inc_dir[0];
    result.dir[1] = inc_dir[1];
    result.dir[2] = inc_dir[2];
    result.displacement[0] = result.displacement[1] = result.displacement[2] = 0;

    double const inc_energy = t.energy;
    double end_energy = 0;
    double tau = 0, xmean = 0, x2mean = 0, theta0 = 0;
    double limit_min = limit_min_in;
    bool skip = false;
    if (true_path == t.range || true_path < msc_geom_limit)
    {
        skip = true;
    }
    else
    {
        end_energy = t.calc_end_energy(true_path);
        if (end_energy < msc_min_sampling_energy || true_path <= t.lambda * msc_tau_small)
            skip = true;
    }
    if (!skip)
    {
        double const lambda = t.lambda;
        double const lambda_end = t.calc_msc_mfp(end_energy);
        double const denom = std::fabs(lambda - lambda_end) < lambda * 0.01
                                 ? lambda
                                 : (lambda - lambda_end) / mth::log(lambda / lambda_end);
        tau = true_path / denom;
        if (tau < msc_tau_big)
        {
            xmean = mth::exp(-tau);
            x2mean = (1 + 2 * mth::exp(-2.5 * tau)) / 3;
            if (CELER_UNLIKELY(limit_min == 0))
                limit_min = msc_limit_min;
            limit_min = limit_min < p.msc.lambda_limit ? limit_min : p.msc.lambda_limit;
            // compute_theta0
            double const mass = p.model.electron_mass;
            double const tp = limit_min > true_path ? limit_min : true_path;
            double y = tp / material.rad_length;
            if (is_positron)
                y *= positron_correction(material.zeff,
                                         mth::sqrt(inc_energy * end_energy) / mass);
            double const invbetacp
                = mth::sqrt((inc_energy + mass) * (end_energy + mass)
                            / (inc_energy * (inc_energy + 2 * mass) * end_energy
                               * (end_energy + 2 * mass)));
            theta0 = 13.6 * mth::sqrt(y) * invbetacp;
            theta0 *= std::fma(mth::log(y), msc.theta_coeff[1], msc.theta_coeff[0]);
            if (true_path < limit_min)
                theta0 *= mth::sqrt(true_path / limit_min);
            theta0 = theta0 > 0 ? theta0 : 0;
            if (theta0 < 1e-8)
            {
                if (!is_displaced)
                    skip = true;
                else
                    theta0 = 0;
            }
        }
    }
    if (skip)
        return result;

    // Polar angle
    auto simple_scattering = [&]() {
        double const a = (2 * xmean + 9 * x2mean - 3) / (2 * xmean - 3 * x2mean + 1);
        double const p_pow = (a + 2) * xmean / a;
        double r;
        unsigned int iter = 0;
        do
        {
            double const rdm = draw_canonical(rng);
            bool const use_pow = draw_canonical(rng) < p_pow;
            r = 2 * (use_pow ? fastpow(rdm, 1 / (a + 1)) : rdm) - 1;
        } while (std::fabs(r) > 1 && within_budget(iter, rng));
        return r;
    };
    double costheta;
    if (theta0 <= 0)
    {
        costheta = 1;
    }
    else if (tau >= msc_tau_big)
    {
        costheta = uniform(-1.0, 1.0, rng);
    }
    else if (2 * end_energy < inc_energy || theta0 > phys::pi / 6)
    {
        costheta = simple_scattering();
    }
    else
    {
        // sample_cos_theta
        double const maxtau = true_path < limit_min ? limit_min / t.lambda : tau;
        double const u = fastpow(maxtau, 1.0 / 6);
        double const radlen_mfp = true_path / (tau * material.rad_length);
        double xsi = std::fma(u, std::fma(u, msc.tail_coeff[2], msc.tail_coeff[1]),
                              msc.tail_coeff[0])
                     + msc.tail_corr * mth::log(radlen_mfp);
        xsi = xsi > 1.9 ? xsi : 1.9;
        double const ea = mth::exp(-xsi);
        double const sh = 2 * mth::sin(0.5 * theta0);
        double const x = sh * sh;
        double const xmean_1 = 1 - x * (1 + (xsi * ea) / (1 - ea));
        if (xmean_1 <= 0.999 * xmean)
        {
            costheta = simple_scattering();
        }
        else
        {
            double c = xsi;
            if (std::fabs(xsi - 3) < 0.001)
                c = 3.001;
            else if (std::fabs(xsi - 2) < 0.001)
                c = 2.001;
            double const b1 = 2 + (c - xsi) * x;
            double const d = fastpow(c * x / b1, c - 1);
            double const x0 = 1 - xsi * x;
            double const xmean_2 = (x0 + d - (c * x - b1 * d) / (c - 2)) / (1 - d);
            double const f2x0 = (c - 1) / (c * (1 - d));
            double const prob = f2x0 / (ea / (1 - ea) + f2x0);
            double const qprob = xmean / (prob * xmean_1 + (1 - prob) * xmean_2);
            if (draw_canonical(rng) >= qprob)
            {
                costheta = uniform(-1.0, 1.0, rng);
            }
            else if (draw_canonical(rng) < prob)
            {
                costheta = 1 + mth::log(uniform(ea, 1.0, rng)) * x;
            }
            else
            {
                double var = (1 - d) * draw_canonical(rng);
                if (var < 0.01 * d)
                {
                    var /= (d * (c - 1));
                    costheta = -1 + var * (1 - 0.5 * var * c) * (2 + (c - xsi) * x);
                }
                else
                {
                    costheta = x * (c - xsi - c * fastpow(var + d, -1 / (c - 1))) + 1;
                }
            }
        }
    }
    double phi = uniform(0.0, 2 * phys::pi, rng);

    result.action = 1;
    if (is_displaced)
    {
        double length = calc_displacement(geom_path, true_path);
        double const lim = (1 - msc_safety_tol) * safety;
        length = length < lim ? length : lim;
        if (length >= msc_geom_limit)
        {
            // sample_displacement_dir
            double const cbeta = 2.160;
            double const cbeta1 = 0.9988703417569197;
            double const psi = -mth::log(1 - draw_canonical(rng) * cbeta1) / cbeta;
            double const dphi = draw_canonical(rng) < 0.5 ? phi + psi : phi - psi;
            double const local[3] = {mth::cos(dphi), mth::sin(dphi), 0};
            rotate(local, inc_dir, result.displacement);
            for (int i = 0; i < 3; ++i)
                result.displacement[i] *= length;
            result.action = 2;
        }
    }
    double local[3];
    from_spherical(costheta, phi, local);
    rotate(local, inc_dir, result.dir);
    return result;
}

//---------------------------------------------------------------------------//
}  // namespace alongstep
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
