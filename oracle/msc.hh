//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/msc.hh
//! Urban multiple scattering pieces as standalone functions, restated from
//! em/params/UrbanMscParams.cc (material coefficients),
//! em/msc/detail/UrbanPositronCorrector.hh, MscStepToGeo.hh,
//! MscStepFromGeo.hh, UrbanMscScatter.hh (theta0, sample_cos_theta,
//! simple_scattering) and UrbanMscHelper.hh.
//!
//! Pinned by (tests/test_oracle_msc.py):
//! - the steel coefficients of UrbanMsc.test.cc:139-167 (stepmin, theta,
//!   tail coefficients, tail correction, d_over_r for Zeff of
//!   G4_STAINLESS-STEEL from four-steel-slabs.gdml:65-72);
//! - the positron correction values of UrbanMsc.test.cc:55-80 (Z = 1, 74);
//! - the reference's step_conversion invariants (UrbanMsc.test.cc:188-283:
//!   geometric <= true step, true step in [gstep, pstep], true -> geo ->
//!   true to 1e-10) on the generated steel tables, plus a closed form of
//!   the small-step regime;
//! - the two moments of simple_scattering (mean e^-tau and second moment
//!   (1 + 2 e^-2.5 tau)/3, which fix its mixture parameter) and the means
//!   of the three components of sample_cos_theta, against integrals of
//!   their densities.
//---------------------------------------------------------------------------//
#pragma once

#include <algorithm>
#include <cmath>

#include "physics.hh"

namespace orc
{
//---------------------------------------------------------------------------//
//! Per-track inputs of the MSC helpers (UrbanMscHelper)
struct MscCtx
{
    PhysicsData const& pd;
    ProcessDef const& eloss;
    std::uint32_t mat;
    int par;  // 0 e-, 1 e+
    double energy;
    double range;
    double lambda = 0;

    //! Transport MFP at an energy (the tabulated sigma_tr E^2 over E^2)
    double mfp(double e) const { return 1 / (pd.msc.xs[par][mat].calc(e) / (e * e)); }
    double inverse_range(double r) const { return calc_inverse_range(eloss.range[mat], r); }
    //! Energy at the end of a true path (UrbanMscHelper::calc_end_energy)
    double end_energy(double step) const
    {
        if (step <= range * 0.05)
            return energy - step * eloss.dedx[mat].calc(energy);
        return inverse_range(range - step);
    }
};

//! Deliberate mistake planted by the tests (0 in normal use): 1 theta
//! coefficient sign, 2 positron correction interval edge, 3 inverse of the
//! constant-MFP regime, 4 inverse of the linear-MFP regime, 5
//! simple-scattering mixture weight, 6 tail-model g1 width, 7 geometric path
//! of the constant-MFP regime
extern int g_msc_mutation;

inline constexpr double msc_geom_limit = 5e-9;  // [cm]
inline constexpr double msc_limit_min_fix = 1e-9;  // [cm]
inline constexpr double msc_min_step = 1e-7;  // [cm]

//! Box-Muller normal deviate keeping its spare (NormalDistribution.hh)
template<class U>
struct NormalDist
{
    double mean, stddev;
    double spare = 0;
    bool has_spare = false;
    double operator()(U& u)
    {
        if (has_spare)
        {
            has_spare = false;
            return std::fma(spare, stddev, mean);
        }
        double const twopi = 2 * 3.14159265358979323846;
        double const theta = twopi * u();
        double const r = std::sqrt(-2 * std::log(u()));
        spare = r * std::cos(theta);
        has_spare = true;
        return std::fma(r * std::sin(theta), stddev, mean);
    }
};

//! Persistent MSC range of a track (MscRange; invalid until first set)
struct MscRangeO
{
    double range_init = 0;
    double range_factor = 0;
    double limit_min = 0;
    bool valid() const { return range_init > 0 && range_factor > 0 && limit_min > 0; }
};

//! Inputs of one MSC true-path limit
struct MscLimitInput
{
    MscLimitAlg alg = MscLimitAlg::safety;
    double energy = 0;  //!< [MeV]
    double range = 0;  //!< continuous-slowing-down range [cm]
    double lambda = 0;  //!< MSC transport MFP [cm]
    MscDataO::Material const* md = nullptr;
    int par = 0;  //!< 0 e-, 1 e+
    double range_factor = 0.04;
    double safety_factor = 0.6;
    double lambda_limit = 0.1;  //!< [cm]
    bool on_boundary = false;
    double safety = 0;  //!< [cm]
    double phys_step = 0;  //!< [cm]
};

/*!
 * True-path limit: UrbanMscMinimalStepLimit.hh:53-126 ("minimal") or
 * UrbanMscSafetyStepLimit.hh:102-213 ("safety", "safety_plus"). Updates the
 * track's MSC range; draws two uniforms when it samples the limit.
 */
template<class U>
double msc_step_limit(MscLimitInput const& in, MscRangeO& r, U& u)
{
    double max_step = in.phys_step;
    double limit_min, limit;
    if (in.alg == MscLimitAlg::minimal)
    {
        if (!r.valid())
        {
            r.range_init = HUGE_VAL;
            r.range_factor = in.range_factor;
            r.limit_min = 10 * msc_limit_min_fix;
        }
        limit_min = r.limit_min;
        if (in.on_boundary)
        {
            double const init = r.range_factor * std::max(in.range, in.lambda);
            r.range_init = std::max(init, limit_min);
        }
        limit = r.range_init;
    }
    else
    {
        bool const plus = in.alg == MscLimitAlg::safety_plus;
        if (!r.valid() || in.on_boundary)
        {
            double factor = in.range_factor;
            double const init = plus && g_msc_mutation != 11
                                    ? in.range
                                    : std::max(in.range, in.lambda);
            if (in.lambda > in.lambda_limit)
            {
                double const c = plus && g_msc_mutation != 12 ? 0.84 : 0.75;
                factor *= c + (1 - c) * in.lambda / in.lambda_limit;
            }
            // calc_limit_min (G4UrbanMscModel::ComputeTlimitmin)
            MscDataO::Material const& md = *in.md;
            double const e = in.energy;
            double xm = in.lambda / std::fma(e, std::fma(e, md.stepmin[1], md.stepmin[0]), 2.0);
            xm *= md.scaled_zeff[in.par];
            if (e < 5e-3)
                xm *= 0.5 + 0.5 * e / 5e-3;
            r.range_init = init;
            r.range_factor = factor;
            r.limit_min = std::max(xm, msc_limit_min_fix);
        }
        limit_min = r.limit_min;
        limit = in.range;
        if (in.safety < in.range)
            limit = std::max(r.range_factor * r.range_init, in.safety_factor * in.safety);
        limit = std::max(limit, limit_min);
        if (plus && g_msc_mutation != 13)
        {
            // Scaled step range alpha r + rho (1 - alpha)(2 - rho / r)
            double const rho = 1e-3;  // [cm]
            double const alpha = 0.35;
            if (in.range > rho)
            {
                double const limit_step
                    = alpha * in.range + rho * (1 - alpha) * (2 - rho / in.range);
                max_step = std::min(max_step, limit_step);
            }
        }
    }
    if (max_step <= limit)
        return max_step;
    if (limit == limit_min)
        return limit_min;
    NormalDist<U> n{limit, 0.1 * (limit - limit_min)};
    return std::clamp(n(u), limit_min, max_step);
}


//! Material coefficients (UrbanMscParams::calc_material_data and the
//! per-particle scaled Zeff and d_over_r)
MscDataO::Material msc_material_coeffs(double zeff);

//! Positron correction of theta0 (UrbanPositronCorrector)
double urban_positron_correction(double zeff, double y);

//! Geometric path of a true path (MscStepToGeo); sets alpha
double msc_step_to_geo(MscCtx const& ctx, double electron_mass, double tstep, double* alpha);

//! True path of a geometric path (MscStepFromGeo), clamped to [gstep, tstep]
double msc_step_from_geo(MscCtx const& ctx, double true_path, double alpha, double gstep);

//! Width of the central part of the angular distribution (compute_theta0);
//! y = path / radiation length, corrected for positrons by the caller
double msc_theta0(MscDataO::Material const& md,
                  double y,
                  double e_in,
                  double e_end,
                  double electron_mass);

//! Large-angle scattering from two model functions (simple_scattering)
template<class U>
double msc_simple_scattering(double xmean, double x2mean, U& u);

//! Inputs of the polar-angle sampling with the tail model
struct MscAngleInput
{
    double xmean;  //!< e^-tau
    double x2mean;  //!< (1 + 2 e^-2.5 tau) / 3
    double tau;
    double theta0;
    double true_path;
    double limit_min;
    double lambda;
    double rad_length;
    MscDataO::Material const* md;
};

//! Polar cosine from the three-component model (sample_cos_theta)
template<class U>
double msc_sample_cos_theta(MscAngleInput const& in, U& u);

//---------------------------------------------------------------------------//
// INLINE DEFINITIONS
//---------------------------------------------------------------------------//
namespace detail
{
inline double fastpow(double a, double b)
{
    return std::exp(b * std::log(a));
}
template<class U>
double uniform_real(double a, double b, U& u)
{
    return std::fma(b - a, u(), a);
}
}  // namespace detail

template<class U>
double msc_simple_scattering(double xmean, double x2mean, U& u)
{
    double const a = (2 * xmean + 9 * x2mean - 3) / (2 * xmean - 3 * x2mean + 1);
    double const ppow = (a + 2) * xmean / (g_msc_mutation == 5 ? a + 1 : a);
    double r;
    do
    {
        double const rdm = u();
        r = 2 * (u() < ppow ? detail::fastpow(rdm, 1 / (a + 1)) : rdm) - 1;
    } while (std::fabs(r) > 1);
    return r;
}

template<class U>
double msc_sample_cos_theta(MscAngleInput const& in, U& u)
{
    MscDataO::Material const& md = *in.md;
    double const maxtau = in.true_path < in.limit_min ? in.limit_min / in.lambda : in.tau;
    double const uu = detail::fastpow(maxtau, 1.0 / 6);
    double const radlen_mfp = in.true_path / (in.tau * in.rad_length);
    double xsi = std::fma(uu, std::fma(uu, md.tail[2], md.tail[1]), md.tail[0])
                 + md.tail_corr * std::log(radlen_mfp);
    xsi = std::max(xsi, 1.9);
    double const ea = std::exp(-xsi);
    double const sh = 2 * std::sin(0.5 * in.theta0);
    double const x = sh * sh;
    // The reference's estimate of the mean of g1 (UrbanMscScatter.hh:438)
    double const xmean_1 = 1 - x * (1 + (xsi * ea) / (1 - ea));
    if (xmean_1 <= 0.999 * in.xmean)
        return msc_simple_scattering(in.xmean, in.x2mean, u);
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
    double const qprob = in.xmean / (prob * xmean_1 + (1 - prob) * xmean_2);
    if (u() >= qprob)
        return detail::uniform_real(-1.0, 1.0, u);
    if (u() < prob)
        return 1 + std::log(detail::uniform_real(ea, 1.0, u)) * (g_msc_mutation == 6 ? 0.5 * x : x);
    double var = (1 - d) * u();
    if (var < 0.01 * d)
    {
        var /= (d * (c - 1));
        return -1 + var * (1 - 0.5 * var * c) * (2 + (c - xsi) * x);
    }
    return x * (c - xsi - c * detail::fastpow(var + d, -1 / (c - 1))) + 1;
}
}  // namespace orc
