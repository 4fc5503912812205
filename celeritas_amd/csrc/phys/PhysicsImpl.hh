//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/PhysicsImpl.hh
//! Cross-section lookup and step-limit helpers used by the kernels.
//---------------------------------------------------------------------------//
#pragma once

#include <cmath>

#include "corecel/Precision.hh"
#include "global/CoreData.hh"
#include "LivermoreXs.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
namespace phys
{
//---------------------------------------------------------------------------//
inline constexpr double c_light = 29979245800.0;  //!< [cm/s]
inline constexpr double pi = 3.14159265358979323846;

/*!
 * An energy and its bracket on a log-uniform grid: the log and the two grid
 * energies (one FP64 log and two exps) are the costly part of a lookup, and
 * grids with the same front and spacing share them. Reusing them repeats
 * exactly the same operations, so results are bit-identical.
 */
struct GridPoint
{
    double energy;
    double loge;
    double front{HUGE_VAL};  //!< grid front/spacing of the cached bracket
    double delta{HUGE_VAL};
    std::uint32_t lower{0};
    double lower_e{0};
    double upper_e{0};

    CELER_FUNCTION explicit GridPoint(double e) : energy(e), loge(mth::log(e)) {}
};

/*!
 * Interpolate a log-uniform grid at an energy (grid/XsCalculator.hh:107-150).
 *
 * Linear in E between the two bracketing points; values at and above the
 * prime index are stored multiplied by E. Outside the grid the end value is
 * used (with the 1/E scaling where applicable).
 */
CELER_FUNCTION double
calc_grid(XsGrid const& g, double const* __restrict__ reals, GridPoint& gp)
{
    double const* v = reals + g.offset;
    double const energy = gp.energy;
    double const loge = gp.loge;
    if (!(loge > g.loge_front))  // also catches NaN: never index with it
    {
        double r = v[0];
        return (0 >= g.prime_index) ? r / energy : r;
    }
    if (loge >= g.loge_back)
    {
        std::uint32_t const i = g.size - 1;
        double r = v[i];
        return (i >= g.prime_index) ? r / energy : r;
    }
    if (!(gp.front == g.loge_front && gp.delta == g.loge_delta))
    {
        gp.front = g.loge_front;
        gp.delta = g.loge_delta;
        gp.lower = static_cast<std::uint32_t>((loge - g.loge_front) / g.loge_delta);
        gp.upper_e = mth::exp(g.loge_front + g.loge_delta * (gp.lower + 1));
        gp.lower_e = mth::exp(g.loge_front + g.loge_delta * gp.lower);
    }
    std::uint32_t lower = gp.lower;
    double upper_e = gp.upper_e;
    double lower_e = gp.lower_e;
    if (CELER_UNLIKELY(lower + 2 > g.size))
    {
        // A log a few ulps below the stored back can round to the last
        // point: use the last interval
        lower = g.size - 2;
        upper_e = mth::exp(g.loge_front + g.loge_delta * (lower + 1));
        lower_e = mth::exp(g.loge_front + g.loge_delta * lower);
    }
    double upper_xs = v[lower + 1];
    if (lower + 1 == g.prime_index)
        upper_xs /= upper_e;
    double const lower_xs = v[lower];
    double const slope = ((-lower_xs) + upper_xs) / ((-lower_e) + upper_e);
    double result = std::fma(slope, (-lower_e) + energy, lower_xs);
    if (lower >= g.prime_index)
        result /= energy;
    return result;
}

CELER_FUNCTION double
calc_grid(XsGrid const& g, double const* __restrict__ reals, double energy)
{
    GridPoint gp(energy);
    return calc_grid(g, reals, gp);
}

inline constexpr double r_electron = 2.8179403262e-13;  //!< [cm]

//! Heitler e+ annihilation macroscopic cross section [1/cm]
//! (em/xs/EPlusGGMacroXsCalculator.hh)
CELER_FUNCTION double
eplusgg_macro_xs(double electron_mass, double electron_density, double energy)
{
    double const gamma = (energy > 1e-6 ? energy : 1e-6) / electron_mass;
    double const sqrt_gg2 = mth::sqrt(gamma * (gamma + 2));
    double const gp1 = gamma + 1;
    double const poly = 1 + gp1 * (4 + gp1 * 1);
    return pi * r_electron * r_electron * electron_density
           * (poly * mth::log(gamma + 1 + sqrt_gg2) - (gamma + 4) * sqrt_gg2)
           / (gamma * (gamma + 2) * (gamma + 2));
}

//! Macroscopic cross section of a process in a material
CELER_FUNCTION double calc_xs(PhysicsParamsRef const& p,
                              ProcessRecord const& proc,
                              std::uint32_t mat,
                              GridPoint& gp)
{
    double const energy = gp.energy;
    if (proc.xs_grid != invalid_id)
        return calc_grid(p.grids[proc.xs_grid + mat], p.grid_reals, gp);
    if (proc.process_class == static_cast<std::uint32_t>(ProcessClass::annihilation))
    {
        return eplusgg_macro_xs(p.model.electron_mass,
                                p.materials[mat].electron_density, energy);
    }
    if (proc.process_class == static_cast<std::uint32_t>(ProcessClass::photoelectric))
        return livermore_pe_macro_xs(p, mat, energy);
    return 0;
}

CELER_FUNCTION double calc_xs(PhysicsParamsRef const& p,
                              ProcessRecord const& proc,
                              std::uint32_t mat,
                              double energy)
{
    GridPoint gp(energy);
    return calc_xs(p, proc, mat, gp);
}

//! Upper bound of the XS over the step (PhysicsTrackView.hh:446-478)
CELER_FUNCTION double calc_max_xs(PhysicsParamsRef const& p,
                                  ProcessRecord const& proc,
                                  std::uint32_t mat,
                                  GridPoint& gp)
{
    double const energy = gp.energy;
    double const emax = p.grid_reals[proc.energy_max_xs + mat];
    double const exi = energy * p.scalars.min_eprime_over_e;
    if (emax >= exi && emax < energy)
        return calc_xs(p, proc, mat, emax);
    double const a = calc_xs(p, proc, mat, gp);
    double const b = calc_xs(p, proc, mat, exi);
    return a > b ? a : b;
}

/*!
 * Range at an energy (grid/RangeCalculator.hh:80-110): sqrt(E/Emin) scaling
 * below the grid, clipped above it, linear in E between grid points.
 */
CELER_FUNCTION double
calc_range(XsGrid const& g, double const* __restrict__ reals, double energy)
{
    double const* v = reals + g.offset;
    double const loge = mth::log(energy);
    if (!(loge > g.loge_front))  // also catches NaN: never index with it
        return v[0] * mth::exp(0.5 * (loge - g.loge_front));
    if (loge >= g.loge_back)
        return v[g.size - 1];
    auto i = static_cast<std::uint32_t>((loge - g.loge_front) / g.loge_delta);
    i = i + 2 > g.size ? g.size - 2 : i;
    double const xl = mth::exp(g.loge_front + g.loge_delta * i);
    double const xr = mth::exp(g.loge_front + g.loge_delta * (i + 1));
    double const slope = ((-v[i]) + v[i + 1]) / ((-xl) + xr);
    return std::fma(slope, (-xl) + energy, v[i]);
}

/*!
 * Energy for a range (grid/InverseRangeCalculator.hh:87-117): quadratic
 * below the first range value, linear in E between tabulated ranges.
 */
CELER_FUNCTION double
calc_inverse_range(XsGrid const& g, double const* __restrict__ reals, double range)
{
    double const* r = reals + g.offset;
    std::uint32_t const n = g.size;
    if (!(range >= r[0]))  // also catches NaN: never index with it
    {
        double const f = range / r[0];
        return mth::exp(g.loge_front) * (f * f);
    }
    if (range >= r[n - 1])
        return mth::exp(g.loge_back);
    std::uint32_t lo = 0;
    std::uint32_t len = n;
    while (len > 0)
    {
        std::uint32_t const half = len / 2;
        if (r[lo + half] < range)
        {
            lo += half + 1;
            len -= half + 1;
        }
        else
        {
            len = half;
        }
    }
    std::uint32_t i = lo;
    if (r[i] != range)
        --i;
    double const el = mth::exp(g.loge_front + g.loge_delta * i);
    double const er = mth::exp(g.loge_front + g.loge_delta * (i + 1));
    double const slope = ((-el) + er) / ((-r[i]) + r[i + 1]);
    return std::fma(slope, (-r[i]) + range, el);
}

/*!
 * Mean energy loss over a step (PhysicsStepUtils.hh:177-233): dE/dx times
 * the step, or the inverse-range difference when that exceeds the linear
 * loss limit.
 */
CELER_FUNCTION double calc_mean_eloss(PhysicsParamsRef const& p,
                                      ParticlePhysics const& part,
                                      std::uint32_t mat,
                                      double energy,
                                      double range,
                                      double step)
{
    double eloss = step * calc_grid(p.grids[part.dedx_grid + mat], p.grid_reals, energy);
    if (eloss >= energy * p.scalars.linear_loss_limit)
    {
        if (step == range)
            return energy;
        eloss = energy
                - calc_inverse_range(p.grids[part.range_grid + mat], p.grid_reals, range - step);
    }
    return eloss;
}

//! Scaled range limit (PhysicsTrackView.hh:624-640)
CELER_FUNCTION double range_to_step(PhysicsScalars const& sc, double range)
{
    double const rho = sc.min_range;
    // 1 + sqrt(epsilon) for double precision
    if (range < rho * (1 + 1.4901161193847656e-08))
        return range;
    double const alpha = sc.max_step_over_range;
    return alpha * range + rho * (1 - alpha) * (2 - rho / range);
}

//! Speed [cm/s] (ParticleTrackView.hh:303-323)
CELER_FUNCTION double speed(double energy, double mass)
{
    double const inv_gamma = mass / (energy + mass);
    double const beta_sq = 1 - inv_gamma * inv_gamma;
    return mth::sqrt(beta_sq) * c_light;
}

//---------------------------------------------------------------------------//
}  // namespace phys
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
