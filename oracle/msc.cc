//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/msc.cc
//---------------------------------------------------------------------------//
#include "msc.hh"

#include <algorithm>

namespace orc
{
//! Deliberate mistakes for the tests that show the pins catch them (0: none)
int g_msc_mutation = 0;

namespace
{
double poly2(double c0, double c1, double c2, double x)
{
    return std::fma(x, std::fma(x, c2, c1), c0);
}
}  // namespace

//---------------------------------------------------------------------------//
MscDataO::Material msc_material_coeffs(double zeff)
{
    // UrbanMscParams::calc_material_data: step minimum, Highland theta0
    // correction and tail parameters as functions of Zeff (the Urban
    // model's fits, Geant4 PRM §8.1)
    MscDataO::Material d;
    d.stepmin[0] = 1e3 * 27.725 / (1 + 0.203 * zeff);
    d.stepmin[1] = 1e3 * 6.152 / (1 + 0.111 * zeff);
    double const z16 = detail::fastpow(zeff, 1.0 / 6.0);
    double const fz = poly2(0.990395, -0.168386, 0.093286, z16);
    d.theta[0] = fz * (1 - 8.7780e-2 / zeff);
    d.theta[1] = fz * (4.0780e-2 + (g_msc_mutation == 1 ? -1.7315e-4 : 1.7315e-4) * zeff);
    double const z13 = z16 * z16;
    d.tail[0] = poly2(2.3785, -4.1981e-1, 6.3100e-2, z13);
    d.tail[1] = poly2(4.7526e-1, 1.7694, -3.3885e-1, z13);
    d.tail[2] = poly2(2.3683e-1, -1.8111, 3.2774e-1, z13);
    d.tail_corr = poly2(1.7888e-2, 1.9659e-2, -2.6664e-3, z13);
    // calc_par_mat_data: scaled Zeff for the minimum step (e-, e+) and the
    // maximum step over range
    d.scaled_zeff[0] = 0.87 * detail::fastpow(zeff, 2.0 / 3);
    d.scaled_zeff[1] = 0.70 * detail::fastpow(zeff, 1.0 / 2);
    d.d_over_r = 9.6280e-1 - 8.4848e-2 * std::sqrt(zeff) + 4.3769e-3 * zeff;
    return d;
}

//---------------------------------------------------------------------------//
double urban_positron_correction(double z, double y)
{
    double const a = std::fma(z, -4.08e-3, 0.994);
    double const b = std::fma(1 / z, std::fma(1 / z, 365.0, 52.6), 7.16);
    double const c = std::fma(z, -4.47e-3, 1.0);
    double const d = 1.21e-3 * z;
    double const mult = std::fma(z, std::fma(z, 1.84035e-4, -1.86427e-2), 1.41125);
    double const x = std::sqrt(y * (y + 2) / ((y + 1) * (y + 1)));
    double corr;
    if (x < 0.6)
    {
        corr = a * (1 - std::exp(-b * x));
    }
    else if (x > 0.9)
    {
        corr = c + d * std::exp(113 * (x - 1));
    }
    else
    {
        double const yl = a * (1 - std::exp(-b * 0.6));
        double const xh = g_msc_mutation == 2 ? 0.6 : 0.9;
        double const yh = c + d * std::exp(113 * (xh - 1));
        double const y0 = (yh - yl) / (0.9 - 0.6);
        corr = y0 * x + (yl - y0 * 0.6);
    }
    return corr * mult;
}

//---------------------------------------------------------------------------//
double msc_step_to_geo(MscCtx const& ctx, double electron_mass, double tstep, double* alpha)
{
    double step;
    *alpha = 0;
    if (tstep < msc_min_step)
    {
        // Very small: geometric and true paths coincide
        step = tstep;
    }
    else if (tstep < ctx.range * 0.05)
    {
        // Small with respect to the range: constant MFP
        step = g_msc_mutation == 7 ? ctx.lambda * std::log1p(tstep / ctx.lambda)
                                   : -ctx.lambda * std::expm1(-tstep / ctx.lambda);
    }
    else
    {
        // MFP linear in the path length
        double slope;
        if (ctx.energy < electron_mass || tstep == ctx.range)
        {
            *alpha = 1 / ctx.range;
            slope = std::max(1 - *alpha * tstep, 0.0);
        }
        else
        {
            double const e1 = ctx.inverse_range(ctx.range - tstep);
            double const lambda1 = ctx.mfp(e1);
            *alpha = (ctx.lambda - lambda1) / (ctx.lambda * tstep);
            slope = lambda1 / ctx.lambda;
        }
        double const w = 1 + 1 / (*alpha * ctx.lambda);
        step = (1 - detail::fastpow(slope, w)) / (*alpha * w);
    }
    return std::min(step, tstep);
}

//---------------------------------------------------------------------------//
double msc_step_from_geo(MscCtx const& ctx, double true_path, double alpha, double gstep)
{
    if (gstep < msc_min_step)
        return gstep;
    double tstep;
    if (alpha == 0)
    {
        tstep = g_msc_mutation == 3 ? ctx.lambda * std::expm1(gstep / ctx.lambda)
                                    : -ctx.lambda * std::log1p(-gstep / ctx.lambda);
        if (tstep < msc_min_step)
            tstep = gstep;
    }
    else
    {
        double const w = 1 + 1 / (alpha * ctx.lambda);
        double const x = std::min(alpha * (g_msc_mutation == 4 ? 1.0 : w) * gstep, 1.0);
        tstep = std::min((1 - detail::fastpow(1 - x, 1 / w)) / alpha, ctx.range);
    }
    return std::clamp(tstep, gstep, true_path);
}

//---------------------------------------------------------------------------//
double msc_theta0(MscDataO::Material const& md,
                  double y,
                  double e_in,
                  double e_end,
                  double me)
{
    // Highland formula with the Urban model's correction in log(y)
    double const invbetacp = std::sqrt((e_in + me) * (e_end + me)
                                       / (e_in * (e_in + 2 * me) * e_end * (e_end + 2 * me)));
    double theta0 = 13.6 * std::sqrt(y) * invbetacp;
    theta0 *= std::fma(std::log(y), md.theta[1], md.theta[0]);
    return theta0;
}
}  // namespace orc

extern "C" void orc_msc_set_mutation(int m)
{
    orc::g_msc_mutation = m;
}
