//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/LivermoreXs.hh
//! Livermore photoelectric cross sections (computed on the fly).
//!
//! Follows em/xs/LivermorePEMicroXsCalculator.hh, grid/GenericCalculator.hh
//! and phys/MacroXsCalculator.hh: above thresh_lo a cumulative polynomial in
//! 1/E (the last shell's row is the total), from the K edge to thresh_lo a
//! table of sigma E^3, below the K edge a second table; tables interpolate
//! linearly and clamp at their ends.
//---------------------------------------------------------------------------//
#pragma once

#include <cmath>

#include "corecel/Precision.hh"
#include "PhysicsData.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
namespace phys
{
//---------------------------------------------------------------------------//
//! GenericCalculator::operator()
CELER_FUNCTION double
generic_calc(GenericGrid const& g, double const* __restrict__ reals, double x)
{
    double const* xs = reals + g.x;
    double const* ys = reals + g.y;
    std::uint32_t const n = g.size;
    if (!(x > xs[0]))  // also catches NaN: never index with it
        return ys[0];
    if (x >= xs[n - 1])
        return ys[n - 1];
    // NonuniformGrid::find: lower_bound, one back unless exact
    std::uint32_t lo = 0;
    std::uint32_t len = n;
    while (len > 0)
    {
        std::uint32_t const half = len / 2;
        if (xs[lo + half] < x)
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
    if (xs[i] != x)
        --i;
    double const slope = ((-ys[i]) + ys[i + 1]) / ((-xs[i]) + xs[i + 1]);
    return std::fma(slope, (-xs[i]) + x, ys[i]);
}

//! PolyEvaluator<real_type, 5>: Horner with fma
CELER_FORCEINLINE_FUNCTION double poly5(double const* a, double x)
{
    double r = std::fma(x, a[5], a[4]);
    r = std::fma(x, r, a[3]);
    r = std::fma(x, r, a[2]);
    r = std::fma(x, r, a[1]);
    return std::fma(x, r, a[0]);
}

//! Microscopic photoelectric cross section [barn]
CELER_FUNCTION double livermore_pe_micro_xs(PhysicsParamsRef const& p,
                                            std::uint32_t element,
                                            double inc_energy)
{
    LivermoreElement const& el = p.pe_elements[element];
    LivermoreShell const* shells = p.pe_shells + el.shell_begin;
    LivermoreShell const& last = shells[el.num_shells - 1];
    double const energy = inc_energy > last.binding_energy ? inc_energy
                                                           : last.binding_energy;
    double const inv_energy = 1. / energy;
    if (energy >= el.thresh_lo)
    {
        double const* param = last.param[energy < el.thresh_hi ? 0 : 1];
        return inv_energy * poly5(param, inv_energy);
    }
    double const inv3 = inv_energy * inv_energy * inv_energy;
    if (energy >= shells[0].binding_energy)
        return inv3 * generic_calc(el.xs_hi, p.reals, energy);
    return inv3 * generic_calc(el.xs_lo, p.reals, energy);
}

//! Macroscopic photoelectric cross section [1/cm]
CELER_FUNCTION double
livermore_pe_macro_xs(PhysicsParamsRef const& p, std::uint32_t mat, double energy)
{
    MaterialRecord const& m = p.materials[mat];
    double result = 0.;
    for (std::uint32_t k = 0; k < m.num_elements; ++k)
    {
        ElementComponent const& c = p.components[m.elem_begin + k];
        result += livermore_pe_micro_xs(p, c.element, energy) * c.fraction;
    }
    return result * 1e-24 * m.number_density;
}

//---------------------------------------------------------------------------//
}  // namespace phys
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
