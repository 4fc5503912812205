//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file em/Brems.hh
//! Electron and positron bremsstrahlung: Seltzer-Berger tabulated DCS,
//! relativistic (Tsai) DCS and the combined model.
//!
//! Follows em/interactor/{SeltzerBerger,RelativisticBrem,CombinedBrem}
//! Interactor.hh, em/interactor/detail/{SBEnergySampler,RBEnergySampler,
//! BremFinalStateHelper,SBPositronXsCorrector}.hh,
//! em/distribution/{SBEnergyDistHelper,SBEnergyDistribution}.hh and
//! em/xs/RBDiffXsCalculator.hh: the photon energy is drawn from a
//! reciprocal distribution in k^2 + (density correction) and accepted
//! against the DCS bound, then the Tsai-Urban polar angle and an azimuth
//! place the photon and momentum conservation the outgoing lepton. The
//! draw order is the reference's, so a replayed uniform stream reproduces
//! its outputs. The relativistic DCS applies the LPM suppression (with
//! dielectric suppression) when the physics enables it and the total
//! energy exceeds the material's threshold (RBDiffXsCalculator.hh:120-135).
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Precision.hh"
#include "LPM.hh"
#include "StandardEm.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
namespace brems
{
//---------------------------------------------------------------------------//
inline constexpr double lambdabar_electron = 3.8615926796e-11;  //!< [cm]
inline constexpr double alpha_fine = 7.2973525693e-3;
inline constexpr double sb_upper_limit = 1e3;  //!< [MeV]
inline constexpr double high_energy_limit = 1e8;  //!< [MeV]

using lpm::migdal_constant;

struct GridInterp
{
    std::uint32_t index;
    double fraction;
};

/*!
 * Lower index and fraction on a nonuniform grid: lower_bound, then one step
 * back unless the value is a grid point (NonuniformGrid::find). Values
 * outside [front, back) are clamped to the end intervals.
 */
CELER_FUNCTION GridInterp find_interp(double const* g, std::uint32_t n, double v)
{
    std::uint32_t lo = 0;
    std::uint32_t len = n;
    while (len > 0)
    {
        std::uint32_t const half = len / 2;
        if (g[lo + half] < v)
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
    if (i >= n || g[i] != v)
        i = (i == 0) ? 0 : i - 1;
    if (i + 1 >= n)
        i = n - 2;
    return {i, (v - g[i]) / (g[i + 1] - g[i])};
}

CELER_FORCEINLINE_FUNCTION double
positron_invbeta(double inc_energy, double mass, double gamma_energy)
{
    double const e = inc_energy - gamma_energy;
    return (e + mass) / mth::sqrt(e * (e + 2 * mass));
}

//! Tsai's DCS without LPM (RBDiffXsCalculator::dxsec_per_atom, Z >= 5)
CELER_FUNCTION double
rb_dxsec(ElementRecord const& el, double total_energy, double gamma_energy)
{
    double const y = gamma_energy / total_energy;
    double const onemy = 1 - y;
    double const term0 = onemy + 0.75 * (y * y);
    if (el.z < 5)
    {
        // Dirac-Fock model
        double const d = term0 * el.rb_factor1 + onemy * el.rb_factor2;
        return d > 0 ? d : 0;
    }
    double const invz = 1 / el.z;
    double const term1 = y / (total_energy - gamma_energy);
    double const gam = term1 * el.rb_gamma_factor;
    double const eps = term1 * el.rb_epsilon_factor;
    double const gam2 = gam * gam;
    double const eps2 = eps * eps;
    double const phi1 = 16.863 - 2 * mth::log(1 + 0.311877 * gam2)
                        + 2.4 * mth::exp(-0.9 * gam) + 1.6 * mth::exp(-1.5 * gam);
    double const phi2 = 2 / (3 + 19.5 * gam + 18 * gam2);
    double const psi1 = 24.34 - 2 * mth::log(1 + 13.111641 * eps2)
                        + 2.8 * mth::exp(-8 * eps) + 1.2 * mth::exp(-29.2 * eps);
    double const psi2 = 2 / (3 + 120 * eps + 1200 * eps2);
    double const dxsec
        = term0
              * ((0.25 * phi1 - el.rb_fz) + (0.25 * psi1 - 2 * el.log_z / 3) * invz)
          + 0.125 * onemy * (phi2 + psi2 * invz);
    return dxsec > 0 ? dxsec : 0;
}

/*!
 * Relativistic DCS of a lepton of fixed total energy (RBDiffXsCalculator):
 * Tsai's form, or with LPM suppression above the material's threshold
 * E_LPM sqrt(n_e 4 pi r_e lambdabar^2) when enabled.
 */
struct RbDxsec
{
    ElementRecord const& el;
    MaterialRecord const& mat;
    double total_energy;
    double density_corr;
    double lpm_energy;
    bool use_lpm;

    CELER_FUNCTION
    RbDxsec(ElementRecord const& e, MaterialRecord const& m, bool enable_lpm, double total)
        : el(e), mat(m), total_energy(total)
    {
        double const density_factor = m.electron_density * migdal_constant();
        density_corr = density_factor * (total * total);
        lpm_energy = m.rad_length * lpm::lpm_constant();
        use_lpm = enable_lpm && total > lpm_energy * mth::sqrt(density_factor);
    }

    CELER_FUNCTION double operator()(double gamma_energy) const
    {
        if (!use_lpm)
            return rb_dxsec(el, total_energy, gamma_energy);
        double const epsilon = total_energy / gamma_energy;
        lpm::Functions const f = lpm::calc_functions(
            el.cbrt_z, mat.electron_density, lpm_energy, true, gamma_energy, epsilon);
        double const y = gamma_energy / total_energy;
        double const onemy = 1 - y;
        double const y2 = 0.25 * (y * y);
        double const term = f.xi * (y2 * f.g + (onemy + 2 * y2) * f.phi);
        double const dxsec = term * el.rb_factor1 + onemy * el.rb_factor2;
        return dxsec > 0 ? dxsec : 0;
    }
};

//! Photon energy from the relativistic DCS (RBEnergySampler)
template<class Engine>
CELER_FUNCTION double sample_rb_energy(ElementRecord const& el,
                                       MaterialRecord const& mat,
                                       bool enable_lpm,
                                       double inc_energy,
                                       double mass,
                                       double kc,
                                       Engine& rng)
{
    double const total = inc_energy + mass;
    RbDxsec const calc_dxsec(el, mat, enable_lpm, total);
    double const dens = calc_dxsec.density_corr;
    double const tmin = kc < inc_energy ? kc : inc_energy;
    double const tmax = high_energy_limit < inc_energy ? high_energy_limit : inc_energy;
    double const a = tmin * tmin + dens;
    double const logratio = mth::log((1 / a) * (tmax * tmax + dens));
    double const max_xs = el.rb_factor1 + el.rb_factor2;
    double k, dsigma;
    unsigned int iter = 0;
    do
    {
        k = mth::sqrt(a * mth::exp(logratio * draw_canonical(rng)) - dens);
#if CELER_PREC_IS_MIXED
        // FP32 exp/log round trips may overshoot the upper edge by ~1e-7
        k = k < tmax ? k : tmax;
#endif
        dsigma = calc_dxsec(k);
    } while (dsigma < max_xs * draw_canonical(rng) && within_budget(iter, rng));
    return k;
}

//! Photon energy from the tabulated SB DCS (SBEnergySampler)
template<class Engine>
CELER_FUNCTION double sample_sb_energy(PhysicsParamsRef const& p,
                                       ElementRecord const& el,
                                       double electron_density,
                                       double inc_energy,
                                       double mass,
                                       double kc,
                                       bool is_electron,
                                       Engine& rng)
{
    double const* xg = p.reals + el.sb_x;
    double const* yg = p.reals + el.sb_y;
    double const* vals = p.reals + el.sb_values;
    std::uint32_t const* argmax = p.sb_argmax + el.sb_argmax;
    std::uint32_t const nx = el.sb_nx;
    std::uint32_t const ny = el.sb_ny;
    double const total = inc_energy + mass;
    double const dens = electron_density * migdal_constant() * (total * total);
    GridInterp const xl = find_interp(xg, nx, mth::log(inc_energy));
    double const* row0 = vals + std::size_t(xl.index) * ny;
    double const* row1 = row0 + ny;
    double const max_xs = (1 - xl.fraction) * row0[argmax[xl.index]]
                          + xl.fraction * row1[argmax[xl.index + 1]];
    double const inv_e = 1 / inc_energy;
    double const a = kc * kc + dens;
    double const logratio = mth::log((1 / a) * (inc_energy * inc_energy + dens));
    double const alpha_z = 2 * phys::pi * alpha_fine * el.z;
    double const cutoff_invbeta
        = is_electron ? 0 : positron_invbeta(inc_energy, mass, kc);
    double exit, xs;
    unsigned int iter = 0;
    do
    {
        exit = mth::sqrt(a * mth::exp(logratio * draw_canonical(rng)) - dens);
#if CELER_PREC_IS_MIXED
        // FP32 exp/log round trips may overshoot the upper edge by ~1e-7
        exit = exit < inc_energy ? exit : inc_energy;
#endif
        GridInterp const yl = find_interp(yg, ny, exit * inv_e);
        std::uint32_t const j = yl.index;
        double const v = (1 - xl.fraction)
                             * ((1 - yl.fraction) * row0[j] + yl.fraction * row0[j + 1])
                         + xl.fraction
                               * ((1 - yl.fraction) * row1[j] + yl.fraction * row1[j + 1]);
        double scale = 1;
        if (!is_electron)
        {
            double const delta = cutoff_invbeta - positron_invbeta(inc_energy, mass, exit);
            scale = mth::exp(alpha_z * (delta < 0 ? delta : 0.0));
        }
        xs = v * scale;
    } while (xs < max_xs * draw_canonical(rng) && within_budget(iter, rng));
    return exit;
}

//---------------------------------------------------------------------------//
}  // namespace brems

/*!
 * Sample a bremsstrahlung photon; which = 0 (SB), 1 (relativistic) or
 * 2 (combined: SB below 1 GeV).
 */
template<class Engine>
CELER_FUNCTION Interaction sample_brems(PhysicsParamsRef const& p,
                                        CoreStateRef const& s,
                                        std::uint32_t particle,
                                        std::uint32_t material,
                                        std::uint32_t element,
                                        double energy,
                                        double const* dir,
                                        int which,
                                        Engine& rng)
{
    Interaction r;
    std::uint32_t const start = allocate_secondaries(s, 1);
    if (start == invalid_id)
    {
        r.action = Interaction::failed;
        return r;
    }
    MaterialRecord const& m = p.materials[material];
    ElementRecord const& el = p.elements[p.components[m.elem_begin + element].element];
    double const mass = p.mass[particle];
    double const kc = p.cutoff_energy[p.gamma_id * p.num_materials + material];
    bool const use_rb = which == 1 || (which == 2 && energy >= brems::sb_upper_limit);
    double const k
        = use_rb ? brems::sample_rb_energy(el, m, p.scalars.enable_lpm != 0, energy, mass,
                                           kc, rng)
                 : brems::sample_sb_energy(p, el, m.electron_density, energy, mass, kc,
                                           particle == p.electron_id, rng);
    double const costheta = sample_tsai_urban(energy, mass, rng);
    // BremFinalStateHelper
    double const inc_momentum = mth::sqrt(energy * energy + 2 * mass * energy);
    Secondary& sec = s.secondaries.storage[start];
    sample_exiting_direction(costheta, dir, rng, sec.dir);
    sec.particle_id = p.gamma_id;
    sec.energy = k;
    r.energy = energy - k;
    momentum_direction(inc_momentum, dir, k, sec.dir, r.dir);
    r.sec_start = start;
    r.num_secondaries = 1;
    return r;
}

//---------------------------------------------------------------------------//
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
