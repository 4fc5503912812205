//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see em.hh).
//---------------------------------------------------------------------------//
#include "em.hh"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "problem.hh"

namespace orc
{
namespace
{
constexpr double pi = 3.14159265358979323846;
constexpr double r_electron = 2.8179403262e-13;  // [cm] (CODATA 2014)
constexpr double alpha_fine = 7.2973525693e-3;  // CODATA 2018

double dot3(double const* a, double const* b)
{
    double r = 0;
    for (int i = 0; i < 3; ++i)
        r = std::fma(a[i], b[i], r);
    return r;
}

//! Exiting direction at polar cosine mu about dir (InteractionUtils.hh)
template<class Engine>
void exiting_direction(double mu, double const* dir, Engine& rng, double* out)
{
    double phi = std::fma(2 * pi, rng(), 0.0);
    double local[3];
    from_spherical(mu, phi, local);
    rotate(local, dir, out);
}

//! Momentum-conserving exiting direction (calc_exiting_direction)
void momentum_direction(double p_in, double const* d_in, double p_out,
                        double const* d_out, double* out)
{
    for (int i = 0; i < 3; ++i)
        out[i] = d_in[i] * p_in - d_out[i] * p_out;
    make_unit(out);
}

//! TsaiUrbanDistribution (em/distribution/TsaiUrbanDistribution.hh)
template<class Engine>
double tsai_urban(double energy, double mass, Engine& rng)
{
    double umax = 2 * (1 + energy / mass);
    double u;
    do
    {
        double a = rng();
        double b = rng();
        double uu = -std::log(a * b);
        u = uu * (rng() < 0.25 ? 1.6 : 1.6 / 3);
    } while (u > umax);
    double r = u / umax;
    return 1 - 2 * r * r;
}

//---------------------------------------------------------------------------//
// KLEIN-NISHINA (KleinNishinaInteractor.hh:104-190)
//---------------------------------------------------------------------------//
template<class Engine>
Interaction klein_nishina(InteractorInput const& in, State& s, Engine& rng)
{
    PhysicsData const& pd = *in.pd;
    Interaction r;
    std::uint32_t start;
    SecondaryP* sec = allocate_secondaries(s, 1, &start);
    if (!sec)
    {
        r.action = Interaction::failed;
        return r;
    }
    double const inv_mass = 1 / pd.particles[pd.electron].mass;
    double const E = in.energy;
    double const k = E * inv_mass;
    double const eps0 = 1 / (1 + 2 * k);
    double const a = -std::log(eps0);
    double const b = 0.5 * (1 - eps0 * eps0);
    double const p_f1 = a / (a + b);
    double const logratio = std::log(1.0 * eps0);
    double const f2_lo = eps0 * eps0;
    double const f2_delta = 1 - f2_lo;
    double eps, omc;
    while (true)
    {
        double eps_sq;
        if (rng() < p_f1)
        {
            eps = std::exp(logratio * rng());
            eps_sq = eps * eps;
        }
        else
        {
            eps_sq = std::fma(f2_delta, rng(), f2_lo);
            eps = std::sqrt(eps_sq);
        }
        omc = (1 - eps) / (eps * k);
        double sin2 = omc * (2 - omc);
        double reject = eps * sin2 / (1 + eps_sq);
        if (!(rng() < reject))
            break;
    }
    r.energy = eps * E;
    exiting_direction(1 - omc, in.dir, rng, r.dir);
    r.sec_start = start;
    r.num_secondaries = 1;
    sec->energy = E - r.energy;
    if (sec->energy < 1e-4)
    {
        r.energy_deposition = sec->energy;
        *sec = SecondaryP{};
        return r;
    }
    sec->particle = pd.electron;
    momentum_direction(E, in.dir, r.energy, r.dir, sec->dir);
    return r;
}

//---------------------------------------------------------------------------//
// BETHE-HEITLER (BetheHeitlerInteractor.hh:165-321), LPM off below 100 GeV
//---------------------------------------------------------------------------//
double bh_f1(double d)
{
    return d > 1.4 ? 42.038 - 8.29 * std::log(d + 0.958)
                   : 42.184 - d * (7.444 - 1.623 * d);
}
double bh_f2(double d)
{
    return d > 1.4 ? 42.038 - 8.29 * std::log(d + 0.958)
                   : 41.326 - d * (5.848 - 0.902 * d);
}

//! Screening functions Phi_1, Phi_2 (BetheHeitlerInteractor::screening_phi1_phi2)
struct BhPhi
{
    double phi1, phi2;
};
BhPhi bh_phi1_phi2(double d)
{
    BhPhi r;
    if (d > 1.4)
    {
        r.phi1 = 21.0190 - 4.145 * std::log(d + 0.958);
        r.phi2 = r.phi1;
    }
    else
    {
        r.phi1 = 20.806 - d * (3.190 - 0.5710 * d);
        r.phi2 = 20.234 - d * (2.126 - 0.0903 * d);
    }
    return r;
}

template<class Engine>
Interaction bethe_heitler(InteractorInput const& in, State& s, Engine& rng)
{
    PhysicsData const& pd = *in.pd;
    Interaction r;
    std::uint32_t start;
    SecondaryP* sec = allocate_secondaries(s, 2, &start);
    if (!sec)
    {
        r.action = Interaction::failed;
        return r;
    }
    MaterialDef const& mat = pd.materials[in.material];
    ElementDef const& el = pd.elements[mat.elements[in.element].first];
    double const me = pd.particles[pd.electron].mass;
    double const E = in.energy;
    double const eps0 = me / E;
    double const half = 0.5;
    double eps;
    if (E < 2)
    {
        eps = std::fma(half - eps0, rng(), eps0);
    }
    else
    {
        double const delta_min = 4 * 136 / el.cbrt_z * eps0;
        double f_z = 8.0 / 3.0 * el.log_z;
        if (E > 50)
            f_z += 8 * el.coulomb_correction;
        double const delta_max = std::exp((42.038 - f_z) / 8.29) - 0.958;
        double const eps1 = half - half * std::sqrt(1 - delta_min / delta_max);
        double const eps_min = std::max(eps0, eps1);
        double const f10 = bh_f1(delta_min) - f_z;
        double const f20 = bh_f2(delta_min) - f_z;
        double const st = (half - eps_min) * (half - eps_min) * f10;
        double const sf = 1.5 * f20;
        double const p_f1 = st / (st + sf);
        // LPM above 100 GeV when enabled; no dielectric suppression
        // (BetheHeitlerInteractor.hh:105-109,147-153, BetheHeitlerData.hh:51)
        bool const lpm = pd.enable_lpm && E > 1e5;
        double g;
        do
        {
            if (rng() < p_f1)
            {
                eps = half - (half - eps_min) * std::cbrt(rng());
                double delta = 136 / el.cbrt_z * eps0 / (eps * (1 - eps));
                if (lpm)
                {
                    BhPhi const sc = bh_phi1_phi2(delta);
                    LpmFunctions const f = lpm_functions(el, mat, false, E, eps);
                    g = f.xi * ((2 * f.phi + f.g) * sc.phi1 - f.g * sc.phi2 - f.phi * f_z)
                        / f10;
                }
                else
                {
                    g = (bh_f1(delta) - f_z) / f10;
                }
            }
            else
            {
                eps = eps_min + (half - eps_min) * rng();
                double delta = 136 / el.cbrt_z * eps0 / (eps * (1 - eps));
                if (lpm)
                {
                    BhPhi const sc = bh_phi1_phi2(delta);
                    LpmFunctions const f = lpm_functions(el, mat, false, E, eps);
                    g = f.xi
                        * ((f.phi + half * f.g) * sc.phi1 + half * f.g * sc.phi2
                           - half * (f.g + f.phi) * f_z)
                        / f20;
                }
                else
                {
                    g = (bh_f2(delta) - f_z) / f20;
                }
            }
        } while (g < rng());
    }
    r.action = Interaction::absorbed;
    r.energy = 0;
    r.sec_start = start;
    r.num_secondaries = 2;
    sec[0].particle = pd.electron;
    sec[1].particle = pd.positron;
    sec[0].energy = (1 - eps) * E - me;
    sec[1].energy = eps * E - me;
    if (rng() < half)
        std::swap(sec[0].energy, sec[1].energy);
    double phi = std::fma(2 * pi, rng(), 0.0);
    double local[3];
    double cost = tsai_urban(sec[0].energy, me, rng);
    from_spherical(cost, phi, local);
    rotate(local, in.dir, sec[0].dir);
    cost = tsai_urban(sec[1].energy, me, rng);
    from_spherical(cost, phi + pi, local);
    rotate(local, in.dir, sec[1].dir);
    return r;
}

//---------------------------------------------------------------------------//
// MOLLER-BHABHA (MollerBhabhaInteractor.hh:119-157 + distributions)
//---------------------------------------------------------------------------//
double moller_g(double gamma, double eps)
{
    double two_gamma_term = (2 * gamma - 1) / (gamma * gamma);
    double comp = 1 - eps;
    return 1 - two_gamma_term * eps
           + eps * eps
                 * (1 - two_gamma_term
                    + (1 - two_gamma_term * comp) / (comp * comp));
}

double bhabha_g(double gamma, double eps_min, double eps_max)
{
    double y = 1 / (1 + gamma);
    double y_sq = y * y;
    double one_minus_2y = 1 - 2 * y;
    double b1 = 2 - y_sq;
    double b2 = one_minus_2y * (3 + y_sq);
    double b4 = one_minus_2y * one_minus_2y * one_minus_2y;
    double b3 = one_minus_2y * one_minus_2y + b4;
    double beta_sq = 1 - 1 / (gamma * gamma);
    double em2 = eps_max * eps_max;
    return 1
           + (em2 * em2 * b4 - eps_min * eps_min * eps_min * b3 + em2 * b2
              - eps_min * b1)
                 * beta_sq;
}

template<class Engine>
Interaction moller_bhabha(InteractorInput const& in, State& s, Engine& rng)
{
    PhysicsData const& pd = *in.pd;
    Interaction r;
    std::uint32_t start;
    SecondaryP* sec = allocate_secondaries(s, 1, &start);
    if (!sec)
    {
        r.action = Interaction::failed;
        return r;
    }
    double const me = pd.particles[pd.electron].mass;
    double const E = in.energy;
    double const cut = pd.cutoff_energy[pd.electron][in.material];
    bool const is_electron = in.particle == pd.electron;
    double const min_frac = cut / E;
    double const gamma = 1 + E / me;
    double eps;
    if (is_electron)
    {
        double const max_frac = 0.5;
        double const g_den = moller_g(gamma, max_frac);
        double const a = 1 / max_frac;
        double const delta = 1 / min_frac - a;
        do
        {
            eps = 1 / std::fma(delta, rng(), a);
        } while (g_den * rng() > moller_g(gamma, eps));
    }
    else
    {
        double const max_frac = 1;
        double const g_den = bhabha_g(gamma, min_frac, max_frac);
        double const a = 1 / max_frac;
        double const delta = 1 / min_frac - a;
        do
        {
            eps = 1 / std::fma(delta, rng(), a);
        } while (g_den * rng() > bhabha_g(gamma, eps, eps));
    }
    double const e_sec = E * eps;
    double const inc_momentum = std::sqrt(E * E + 2 * me * E);
    // IoniFinalStateHelper (em/interactor/detail/IoniFinalStateHelper.hh)
    double const momentum = std::sqrt(e_sec * (e_sec + 2 * me));
    double const costheta = e_sec * (E + me + me) / (momentum * inc_momentum);
    sec->energy = e_sec;
    exiting_direction(costheta, in.dir, rng, sec->dir);
    sec->particle = pd.electron;
    r.energy = E - e_sec;
    momentum_direction(inc_momentum, in.dir, momentum, sec->dir, r.dir);
    r.sec_start = start;
    r.num_secondaries = 1;
    return r;
}

//---------------------------------------------------------------------------//
// POSITRON ANNIHILATION (EPlusGGInteractor.hh:99-170)
//---------------------------------------------------------------------------//
template<class Engine>
Interaction eplusgg(InteractorInput const& in, State& s, Engine& rng)
{
    PhysicsData const& pd = *in.pd;
    Interaction r;
    std::uint32_t start;
    SecondaryP* sec = allocate_secondaries(s, 2, &start);
    if (!sec)
    {
        r.action = Interaction::failed;
        return r;
    }
    double const me = pd.particles[pd.electron].mass;
    double const E = in.energy;
    r.action = Interaction::absorbed;
    r.energy = 0;
    r.sec_start = start;
    r.num_secondaries = 2;
    sec[0].particle = sec[1].particle = pd.gamma;
    if (E == 0)
    {
        sec[0].energy = sec[1].energy = me;
        double costheta = std::fma(2.0, rng(), -1.0);
        double phi = std::fma(2 * pi, rng(), 0.0);
        from_spherical(costheta, phi, sec[0].dir);
        for (int i = 0; i < 3; ++i)
            sec[1].dir[i] = -sec[0].dir[i];
        return r;
    }
    double const half = 0.5;
    double const tau = E / me;
    double const tau2 = tau + 2;
    double const sqgrate = std::sqrt(tau / tau2) * half;
    double const lo = half - sqgrate;
    double const logratio = std::log((1 / lo) * (half + sqgrate));
    double epsil;
    do
    {
        epsil = lo * std::exp(logratio * rng());
    } while (rng() < epsil - (2 * (tau + 1) * epsil - 1) / (epsil * tau2 * tau2));
    double const cost = (epsil * tau2 - 1) / (epsil * std::sqrt(tau * tau2));
    double const total = E + 2 * me;
    double const eg = epsil * total;
    double const eplus_moment = std::sqrt(E * total);
    sec[0].energy = eg;
    exiting_direction(cost, in.dir, rng, sec[0].dir);
    sec[1].energy = total - eg;
    momentum_direction(eplus_moment, in.dir, E, in.dir, sec[1].dir);
    return r;
}

//---------------------------------------------------------------------------//
// BREMSSTRAHLUNG
//---------------------------------------------------------------------------//
constexpr double lambdabar_electron = 3.8615926796e-11;  // [cm] (CODATA 2018)
constexpr double sb_upper_limit = 1e3;  // [MeV] (PhysicsConstants.hh:62-65)
constexpr double high_energy_limit = 1e8;  // [MeV]

struct Interp
{
    std::size_t index;
    double fraction;
};

//! Lower grid index and fraction (corecel/grid/FindInterp.hh and
//! NonuniformGrid::find: lower_bound, stepping back unless exact)
Interp find_interp(std::vector<double> const& g, double v)
{
    if (!(v >= g.front() && v < g.back()))
        throw std::runtime_error("oracle: value outside interpolation grid");
    auto it = std::lower_bound(g.begin(), g.end(), v);
    std::size_t i = static_cast<std::size_t>(it - g.begin());
    if (*it != v)
        --i;
    return {i, (v - g[i]) / (g[i + 1] - g[i])};
}

double positron_invbeta(double inc_energy, double mass, double gamma_energy)
{
    double const e = inc_energy - gamma_energy;
    return (e + mass) / std::sqrt(e * (e + 2 * mass));
}

//! BremFinalStateHelper (em/interactor/detail/BremFinalStateHelper.hh)
template<class Engine>
Interaction brem_final_state(InteractorInput const& in, double mass,
                             double gamma_energy, double costheta,
                             SecondaryP* sec, std::uint32_t start, Engine& rng)
{
    PhysicsData const& pd = *in.pd;
    double const E = in.energy;
    double const inc_momentum = std::sqrt(E * E + 2 * mass * E);
    exiting_direction(costheta, in.dir, rng, sec->dir);
    sec->particle = pd.gamma;
    sec->energy = gamma_energy;
    Interaction r;
    r.energy = E - gamma_energy;
    momentum_direction(inc_momentum, in.dir, gamma_energy, sec->dir, r.dir);
    r.sec_start = start;
    r.num_secondaries = 1;
    return r;
}

//! RBEnergySampler (em/interactor/detail/RBEnergySampler.hh) over
//! RBDiffXsCalculator (LPM when the model enables it and the total energy
//! is above the material's threshold)
template<class Engine>
double sample_rb_energy(ElementDef const& el, MaterialDef const& mat, bool enable_lpm,
                        double inc_energy, double mass, double kc, Engine& rng)
{
    double const total = inc_energy + mass;
    RbDxsec const dxsec(el, mat, enable_lpm, total);
    double const dens = dxsec.density_corr;
    double const tmin = std::min(kc, inc_energy);
    double const tmax = std::min(high_energy_limit, inc_energy);
    double const a = tmin * tmin + dens;
    double const logratio = std::log((1 / a) * (tmax * tmax + dens));
    double const max_xs = el.rb_factor1 + el.rb_factor2;
    double k, dsigma;
    do
    {
        k = std::sqrt(a * std::exp(logratio * rng()) - dens);
        dsigma = dxsec(k);
    } while (dsigma < max_xs * rng());
    return k;
}

//! Shared body of the SB, relativistic and combined brems interactors
//! (SeltzerBergerInteractor.hh:142-163, RelativisticBremInteractor.hh,
//! CombinedBremInteractor.hh:129-170)
template<class Engine>
Interaction brems(InteractorInput const& in, State& s, Engine& rng, int which)
{
    PhysicsData const& pd = *in.pd;
    Interaction r;
    std::uint32_t start;
    SecondaryP* sec = allocate_secondaries(s, 1, &start);
    if (!sec)
    {
        r.action = Interaction::failed;
        return r;
    }
    MaterialDef const& mat = pd.materials[in.material];
    ElementDef const& el = pd.elements[mat.elements[in.element].first];
    double const mass = pd.particles[in.particle].mass;
    double const kc = pd.cutoff_energy[pd.gamma][in.material];
    bool const use_rb = which == 1 || (which == 2 && in.energy >= sb_upper_limit);
    double k;
    if (use_rb)
        k = sample_rb_energy(el, mat, pd.enable_lpm, in.energy, mass, kc, rng);
    else
        k = sample_sb_energy(el, mat.electron_density, in.energy, mass, kc,
                             in.particle == pd.electron ? 0 : 1, rng);
    double const costheta = tsai_urban(in.energy, mass, rng);
    return brem_final_state(in, mass, k, costheta, sec, start, rng);
}

//---------------------------------------------------------------------------//
// LIVERMORE PHOTOELECTRIC (LivermorePEInteractor.hh:145-358)
//---------------------------------------------------------------------------//
double poly5(double const* a, double x)
{
    double r = std::fma(x, a[5], a[4]);
    for (int i = 3; i >= 0; --i)
        r = std::fma(x, r, a[i]);
    return r;
}

template<class Engine>
Interaction livermore_pe(InteractorInput const& in, State& s, Engine& rng)
{
    PhysicsData const& pd = *in.pd;
    Interaction r;
    std::uint32_t start;
    SecondaryP* sec = allocate_secondaries(s, 1, &start);
    if (!sec)
    {
        r.action = Interaction::failed;
        return r;
    }
    MaterialDef const& mat = pd.materials[in.material];
    double const E = in.energy;
    std::size_t comp = in.element;
    if (mat.elements.size() > 1)
    {
        // ElementSelector (mat/ElementSelector.hh) by micro cross section
        double total = 0;
        for (auto const& c : mat.elements)
            total += c.second * livermore_pe_micro_xs(pd.elements[c.first], E);
        double accum = -total * rng();
        std::size_t i = 0;
        for (; i + 1 < mat.elements.size(); ++i)
        {
            accum += mat.elements[i].second
                     * livermore_pe_micro_xs(pd.elements[mat.elements[i].first], E);
            if (accum > 0)
                break;
        }
        comp = i;
    }
    ElementDef const& el = pd.elements[mat.elements[comp].first];
    LivermoreO const& pe = el.pe;
    double const inv_energy = 1 / E;
    std::size_t shell = 0;
    double const cutoff = rng() * livermore_pe_micro_xs(el, E);
    if (E < pe.thresh_lo)
    {
        double xs = 0;
        double const inv_cube = inv_energy * inv_energy * inv_energy;
        for (; shell < pe.shells.size(); ++shell)
        {
            if (E < pe.shells[shell].binding)
                continue;
            xs += inv_cube * pe.shells[shell].xs(E);
            if (xs > cutoff)
                break;
        }
    }
    else
    {
        int const pidx = E < pe.thresh_hi ? 0 : 1;
        for (; shell + 1 < pe.shells.size(); ++shell)
        {
            double xs = inv_energy * poly5(pe.shells[shell].param[pidx], inv_energy);
            if (xs > cutoff)
                break;
        }
    }
    r.action = Interaction::absorbed;
    r.energy = 0;
    r.sec_start = start;
    r.num_secondaries = 1;
    if (shell == pe.shells.size())
    {
        r.energy_deposition = E;
        r.num_secondaries = 0;
        return r;
    }
    double const binding = pe.shells[shell].binding;
    sec->particle = pd.electron;
    sec->energy = E - binding;
    // Sauter-Gavrila direction (sample_direction)
    if (E > 100.)
    {
        for (int i = 0; i < 3; ++i)
            sec->dir[i] = in.dir[i];
    }
    else
    {
        double const epm = std::max(1e-6, E) / pd.particles[pd.electron].mass;
        double const gamma = epm + 1;
        double const beta = std::sqrt(epm * (gamma + 1)) / gamma;
        double const a = (1 - beta) / beta;
        double const b = 0.5 * beta * gamma * epm * (gamma - 2);
        double const g_max = 2 * (1 / a + b);
        double g, nu;
        do
        {
            double const u = rng();
            nu = 2 * a * (2 * u + (a + 2) * std::sqrt(u)) / ((a + 2) * (a + 2) - 4 * u);
            g = (2 - nu) * (1 / (a + nu) + b);
        } while (g < g_max * rng());
        exiting_direction(1 - nu, in.dir, rng, sec->dir);
    }
    r.energy_deposition = binding;
    return r;
}

//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
double GenericO::operator()(double v) const
{
    if (v <= x.front())
        return y.front();
    if (v >= x.back())
        return y.back();
    auto it = std::lower_bound(x.begin(), x.end(), v);
    std::size_t i = static_cast<std::size_t>(it - x.begin());
    if (*it != v)
        --i;
    double const slope = ((-y[i]) + y[i + 1]) / ((-x[i]) + x[i + 1]);
    return std::fma(slope, (-x[i]) + v, y[i]);
}

LivermoreO make_livermore(Json const& j)
{
    auto generic = [](Json const& t) {
        GenericO g;
        g.x = t("x").nums();
        g.y = t("y").nums();
        if (g.x.size() < 2 || g.x.size() != g.y.size())
            throw std::runtime_error("oracle: bad Livermore table");
        return g;
    };
    LivermoreO r;
    r.thresh_lo = j("thresh_lo").n;
    r.thresh_hi = j("thresh_hi").n;
    if (auto lo = j.find("xs_lo"))
        r.xs_lo = generic(*lo);
    r.xs_hi = generic(j("xs_hi"));
    for (auto const& sj : j("shells").a)
    {
        LivermoreO::Shell sh;
        sh.binding = sj("binding_energy").n;
        auto lo = sj("param_lo").nums();
        auto hi = sj("param_hi").nums();
        for (int i = 0; i < 6; ++i)
        {
            sh.param[0][i] = lo.at(i);
            sh.param[1][i] = hi.at(i);
        }
        sh.xs = generic(sj);
        r.shells.push_back(sh);
    }
    return r;
}

double livermore_pe_micro_xs(ElementDef const& el, double inc_energy)
{
    LivermoreO const& pe = el.pe;
    if (!pe)
        throw std::runtime_error("oracle: element has no Livermore data");
    double const energy = std::max(inc_energy, pe.shells.back().binding);
    double const inv_energy = 1. / energy;
    if (energy >= pe.thresh_lo)
    {
        auto const& param = pe.shells.back().param[energy < pe.thresh_hi ? 0 : 1];
        return inv_energy * poly5(param, inv_energy);
    }
    double const inv3 = inv_energy * inv_energy * inv_energy;
    if (energy >= pe.shells.front().binding)
        return inv3 * pe.xs_hi(energy);
    return inv3 * pe.xs_lo(energy);
}

double livermore_pe_macro_xs(PhysicsData const& pd, std::uint32_t mat, double energy)
{
    MaterialDef const& m = pd.materials[mat];
    double result = 0.;
    for (auto const& c : m.elements)
        result += livermore_pe_micro_xs(pd.elements[c.first], energy) * c.second;
    return result * 1e-24 * m.number_density;
}

//---------------------------------------------------------------------------//
double migdal_constant()
{
    return 4 * pi * r_electron * (lambdabar_electron * lambdabar_electron);
}

SBTableO make_sb_table(Json const& j)
{
    SBTableO t;
    auto vec = [&j](char const* k) { return j(k).nums(); };
    t.x = vec("x");
    t.y = vec("y");
    t.values = vec("values");
    if (t.x.size() < 2 || t.y.size() < 2 || t.values.size() != t.x.size() * t.y.size())
        throw std::runtime_error("oracle: inconsistent SB table");
    for (std::size_t i = 0; i < t.x.size(); ++i)
    {
        auto first = t.values.begin() + i * t.y.size();
        t.argmax.push_back(static_cast<std::uint32_t>(
            std::max_element(first, first + t.y.size()) - first));
    }
    return t;
}

void set_rb_element_data(double electron_mass, ElementDef& el)
{
    // Dirac-Fock form factors for Z < 5 (RelativisticBremModel.cc:190-200)
    static double const dirac_fock[4][2]
        = {{5.3104, 5.9173}, {4.7935, 5.6125}, {4.7402, 5.5377}, {4.7112, 5.4728}};
    if (el.z < 1)
        return;
    double const ff_el = el.z < 5 ? dirac_fock[el.z - 1][0]
                                  : std::log(184.15) - el.log_z / 3;
    double const ff_inel = el.z < 5 ? dirac_fock[el.z - 1][1]
                                    : std::log(1194.0) - 2 * el.log_z / 3;
    double const fc = el.coulomb_correction;
    double const invz = 1.0 / el.z;
    double const z13 = el.cbrt_z;
    double const z23 = z13 * z13;
    el.rb_fz = el.log_z / 3 + fc;
    el.rb_factor1 = (ff_el - fc) + ff_inel * invz;
    el.rb_factor2 = (1 + invz) / 12;
    el.rb_gamma_factor = 100 * electron_mass / z13;
    el.rb_epsilon_factor = 100 * electron_mass / z23;
}

double rb_dxsec(ElementDef const& el, double total_energy, double gamma_energy)
{
    double const y = gamma_energy / total_energy;
    double const onemy = 1 - y;
    double const term0 = onemy + 0.75 * (y * y);
    if (el.z < 5)
        return std::max(term0 * el.rb_factor1 + onemy * el.rb_factor2, 0.0);
    double const invz = 1 / static_cast<double>(el.z);
    double const term1 = y / (total_energy - gamma_energy);
    double const gam = term1 * el.rb_gamma_factor;
    double const eps = term1 * el.rb_epsilon_factor;
    // Tsai's screening functions (RBDiffXsCalculator::compute_screen_functions)
    double const gam2 = gam * gam;
    double const eps2 = eps * eps;
    double const phi1 = 16.863 - 2 * std::log(1 + 0.311877 * gam2)
                        + 2.4 * std::exp(-0.9 * gam) + 1.6 * std::exp(-1.5 * gam);
    double const phi2 = 2 / (3 + 19.5 * gam + 18 * gam2);
    double const psi1 = 24.34 - 2 * std::log(1 + 13.111641 * eps2)
                        + 2.8 * std::exp(-8 * eps) + 1.2 * std::exp(-29.2 * eps);
    double const psi2 = 2 / (3 + 120 * eps + 1200 * eps2);
    double const dxsec
        = term0 * ((0.25 * phi1 - el.rb_fz) + (0.25 * psi1 - 2 * el.log_z / 3) * invz)
          + 0.125 * onemy * (phi2 + psi2 * invz);
    return std::max(dxsec, 0.0);
}

//---------------------------------------------------------------------------//
// LPM SUPPRESSION (em/xs/LPMCalculator.hh)
//---------------------------------------------------------------------------//
double lpm_constant()
{
    // alpha (m_e c^2)^2 / (2 h c) (PhysicsConstants.hh:46-55), evaluated in
    // CGS from the electron mass in grams (whose rest energy differs from
    // 0.51099895 MeV by 7.5e-12 relative), then converted to MeV/cm
    constexpr double c_light = 2.99792458e10;  // [cm/s]
    constexpr double h_planck = 6.62607015e-27;  // [erg s]
    constexpr double electron_mass = 9.1093837015e-28;  // [g]
    constexpr double mev = 1.602176634e-6;  // [erg]
    double const emc2 = electron_mass * (c_light * c_light);
    return alpha_fine * (emc2 * emc2) / (2 * h_planck * c_light) / mev;
}

namespace
{
//! Horner evaluation with fma, highest coefficient last (PolyEvaluator)
template<std::size_t N>
double poly(double const (&c)[N], double x)
{
    double r = c[N - 1];
    for (std::size_t i = N - 1; i-- > 0;)
        r = std::fma(x, r, c[i]);
    return r;
}

double lpm_phi(double s)
{
    if (s < 0.01)
        return s * poly({6.0, -6 * pi}, s);
    if (s < 1.55)
    {
        double const a = poly({0.623, 0.796, 0.658}, s);
        double const b = poly({-6.0, -6 * (3 - pi), 1 / a}, s);
        return 1 - std::exp(s * b);
    }
    double const s2 = s * s;
    return 1 - 0.01190476 / (s2 * s2);
}

double lpm_g(double s, double phi)
{
    if (s < 0.01)
        return poly({-2 * phi, 12.0}, s);
    if (s < 0.415827)
    {
        double const a = poly({1.0, 3.936, 4.97, -0.05, 7.5}, s);
        double const b = poly({-4.0, -8 / a}, s);
        double const psi = 1 - std::exp(s * b);
        return 3 * psi - 2 * phi;
    }
    if (s < 1.9156)
        return std::tanh(poly({-0.160723, 3.755030, -1.798138, 0.672827, -0.120772}, s));
    double const s2 = s * s;
    return 1 - 0.0230655 / (s2 * s2);
}
}  // namespace

//! Planted mistakes for the LPM pin tests (0: none). 1: no dielectric
//! suppression; 2: xi without Stanev's cubic correction; 3: G and phi
//! exchanged in the brems DCS
int g_lpm_mutation = 0;

LpmFunctions lpm_functions(ElementDef const& el, MaterialDef const& mat,
                           bool dielectric_suppression, double gamma_energy,
                           double epsilon)
{
    if (g_lpm_mutation == 1)
        dielectric_suppression = false;
    double const lpm_energy = mat.rad_length * lpm_constant();
    // s' (Stanev Eq. 21 input)
    double const s_prime = std::sqrt(
        lpm_energy / (8 * epsilon * gamma_energy * std::fabs(epsilon - 1)));
    double const r = el.cbrt_z / 184.15;
    double const s1 = r * r;
    double const sqrt_two = 1.41421356237309504880;
    double xi = 2;
    if (s_prime > 1)
    {
        xi = 1;
    }
    else if (s_prime > sqrt_two * s1)
    {
        double const log_s1 = std::log(sqrt_two * s1);
        double const h = std::log(s_prime) / log_s1;
        xi = 1 + h - (g_lpm_mutation == 2 ? 0.0 : 0.08) * (1 - h) * h * (2 - h) / log_s1;
    }
    double s = s_prime / std::sqrt(xi);
    if (dielectric_suppression)
    {
        double const ek = epsilon * gamma_energy;
        double const k_p_sq = mat.electron_density * migdal_constant() * (ek * ek);
        s *= (1 + k_p_sq / (gamma_energy * gamma_energy));
        xi = 2;
        if (s > 1)
            xi = 1;
        else if (s > s1)
            xi = 1 + std::log(s) / std::log(s1);
    }
    double const phi = lpm_phi(s);
    if (xi * phi > 1 || s > 0.57)
        xi = 1 / phi;
    LpmFunctions out;
    out.phi = phi;
    out.xi = xi;
    out.g = lpm_g(s, phi);
    return out;
}

RbDxsec::RbDxsec(ElementDef const& el_, MaterialDef const& mat_, bool lpm, double total)
    : el(el_), mat(mat_), total_energy(total)
{
    double const density_factor = mat.electron_density * migdal_constant();
    density_corr = density_factor * (total * total);
    double const lpm_energy = mat.rad_length * lpm_constant();
    double const lpm_threshold = lpm_energy * std::sqrt(density_factor);
    enable_lpm = lpm && total > lpm_threshold;
}

double RbDxsec::operator()(double gamma_energy) const
{
    if (!enable_lpm)
        return rb_dxsec(el, total_energy, gamma_energy);
    // RBDiffXsCalculator::dxsec_per_atom_lpm; the relativistic model
    // includes the dielectric suppression
    double const epsilon = total_energy / gamma_energy;
    LpmFunctions lpm = lpm_functions(el, mat, true, gamma_energy, epsilon);
    if (g_lpm_mutation == 3)
        std::swap(lpm.g, lpm.phi);
    double const y = gamma_energy / total_energy;
    double const onemy = 1 - y;
    double const y2 = 0.25 * (y * y);
    double const term = lpm.xi * (y2 * lpm.g + (onemy + 2 * y2) * lpm.phi);
    double const dxsec = term * el.rb_factor1 + onemy * el.rb_factor2;
    return std::max(dxsec, 0.0);
}

double sb_positron_scale(int z, double mass, double kc, double inc_energy,
                         double exit_energy)
{
    double const alpha_z = 2 * pi * alpha_fine * z;
    double const delta = positron_invbeta(inc_energy, mass, kc)
                         - positron_invbeta(inc_energy, mass, exit_energy);
    return std::exp(alpha_z * std::min(delta, 0.0));
}

template<class Engine>
double sample_sb_energy(ElementDef const& el, double electron_density,
                        double inc_energy, double mass, double kc,
                        int corrector, Engine& rng, double* max_xs_out)
{
    SBTableO const& t = el.sb;
    if (!t)
        throw std::runtime_error("oracle: element has no SB table");
    std::size_t const ny = t.y.size();
    auto at = [&t, ny](std::size_t ix, std::size_t iy) { return t.values[ix * ny + iy]; };
    double const total = inc_energy + mass;
    double const dens = electron_density * migdal_constant() * (total * total);
    Interp const xl = find_interp(t.x, std::log(inc_energy));
    double const max_xs = (1 - xl.fraction) * at(xl.index, t.argmax[xl.index])
                          + xl.fraction * at(xl.index + 1, t.argmax[xl.index + 1]);
    if (max_xs_out)
        *max_xs_out = max_xs;
    double const inv_e = 1 / inc_energy;
    double const a = kc * kc + dens;
    double const logratio = std::log((1 / a) * (inc_energy * inc_energy + dens));
    double const alpha_z = 2 * pi * alpha_fine * el.z;
    double const cutoff_invbeta
        = corrector == 1 ? positron_invbeta(inc_energy, mass, kc) : 0;
    double exit, xs;
    do
    {
        exit = std::sqrt(a * std::exp(logratio * rng()) - dens);
        Interp const yl = find_interp(t.y, exit * inv_e);
        double const v
            = (1 - xl.fraction)
                  * ((1 - yl.fraction) * at(xl.index, yl.index)
                     + yl.fraction * at(xl.index, yl.index + 1))
              + xl.fraction
                    * ((1 - yl.fraction) * at(xl.index + 1, yl.index)
                       + yl.fraction * at(xl.index + 1, yl.index + 1));
        double scale = 1;
        if (corrector == 2)
        {
            scale = 0.5;  // constant scaling (SeltzerBerger.test.cc:250-262)
        }
        else if (corrector == 1)
        {
            double const delta
                = cutoff_invbeta - positron_invbeta(inc_energy, mass, exit);
            scale = std::exp(alpha_z * std::min(delta, 0.0));
        }
        xs = v * scale;
    } while (xs < max_xs * rng());
    return exit;
}

template double sample_sb_energy<MtUniform>(ElementDef const&, double, double,
                                            double, double, int, MtUniform&,
                                            double*);

//---------------------------------------------------------------------------//
void make_unit(double* v)
{
    double scale = 1 / std::sqrt(dot3(v, v));
    for (int i = 0; i < 3; ++i)
        v[i] *= scale;
}

void from_spherical(double costheta, double phi, double* out)
{
    double sintheta = std::sqrt(1 - costheta * costheta);
    out[0] = sintheta * std::cos(phi);
    out[1] = sintheta * std::sin(phi);
    out[2] = costheta;
}

//! corecel/math/ArrayUtils.hh rotate(): rotate dir from the frame of rot
void rotate(double const* dir, double const* rot, double* out)
{
    double sintheta = std::sqrt(1 - rot[2] * rot[2]);
    double cosphi, sinphi;
    if (sintheta >= 0.005)
    {
        double inv = 1 / sintheta;
        cosphi = rot[0] * inv;
        sinphi = rot[1] * inv;
    }
    else if (sintheta > 0)
    {
        cosphi = rot[0] / std::sqrt(rot[0] * rot[0] + rot[1] * rot[1]);
        sinphi = std::sqrt(1 - cosphi * cosphi);
    }
    else
    {
        cosphi = 1;
        sinphi = 0;
    }
    out[0] = (rot[2] * dir[0] + sintheta * dir[2]) * cosphi - sinphi * dir[1];
    out[1] = (rot[2] * dir[0] + sintheta * dir[2]) * sinphi + cosphi * dir[1];
    out[2] = -sintheta * dir[0] + rot[2] * dir[2];
    make_unit(out);
}

SecondaryP* allocate_secondaries(State& s, std::uint32_t n, std::uint32_t* start)
{
    if (s.stack_size + n > s.stack.size())
        return nullptr;
    *start = s.stack_size;
    s.stack_size += n;
    for (std::uint32_t i = 0; i < n; ++i)
        s.stack[*start + i] = SecondaryP{};
    return &s.stack[*start];
}

template<class Engine>
Interaction sample_model(std::string const& model, InteractorInput const& in,
                         State& s, Engine& rng)
{
    if (model == "klein_nishina")
        return klein_nishina(in, s, rng);
    if (model == "bethe_heitler")
        return bethe_heitler(in, s, rng);
    if (model == "moller_bhabha")
        return moller_bhabha(in, s, rng);
    if (model == "eplusgg")
        return eplusgg(in, s, rng);
    if (model == "seltzer_berger")
        return brems(in, s, rng, 0);
    if (model == "relativistic_brem")
        return brems(in, s, rng, 1);
    if (model == "combined_brem")
        return brems(in, s, rng, 2);
    if (model == "livermore_pe")
        return livermore_pe(in, s, rng);
    throw std::runtime_error("oracle: unsupported model " + model);
}

template Interaction sample_model<XorwowUniform>(std::string const&,
                                                 InteractorInput const&, State&,
                                                 XorwowUniform&);
template Interaction sample_model<MtUniform>(std::string const&,
                                             InteractorInput const&, State&,
                                             MtUniform&);
template Interaction sample_model<ReplayUniform>(std::string const&,
                                                 InteractorInput const&, State&,
                                                 ReplayUniform&);

Interaction sample_interaction(State& s, Track& t, ModelDef const& model)
{
    InteractorInput in;
    in.pd = s.problem->physics.get();
    in.particle = t.particle;
    in.energy = t.energy;
    for (int i = 0; i < 3; ++i)
        in.dir[i] = t.geo.dir[i];
    in.material = s.problem->volume_to_material[t.geo.vol];
    in.element = t.element == none ? 0 : t.element;
    XorwowUniform rng{t.rng};
    return sample_model(model.model, in, s, rng);
}

double hardwired_xs(State const& s, ProcessDef const& p, std::uint32_t mat, double e)
{
    PhysicsData const& pd = *s.problem->physics;
    if (p.cls == "annihilation")
    {
        // EPlusGGMacroXsCalculator (em/xs/EPlusGGMacroXsCalculator.hh): Heitler
        double const me = pd.particles[pd.electron].mass;
        double const gamma = std::max(e, 1e-6) / me;
        double const sqrt_gg2 = std::sqrt(gamma * (gamma + 2));
        double const gp1 = gamma + 1;
        double const poly = 1 + gp1 * (4 + gp1 * 1);
        return pi * r_electron * r_electron * pd.materials[mat].electron_density
               * (poly * std::log(gamma + 1 + sqrt_gg2) - (gamma + 4) * sqrt_gg2)
               / (gamma * (gamma + 2) * (gamma + 2));
    }
    if (p.cls == "photoelectric")
        return livermore_pe_macro_xs(pd, mat, e);
    throw std::runtime_error("oracle: no cross section for process " + p.label);
}

std::uint32_t sample_element(State& s, Track& t, ProcessDef const& p,
                             std::size_t model_index, std::uint32_t mat)
{
    PhysicsData const& pd = *s.problem->physics;
    if (mat >= pd.materials.size() || pd.materials[mat].elements.size() <= 1)
        return 0;
    // Element selection by the per-element normalised cumulative XS
    // (TabulatedElementSelector.hh:75-95); models without tables select none
    if (model_index >= p.elem_cdf.size() || p.elem_cdf[model_index].size() <= mat
        || p.elem_cdf[model_index][mat].empty())
        return 0;
    auto const& cdf = p.elem_cdf[model_index][mat];
    XorwowUniform rng{t.rng};
    double u = rng();
    std::size_t i = 0;
    for (; i + 1 < cdf.size(); ++i)
    {
        if (cdf[i].calc(t.energy) > u)
            break;
    }
    return static_cast<std::uint32_t>(i);
}

//---------------------------------------------------------------------------//
double coulomb_correction(int z)
{
    // Davies-Bethe-Maximon: f(Z) = a^2 sum_n 1 / (n (n^2 + a^2)), written as
    // a^2 [1/(1+a^2) + sum_k (-a^2)^k (zeta(2k+3) - 1)]
    double const a2 = (alpha_fine * z) * (alpha_fine * z);
    // zeta(n) - 1: partial sum below m = 40 plus the Euler-Maclaurin tail
    auto zeta_minus_one = [](int n) {
        double s = 0;
        for (int m = 39; m >= 2; --m)
            s += std::pow(m, -n);
        double const x = 40;
        double const f = std::pow(x, -n);
        double const d = n;
        double const tail = x * f / (d - 1) + f / 2 + d * f / (12 * x)
                            - d * (d + 1) * (d + 2) * f / (720 * std::pow(x, 3))
                            + d * (d + 1) * (d + 2) * (d + 3) * (d + 4) * f
                                  / (30240 * std::pow(x, 5))
                            - d * (d + 1) * (d + 2) * (d + 3) * (d + 4) * (d + 5)
                                  * (d + 6) * f / (1209600 * std::pow(x, 7));
        return tail + s;
    };
    double fz = 1 / (1 + a2);
    double apow = 1;
    for (int k = 0; k < 16; ++k)
    {
        fz += apow * zeta_minus_one(2 * k + 3);
        apow *= -a2;
    }
    return fz * a2;
}

double mass_radiation_coeff(int z, double atomic_mass)
{
    // Tsai radiation length (PDG review "Passage of particles through matter")
    double lrad, lrad_prime;
    switch (z)
    {
        case 1: lrad = 5.31; lrad_prime = 6.144; break;
        case 2: lrad = 4.79; lrad_prime = 5.621; break;
        case 3: lrad = 4.74; lrad_prime = 5.805; break;
        case 4: lrad = 4.71; lrad_prime = 5.924; break;
        default:
            lrad = std::log(184.15 * std::pow(z, -1.0 / 3));
            lrad_prime = std::log(1194.0 * std::pow(z, -2.0 / 3));
    }
    // Per gram: the atom's mass is A amu (mat/detail/Utils.cc:102-107)
    constexpr double amu_gram = 1.66053906660e-24;
    double const factor = 4 * alpha_fine * r_electron * r_electron;
    return factor / (atomic_mass * amu_gram)
           * (z * z * (lrad - coulomb_correction(z)) + z * lrad_prime);
}

//---------------------------------------------------------------------------//
}  // namespace orc
