//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/TableGen.cc
//---------------------------------------------------------------------------//
#include "TableGen.hh"

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <vector>

#include "corecel/DeviceRuntime.hh"
#include "em/LPM.hh"
#include "io/G4EmlowReaders.hh"

namespace celeritas_amd
{
namespace
{
//---------------------------------------------------------------------------//
// CONSTANTS (CODATA 2018; MeV, cm, g)
constexpr double pi = 3.14159265358979323846;
constexpr double electron_mass = 0.51099895000;  // [MeV]
constexpr double r_e = 2.8179403262e-13;  // [cm]
constexpr double alpha = 7.2973525693e-3;
constexpr double avogadro = 6.02214076e23;
constexpr double amu_gram = 1.66053906660e-24;
constexpr double hbarc = 197.3269804e-13;  // [MeV cm]
constexpr double lambdabar_e = 3.8615926796e-11;  // [cm]
constexpr double barn = 1e-24;  // [cm^2]
constexpr double millibarn = 1e-27;  // [cm^2]
constexpr double sb_upper_limit = 1e3;  // [MeV]
constexpr double ln10 = 2.302585092994045684;

//---------------------------------------------------------------------------//
// QUADRATURE
//! Gauss-Legendre nodes and weights on [-1, 1] by Newton iteration
struct GaussLegendre
{
    std::vector<double> x, w;
    explicit GaussLegendre(int n)
    {
        x.resize(n);
        w.resize(n);
        for (int i = 0; i < n; ++i)
        {
            double z = std::cos(pi * (i + 0.75) / (n + 0.5));
            double dp = 0;
            for (int it = 0; it < 100; ++it)
            {
                double p0 = 1, p1 = z;
                for (int k = 2; k <= n; ++k)
                {
                    double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                    p0 = p1;
                    p1 = p2;
                }
                dp = n * (z * p1 - p0) / (z * z - 1);
                double dz = p1 / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-16)
                    break;
            }
            x[i] = z;
            w[i] = 2 / ((1 - z * z) * dp * dp);
        }
    }
    //! Integrate f over [a, b]
    template<class F>
    double operator()(F&& f, double a, double b) const
    {
        double const h = (b - a) / 2;
        double const c = (b + a) / 2;
        double s = 0;
        for (std::size_t i = 0; i < x.size(); ++i)
            s += w[i] * f(c + h * x[i]);
        return s * h;
    }
};

GaussLegendre const& gl()
{
    static GaussLegendre const g(24);
    return g;
}

//! Integrate f(t) over [a, b] on nsub log-spaced panels (a > 0)
template<class F>
double integrate_log(F&& f, double a, double b, int nsub)
{
    if (!(b > a) || a <= 0)
        return 0;
    double const la = std::log(a);
    double const h = (std::log(b) - la) / nsub;
    double s = 0;
    for (int i = 0; i < nsub; ++i)
    {
        s += gl()([&f](double lt) {
            double const t = std::exp(lt);
            return f(t) * t;
        }, la + i * h, la + (i + 1) * h);
    }
    return s;
}

//---------------------------------------------------------------------------//
// ELEMENT AND MATERIAL DATA
struct Shell
{
    double binding;  // [MeV]
    double weight;
};

struct Element
{
    int z = 0;
    double atomic_mass = 0;
    double mean_exc = 0;  // [MeV]
    double cbrt_z = 0, log_z = 0, fc = 0;
    // Relativistic brems (RelativisticBremModel.cc:150-185)
    double rb_fz = 0, rb_factor1 = 0, rb_factor2 = 0, rb_gamma = 0, rb_eps = 0;
    // Photoelectric shell model, or the Livermore data read from G4EMLOW
    std::vector<Shell> shells;
    double pe_scale = 0;  // C [barn MeV^3]
    JsonValue pe_data;
    // SB table
    std::vector<double> sb_x, sb_y, sb_values;
};

struct Material
{
    std::string label;
    double number_density = 0;  // atoms / cm^3
    double density = 0;
    double electron_density = 0;
    double zeff = 0;
    double mean_exc = 0;
    double rad_length = 0;
    std::vector<std::pair<std::size_t, double>> comps;  // element, fraction
    bool gas = false;
    // Sternheimer density-effect parameters
    double cbar = 0, x0 = 0, x1 = 0, a = 0, m = 3;
};

double zeta_minus_one(int n)
{
    double sum = 0;
    for (int k = 63; k >= 2; --k)
        sum += std::pow(static_cast<double>(k), -n);
    double const x = 64;
    double const xn = std::pow(x, -n);
    double const dn = n;
    return x * xn / (dn - 1) + xn / 2 + dn * xn / (12 * x)
           - dn * (dn + 1) * (dn + 2) * xn / (720 * x * x * x) + sum;
}

double coulomb_correction(int z)
{
    double const a2 = (alpha * z) * (alpha * z);
    double r = 1 / (1 + a2);
    double ap = 1;
    for (int k = 0; k < 16; ++k)
    {
        r += ap * zeta_minus_one(2 * k + 3);
        ap *= -a2;
    }
    return r * a2;
}

double mass_radiation_coeff(Element const& el)
{
    int const z = el.z;
    double lrad, lrad_p;
    switch (z)
    {
        case 1: lrad = 5.31; lrad_p = 6.144; break;
        case 2: lrad = 4.79; lrad_p = 5.621; break;
        case 3: lrad = 4.74; lrad_p = 5.805; break;
        case 4: lrad = 4.71; lrad_p = 5.924; break;
        default:
            lrad = std::log(184.15 * std::pow(z, -1.0 / 3));
            lrad_p = std::log(1194.0 * std::pow(z, -2.0 / 3));
    }
    // Per gram: the atom's mass is A amu (mat/detail/Utils.cc:102-107)
    return 4 * alpha * r_e * r_e / (el.atomic_mass * amu_gram)
           * (z * z * (lrad - el.fc) + z * lrad_p);
}

//! Elastic and inelastic brems form factors for Z = 1-4 (Dirac-Fock atom;
//! the values the relativistic model takes where Thomas-Fermi fails,
//! RelativisticBremModel.cc:190-200)
constexpr double dirac_fock_form_factors[4][2]
    = {{5.3104, 5.9173}, {4.7935, 5.6125}, {4.7402, 5.5377}, {4.7112, 5.4728}};

//---------------------------------------------------------------------------//
// DIFFERENTIAL CROSS SECTIONS
//! Tsai DCS factor (RBDiffXsCalculator::dxsec_per_atom); dsigma/dk =
//! 16 alpha r_e^2 Z^2 / (3 k) * dxsec
double rb_dxsec(Element const& el, double total, double k)
{
    double const y = k / total;
    double const onemy = 1 - y;
    double const term0 = onemy + 0.75 * y * y;
    if (el.z < 5)
        return std::max(term0 * el.rb_factor1 + onemy * el.rb_factor2, 0.0);
    double const invz = 1.0 / el.z;
    double const term1 = y / (total - k);
    double const gam = term1 * el.rb_gamma;
    double const eps = term1 * el.rb_eps;
    double const phi1 = 16.863 - 2 * std::log(1 + 0.311877 * gam * gam)
                        + 2.4 * std::exp(-0.9 * gam) + 1.6 * std::exp(-1.5 * gam);
    double const phi2 = 2 / (3 + 19.5 * gam + 18 * gam * gam);
    double const psi1 = 24.34 - 2 * std::log(1 + 13.111641 * eps * eps)
                        + 2.8 * std::exp(-8 * eps) + 1.2 * std::exp(-29.2 * eps);
    double const psi2 = 2 / (3 + 120 * eps + 1200 * eps * eps);
    double const d = term0
                         * ((0.25 * phi1 - el.rb_fz)
                            + (0.25 * psi1 - 2 * el.log_z / 3) * invz)
                     + 0.125 * onemy * (phi2 + psi2 * invz);
    return std::max(d, 0.0);
}

//! Bilinear SB table lookup (TwodGridCalculator semantics) [mb]
double sb_chi(Element const& el, double energy, double kappa)
{
    auto locate = [](std::vector<double> const& g, double v, double* frac) {
        std::size_t i = static_cast<std::size_t>(
            std::upper_bound(g.begin(), g.end(), v) - g.begin());
        i = std::min(std::max<std::size_t>(i, 1), g.size() - 1) - 1;
        *frac = std::min(std::max((v - g[i]) / (g[i + 1] - g[i]), 0.0), 1.0);
        return i;
    };
    double fx, fy;
    std::size_t const ix = locate(el.sb_x, std::log(energy), &fx);
    std::size_t const iy = locate(el.sb_y, kappa, &fy);
    std::size_t const ny = el.sb_y.size();
    auto v = [&](std::size_t i, std::size_t j) { return el.sb_values[i * ny + j]; };
    return (1 - fx) * ((1 - fy) * v(ix, iy) + fy * v(ix, iy + 1))
           + fx * ((1 - fy) * v(ix + 1, iy) + fy * v(ix + 1, iy + 1));
}

//! Relativistic DCS with LPM suppression (RBDiffXsCalculator.hh:186-205)
//! when enabled and the total energy exceeds the material's threshold
double rb_dxsec_lpm(Element const& el, Material const& mat, double total, double k,
                    bool lpm)
{
    double const density_factor = mat.electron_density * lpm::migdal_constant();
    double const lpm_energy = mat.rad_length * lpm::lpm_constant();
    if (!lpm || !(total > lpm_energy * std::sqrt(density_factor)))
        return rb_dxsec(el, total, k);
    double const epsilon = total / k;
    lpm::Functions const f = lpm::calc_functions(el.cbrt_z, mat.electron_density,
                                                 lpm_energy, true, k, epsilon);
    double const y = k / total;
    double const onemy = 1 - y;
    double const y2 = 0.25 * (y * y);
    double const term = f.xi * (y2 * f.g + (onemy + 2 * y2) * f.phi);
    return std::max(term * el.rb_factor1 + onemy * el.rb_factor2, 0.0);
}

//! k dsigma/dk per atom [cm^2] with dielectric suppression, as sampled by
//! the SB (E < 1 GeV) and relativistic samplers (the latter with LPM
//! suppression when enabled, as Geant4's relativistic model tabulates it)
double brems_k_dsigma(Element const& el, Material const& mat, double energy, double k,
                      bool lpm)
{
    double const total = energy + electron_mass;
    double const dens = mat.electron_density * 4 * pi * r_e * lambdabar_e * lambdabar_e
                        * total * total;
    double const supp = k * k / (k * k + dens);
    if (energy < sb_upper_limit)
    {
        double const beta2 = energy * (energy + 2 * electron_mass) / (total * total);
        return el.z * el.z / beta2 * sb_chi(el, energy, k / energy) * millibarn * supp;
    }
    return 16.0 / 3 * alpha * r_e * r_e * el.z * el.z * rb_dxsec_lpm(el, mat, total, k, lpm)
           * supp;
}

//! Bethe-Heitler screening functions (as sampled by the interactor)
double bh_f1(double d)
{
    return d > 1.4 ? 42.038 - 8.29 * std::log(d + 0.958) : 42.184 - d * (7.444 - 1.623 * d);
}
double bh_f2(double d)
{
    return d > 1.4 ? 42.038 - 8.29 * std::log(d + 0.958) : 41.326 - d * (5.848 - 0.902 * d);
}

//! Pair-production cross section per atom [cm^2]:
//! alpha r_e^2 Z (Z + xi) 2 int [(2/3)(1/2-e)^2 F1' + (1/3) F2'] de
double pair_xs(Element const& el, double energy)
{
    double const eps0 = electron_mass / energy;
    if (eps0 >= 0.5)
        return 0;
    double fz = 8.0 / 3 * el.log_z;
    if (energy > 50)
        fz += 8 * el.fc;
    double const lrad = std::log(184.15 / el.cbrt_z);
    double const xi = std::log(1194.0 / (el.cbrt_z * el.cbrt_z)) / (lrad - el.fc);
    auto f = [&](double e) {
        double const delta = 136 / el.cbrt_z * eps0 / (e * (1 - e));
        double const s = 0.5 - e;
        return 2.0 / 3 * s * s * std::max(bh_f1(delta) - fz, 0.0)
               + 1.0 / 3 * std::max(bh_f2(delta) - fz, 0.0);
    };
    double integral = 0;
    int const npanel = 8;
    double const h = (0.5 - eps0) / npanel;
    for (int i = 0; i < npanel; ++i)
        integral += gl()(f, eps0 + i * h, eps0 + (i + 1) * h);
    return alpha * r_e * r_e * el.z * (el.z + xi) * 2 * integral;
}

//! Klein-Nishina total cross section per electron [cm^2]
double kn_xs(double energy)
{
    double const k = energy / electron_mass;
    double const sigma_t = 8 * pi / 3 * r_e * r_e;
    if (k < 1e-3)
        return sigma_t
               * (1 + k * (-2 + k * (26.0 / 5 + k * (-133.0 / 10 + k * 1144.0 / 35))));
    double const l = std::log1p(2 * k);
    double const a = 1 + 2 * k;
    return 2 * pi * r_e * r_e
           * ((1 + k) / (k * k) * (2 * (1 + k) / a - l / k) + l / (2 * k)
              - (1 + 3 * k) / (a * a));
}

//! Moller (e-) / Bhabha (e+) cross section per electron above cut [cm^2]
double ioni_xs(double energy, double cut, bool electron)
{
    double const xmin = cut / energy;
    double const xmax = electron ? 0.5 : 1.0;
    if (xmin >= xmax)
        return 0;
    double const tau = energy / electron_mass;
    double const gam = tau + 1;
    double const g2 = gam * gam;
    double const beta2 = tau * (tau + 2) / g2;
    double cross;
    if (electron)
    {
        double const gg = (2 * gam - 1) / g2;
        cross = ((xmax - xmin)
                     * (1 - gg + 1 / (xmin * xmax) + 1 / ((1 - xmin) * (1 - xmax)))
                 - gg * std::log(xmax * (1 - xmin) / (xmin * (1 - xmax))))
                / beta2;
    }
    else
    {
        double const y = 1 / (1 + gam);
        double const y2 = y * y;
        double const c = 1 - 2 * y;
        double const b1 = 2 - y2;
        double const b2 = c * (3 + y2);
        double const b4 = c * c * c;
        double const b3 = c * c + b4;
        cross = (xmax - xmin)
                    * (1 / (beta2 * xmin * xmax) + b2 - 0.5 * b3 * (xmin + xmax)
                       + b4 * (xmin * xmin + xmin * xmax + xmax * xmax) / 3)
                - b1 * std::log(xmax / xmin);
    }
    return std::max(cross, 0.0) * 2 * pi * r_e * r_e * electron_mass / energy;
}

//! Sternheimer density correction delta(x), x = log10(beta gamma)
double density_effect(Material const& m, double x)
{
    if (x < m.x0)
        return 0;
    double const d = 2 * ln10 * x - m.cbar;
    if (x < m.x1)
        return d + m.a * std::pow(m.x1 - x, m.m);
    return d;
}

//! Berger-Seltzer restricted collision stopping power [MeV/cm]
double ioni_dedx(Material const& m, double energy, double cut, bool electron)
{
    double const tau = energy / electron_mass;
    double const gam = tau + 1;
    double const g2 = gam * gam;
    double const bg2 = tau * (tau + 2);
    double const beta2 = bg2 / g2;
    double const eexc = m.mean_exc / electron_mass;
    double const eexc2 = eexc * eexc;
    double const max_t = electron ? 0.5 * energy : energy;
    double const d = std::min(cut, max_t) / electron_mass;
    double dedx;
    if (electron)
    {
        dedx = std::log(2 * (tau + 2) / eexc2) - 1 - beta2 + std::log((tau - d) * d)
               + tau / (tau - d)
               + (0.5 * d * d + (2 * tau + 1) * std::log(1 - d / tau)) / g2;
    }
    else
    {
        double const y = 1 / (1 + gam);
        double const d2 = d * d, d3 = d2 * d, d4 = d3 * d;
        dedx = std::log(2 * (tau + 2) / eexc2) + std::log(tau * d)
               - beta2
                     * (tau + 2 * d - y * (3 * d2 + y * (d - d3 + y * (d2 - tau * d3 + d4))))
                     / tau;
    }
    dedx -= density_effect(m, 0.5 * std::log10(bg2));
    return std::max(dedx, 0.0) * 2 * pi * electron_mass * r_e * r_e * m.electron_density
           / beta2;
}

/*!
 * Transport cross section per atom [cm^2] of the screened Rutherford DCS
 * with Moliere's screening angle:
 *   sigma_tr = 2 pi Z(Z+1) r_e^2 (mc^2)^2 / (beta^2 (pc)^2)
 *              [ln(1 + 1/A) - 1/(1 + A)],
 *   A = (hbar c / (2 pc a_TF))^2 (1.13 + 3.76 (alpha Z / beta)^2),
 *   a_TF = 0.8853 a_0 Z^(-1/3).
 */
double msc_transport_xs(Element const& el, double energy)
{
    constexpr double bohr_radius = 0.529177210903e-8;  // [cm]
    double const pc2 = energy * (energy + 2 * electron_mass);
    double const total = energy + electron_mass;
    double const beta2 = pc2 / (total * total);
    double const z = el.z;
    double const a_tf = 0.8853 * bohr_radius / el.cbrt_z;
    double const az = alpha * z;
    double const screen = hbarc * hbarc / (4 * pc2 * a_tf * a_tf)
                          * (1.13 + 3.76 * az * az / beta2);
    double const c = z * (z + 1) * r_e * r_e * electron_mass * electron_mass / (beta2 * pc2);
    return 2 * pi * c * (std::log1p(1 / screen) - 1 / (1 + screen));
}

//---------------------------------------------------------------------------//
// PHOTOELECTRIC SHELL MODEL
double pe_shape(double e)
{
    double const f = 1 + e / electron_mass;
    return f * f / (e * e * e);
}

//! Livermore-layout element data from the shell model
JsonValue pe_element(Element const& el)
{
    auto const& sh = el.shells;
    double const c = el.pe_scale;
    double const b_k = sh.front().binding;
    double const thresh = 2 * b_k;
    JsonValue r = JsonValue::object();
    r.obj["thresh_lo"] = JsonValue::number(thresh);
    r.obj["thresh_hi"] = JsonValue::number(thresh);
    auto table = [](std::vector<double> const& x, std::vector<double> const& y) {
        JsonValue t = JsonValue::object();
        t.obj["x"] = JsonValue::array(x);
        t.obj["y"] = JsonValue::array(y);
        return t;
    };
    auto logspace = [](double a, double b, int n) {
        std::vector<double> v(n);
        for (int i = 0; i < n; ++i)
            v[i] = a * std::pow(b / a, double(i) / (n - 1));
        v.back() = b;
        return v;
    };
    double total_w = 0;
    for (auto const& s : sh)
        total_w += s.weight;
    // Total above the K edge (sigma E^3 tabulated)
    {
        auto x = logspace(b_k, thresh, 24);
        std::vector<double> y;
        for (double e : x)
            y.push_back(total_w * c * pe_shape(e) * e * e * e);
        r.obj["xs_hi"] = table(x, y);
    }
    // Below the K edge: piecewise by open shells
    {
        std::vector<double> edges;
        for (auto const& s : sh)
            edges.push_back(s.binding);
        std::sort(edges.begin(), edges.end());
        std::vector<double> x, y;
        for (std::size_t j = 0; j + 1 < edges.size(); ++j)
        {
            double w = 0;
            for (auto const& s : sh)
                if (s.binding <= edges[j])
                    w += s.weight;
            for (double e : {edges[j], edges[j + 1] * (1 - 1e-9)})
            {
                x.push_back(e);
                y.push_back(w * c * pe_shape(e) * e * e * e);
            }
        }
        if (x.size() >= 2)
            r.obj["xs_lo"] = table(x, y);
    }
    JsonValue shells = JsonValue::array();
    double cum = 0;
    for (auto const& s : sh)
    {
        cum += s.weight;
        JsonValue js = JsonValue::object();
        js.obj["binding_energy"] = JsonValue::number(s.binding);
        double const wc = cum * c;
        std::vector<double> param = {wc / (electron_mass * electron_mass),
                                     2 * wc / electron_mass, wc, 0, 0, 0};
        js.obj["param_lo"] = JsonValue::array(param);
        js.obj["param_hi"] = JsonValue::array(param);
        auto x = logspace(s.binding, thresh, 24);
        std::vector<double> y;
        for (double e : x)
            y.push_back(s.weight * c * pe_shape(e) * e * e * e);
        js.obj["x"] = JsonValue::array(x);
        js.obj["y"] = JsonValue::array(y);
        shells.arr.push_back(js);
    }
    r.obj["shells"] = shells;
    return r;
}

//! Photoelectric cross section per atom [cm^2] of the shell model
double pe_xs(Element const& el, double energy)
{
    double w = 0;
    for (auto const& s : el.shells)
        if (energy >= s.binding)
            w += s.weight;
    return w * el.pe_scale * pe_shape(std::max(energy, el.shells.back().binding)) * barn;
}

//---------------------------------------------------------------------------//
// GRIDS
struct LogGrid
{
    double emin, emax;
    int size;
    double energy(int i) const
    {
        return std::exp(std::log(emin) + (std::log(emax) - std::log(emin)) * i / (size - 1));
    }
};

JsonValue grid_json(LogGrid const& g, std::vector<double> const& values)
{
    JsonValue j = JsonValue::object();
    j.obj["emin"] = JsonValue::number(g.emin);
    j.obj["emax"] = JsonValue::number(g.emax);
    j.obj["values"] = JsonValue::array(values);
    return j;
}

//! Value of a log grid at energy e (linear interpolation in E)
double grid_value(LogGrid const& g, std::vector<double> const& v, double e)
{
    double const t = (std::log(e) - std::log(g.emin))
                     / ((std::log(g.emax) - std::log(g.emin)) / (g.size - 1));
    int i = std::min(std::max(static_cast<int>(t), 0), g.size - 2);
    double const el = g.energy(i), er = g.energy(i + 1);
    return v[i] + (v[i + 1] - v[i]) * (e - el) / (er - el);
}

//! Range table by integrating 1/dedx (linear in E between grid points)
std::vector<double> range_table(LogGrid const& g, std::vector<double> const& dedx)
{
    std::vector<double> r(g.size);
    r[0] = 2 * g.emin / dedx[0];
    for (int i = 1; i < g.size; ++i)
    {
        double const a = g.energy(i - 1), b = g.energy(i);
        r[i] = r[i - 1]
               + gl()([&](double e) { return 1 / grid_value(g, dedx, e); }, a, b);
    }
    return r;
}

JsonValue model_json(char const* label, char const* model)
{
    JsonValue m = JsonValue::object();
    m.obj["label"] = JsonValue::string(label);
    m.obj["model"] = JsonValue::string(model);
    m.obj["low"] = JsonValue::number(0);
    JsonValue arr = JsonValue::array();
    arr.arr.push_back(m);
    return arr;
}

//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
JsonValue generate_physics(JsonValue const& spec)
{
    // Particles
    JsonValue particles = JsonValue::array();
    {
        auto add = [&](char const* label, int pdg, double mass, double charge) {
            JsonValue p = JsonValue::object();
            p.obj["label"] = JsonValue::string(label);
            p.obj["pdg"] = JsonValue::number(pdg);
            p.obj["mass"] = JsonValue::number(mass);
            p.obj["charge"] = JsonValue::number(charge);
            particles.arr.push_back(p);
        };
        add("electron", 11, electron_mass, -1);
        add("positron", -11, electron_mass, 1);
        add("gamma", 22, 0, 0);
    }

    // G4EMLOW data (optional): "g4emlow": {"seltzer_berger": dir,
    // "livermore_pe": dir}, where true takes the directory from $G4LEDATA
    // as the reference's readers do; elements then use the tabulated SB
    // brems DCS and the Livermore photoelectric data instead of the
    // built-in models
    std::unique_ptr<SeltzerBergerReader> sb_reader;
    std::unique_ptr<LivermorePEReader> pe_reader;
    if (spec.has("g4emlow"))
    {
        JsonValue const& g = spec.at("g4emlow");
        if (g.has("seltzer_berger"))
        {
            JsonValue const& v = g.at("seltzer_berger");
            if (v.type == JsonValue::Type::string)
                sb_reader = std::make_unique<SeltzerBergerReader>(v.as_string());
            else if (v.as_bool())
                sb_reader = std::make_unique<SeltzerBergerReader>();
        }
        if (g.has("livermore_pe"))
        {
            JsonValue const& v = g.at("livermore_pe");
            if (v.type == JsonValue::Type::string)
                pe_reader = std::make_unique<LivermorePEReader>(v.as_string());
            else if (v.as_bool())
                pe_reader = std::make_unique<LivermorePEReader>();
        }
    }

    // Elements
    std::vector<Element> elements;
    for (auto const& ej : spec.at("elements").arr)
    {
        Element el;
        el.z = static_cast<int>(ej.get_int("z", 0));
        CELER_VALIDATE(el.z >= 1 && el.z <= 100,
                       << "table generator supports 1 <= Z <= 100 (got " << el.z << ")");
        el.atomic_mass = ej.at("atomic_mass").as_double();
        el.mean_exc = ej.at("mean_exc_energy").as_double();
        el.cbrt_z = std::cbrt(double(el.z));
        el.log_z = std::log(double(el.z));
        el.fc = coulomb_correction(el.z);
        double ff_el = std::log(184.15) - el.log_z / 3;
        double ff_inel = std::log(1194.0) - 2 * el.log_z / 3;
        if (el.z < 5)
        {
            ff_el = dirac_fock_form_factors[el.z - 1][0];
            ff_inel = dirac_fock_form_factors[el.z - 1][1];
        }
        double const invz = 1.0 / el.z;
        el.rb_fz = el.log_z / 3 + el.fc;
        el.rb_factor1 = (ff_el - el.fc) + ff_inel * invz;
        el.rb_factor2 = (1 + invz) / 12;
        el.rb_gamma = 100 * electron_mass / el.cbrt_z;
        el.rb_eps = 100 * electron_mass / (el.cbrt_z * el.cbrt_z);
        if (pe_reader)
        {
            el.pe_data = (*pe_reader)(el.z);
        }
        else
        {
            for (auto const& sj : ej.at("shells").arr)
                el.shells.push_back({sj[0].as_double(), sj[1].as_double()});
            CELER_VALIDATE(!el.shells.empty(), << "element needs photoelectric shells");
            std::sort(el.shells.begin(), el.shells.end(),
                      [](Shell const& a, Shell const& b) { return a.binding > b.binding; });
            auto ref = ej.at("pe_reference").as_doubles();
            CELER_VALIDATE(ref.size() == 2, << "pe_reference is [energy, barn]");
            double w = 0;
            for (auto const& s : el.shells)
                if (ref[0] >= s.binding)
                    w += s.weight;
            el.pe_scale = ref[1] / (w * pe_shape(ref[0]));
        }
        if (sb_reader)
        {
            SBTable t = (*sb_reader)(el.z);
            el.sb_x = std::move(t.x);
            el.sb_y = std::move(t.y);
            el.sb_values = std::move(t.values);
        }
        else
        {
            // Synthetic SB table: scaled DCS (beta^2/Z^2) k dsigma/dk [mb] on
            // 8 log-E points per decade from 1 keV to 10 GeV
            static double const kappa[]
                = {1e-12, 0.025, 0.05, 0.075, 0.1,    0.15,  0.2,     0.25,
                   0.3,   0.35,  0.4,  0.45,  0.5,    0.55,  0.6,     0.65,
                   0.7,   0.75,  0.8,  0.85,  0.9,    0.925, 0.95,    0.97,
                   0.99,  0.995, 0.999, 0.9995, 0.9999, 0.99995, 0.99999, 1};
            el.sb_y.assign(std::begin(kappa), std::end(kappa));
            for (int i = 0; i <= 56; ++i)
            {
                double const e = 1e-3 * std::pow(10.0, i / 8.0);
                el.sb_x.push_back(std::log(e));
                double const total = e + electron_mass;
                double const beta2 = e * (e + 2 * electron_mass) / (total * total);
                for (double kap : el.sb_y)
                {
                    double const k = std::min(kap * e, e * (1 - 1e-12));
                    el.sb_values.push_back(beta2 * 16.0 / 3 * alpha * r_e * r_e
                                           * rb_dxsec(el, total, k) / millibarn);
                }
            }
        }
        elements.push_back(std::move(el));
    }

    // Materials
    std::vector<Material> materials;
    for (auto const& mj : spec.at("materials").arr)
    {
        Material m;
        m.label = mj.get_str("label", "");
        double frac_sum = 0;
        for (auto const& cj : mj.at("elements").arr)
        {
            auto const idx = static_cast<std::size_t>(cj.get_int("element", 0));
            CELER_VALIDATE(idx < elements.size(), << "bad element index in " << m.label);
            m.comps.push_back({idx, cj.get("fraction", 1.0)});
            frac_sum += m.comps.back().second;
        }
        for (auto& c : m.comps)
            c.second /= frac_sum;
        double avg_amu = 0, avg_z = 0, rad_coeff = 0, log_i = 0;
        for (auto const& c : m.comps)
        {
            Element const& el = elements[c.first];
            avg_amu += c.second * el.atomic_mass;
            avg_z += c.second * el.z;
            rad_coeff += c.second * mass_radiation_coeff(el);
            log_i += c.second * el.z * std::log(el.mean_exc);
        }
        m.density = mj.at("density").as_double();
        m.number_density = m.density / (avg_amu * amu_gram);
        m.electron_density = m.number_density * avg_z;
        m.zeff = avg_z;
        m.rad_length = 1 / (rad_coeff * m.density);
        m.mean_exc = mj.has("mean_exc_energy") ? mj.at("mean_exc_energy").as_double()
                                               : std::exp(log_i / avg_z);
        m.gas = mj.get_bool("gas", false);
        // Sternheimer-Peierls (1971) general density-effect parameters
        double const plasma = std::sqrt(4 * pi * m.electron_density * r_e) * hbarc;
        m.cbar = 1 + 2 * std::log(m.mean_exc / plasma);
        double const i_ev = m.mean_exc * 1e6;
        if (m.gas)
        {
            m.x1 = 4;
            m.x0 = m.cbar < 10 ? 1.6 : m.cbar < 12.25 ? 0.326 * m.cbar - 1.5 : 2.5;
        }
        else if (i_ev < 100)
        {
            m.x1 = 2;
            m.x0 = m.cbar < 3.681 ? 0.2 : 0.326 * m.cbar - 1.0;
        }
        else
        {
            m.x1 = 3;
            m.x0 = m.cbar < 5.215 ? 0.2 : 0.326 * m.cbar - 1.5;
        }
        m.m = 3;
        m.a = (m.cbar - 2 * ln10 * m.x0) / std::pow(m.x1 - m.x0, m.m);
        materials.push_back(std::move(m));
    }
    std::size_t const nmat = materials.size();

    // Cutoffs [particle][material]
    auto cut_of = [&](char const* p) {
        auto v = spec.at("cutoffs").at(p).as_doubles();
        CELER_VALIDATE(v.size() == nmat, << "cutoffs for " << p << " need one per material");
        return v;
    };
    auto const cut_gamma = cut_of("gamma");
    auto const cut_electron = cut_of("electron");
    auto const cut_positron = cut_of("positron");

    JsonValue const opts = spec.has("options") ? spec.at("options") : JsonValue::object();
    int const bpd = static_cast<int>(opts.get_int("bins_per_decade", 7));
    // LPM suppression of the relativistic brems and Bethe-Heitler models:
    // on unless the spec turns it off (GeantPhysicsOptions::lpm)
    bool const enable_lpm = opts.get_bool("lpm", true);
    LogGrid const grid{opts.get("emin", 1e-4), opts.get("emax", 1e8), 0};
    int const npts = static_cast<int>(std::lround(bpd * std::log10(grid.emax / grid.emin))) + 1;
    LogGrid const g{grid.emin, grid.emax, npts};

    auto per_material = [&](auto&& fn) {
        JsonValue arr = JsonValue::array();
        for (std::size_t mi = 0; mi < nmat; ++mi)
        {
            std::vector<double> v(npts);
            for (int i = 0; i < npts; ++i)
                v[i] = fn(materials[mi], mi, g.energy(i));
            arr.arr.push_back(grid_json(g, v));
        }
        return arr;
    };
    // Tables of processes with a production threshold start at it (as the
    // Geant4 lambda tables do), so that interpolation gives exactly zero
    // below threshold
    auto per_material_from = [&](auto&& threshold, auto&& fn) {
        JsonValue arr = JsonValue::array();
        for (std::size_t mi = 0; mi < nmat; ++mi)
        {
            double const e0 = std::max(threshold(mi), g.emin);
            int const n = std::max(
                2, static_cast<int>(std::ceil(bpd * std::log10(g.emax / e0) - 1e-9)) + 1);
            LogGrid const gm{e0, g.emax, n};
            std::vector<double> v(n);
            for (int i = 0; i < n; ++i)
                v[i] = fn(materials[mi], mi, gm.energy(i));
            v[0] = 0;
            arr.arr.push_back(grid_json(gm, v));
        }
        return arr;
    };
    auto atom_sum = [&](Material const& m, auto&& micro) {
        double s = 0;
        for (auto const& c : m.comps)
            s += c.second * micro(elements[c.first]);
        return s * m.number_density;
    };
    bool const any_compound = std::any_of(materials.begin(), materials.end(),
                                          [](Material const& m) { return m.comps.size() > 1; });
    /*
     * Element-selection CDFs of a model (PhysicsParams::build_model_xs,
     * PhysicsParams.cc:638-765): for each material with several elements,
     * per element i the cumulative sum over j <= i of fraction_j * sigma_j(E)
     * normalised by the total, on the grid of the material's process table
     * (threshold(mi) is its lowest energy). Single-element materials get
     * null.
     */
    auto element_cdf = [&](auto&& threshold, auto&& micro) {
        JsonValue arr = JsonValue::array();
        for (std::size_t mi = 0; mi < nmat; ++mi)
        {
            Material const& m = materials[mi];
            if (m.comps.size() <= 1)
            {
                arr.arr.push_back(JsonValue());
                continue;
            }
            double const e0 = std::max(threshold(mi), g.emin);
            int const n = std::max(
                2, static_cast<int>(std::ceil(bpd * std::log10(g.emax / e0) - 1e-9)) + 1);
            LogGrid const gm{e0, g.emax, n};
            std::vector<std::vector<double>> cdf(m.comps.size(), std::vector<double>(n));
            for (int i = 0; i < n; ++i)
            {
                double const e = gm.energy(i);
                double cum = 0;
                for (std::size_t c = 0; c < m.comps.size(); ++c)
                {
                    cum += m.comps[c].second * micro(elements[m.comps[c].first], m, mi, e);
                    cdf[c][i] = cum;
                }
                if (cum > 0)
                {
                    for (auto& v : cdf)
                        v[i] /= cum;
                }
            }
            JsonValue jm = JsonValue::object();
            jm.obj["emin"] = JsonValue::number(gm.emin);
            jm.obj["emax"] = JsonValue::number(gm.emax);
            JsonValue jc = JsonValue::array();
            for (auto const& v : cdf)
                jc.arr.push_back(JsonValue::array(v));
            jm.obj["cdf"] = jc;
            arr.arr.push_back(jm);
        }
        return arr;
    };
    auto no_threshold = [](std::size_t) { return 0.0; };

    JsonValue processes = JsonValue::array();
    auto add_process = [&](char const* label, char const* cls, char const* particle,
                           bool integral, JsonValue models) {
        JsonValue p = JsonValue::object();
        p.obj["label"] = JsonValue::string(label);
        p.obj["class"] = JsonValue::string(cls);
        p.obj["particle"] = JsonValue::string(particle);
        p.obj["integral"] = JsonValue::boolean(integral);
        p.obj["models"] = std::move(models);
        processes.arr.push_back(std::move(p));
        return &processes.arr.back();
    };

    // Photon processes (reference creation order: Compton, photoelectric,
    // conversion)
    add_process("compt", "compton", "gamma", false,
                model_json("scat-klein-nishina", "klein_nishina"))
        ->obj["xs"] = per_material([&](Material const& m, std::size_t, double e) {
        return m.electron_density * kn_xs(e);
    });
    add_process("phot", "photoelectric", "gamma", false,
                model_json("photoel-livermore", "livermore_pe"));
    {
        JsonValue models = model_json("conv-bethe-heitler", "bethe_heitler");
        models.arr[0].obj["element_cdf"] = element_cdf(
            no_threshold,
            [](Element const& el, Material const&, std::size_t, double e) {
                return pair_xs(el, e);
            });
        add_process("conv", "conversion", "gamma", false, std::move(models))
            ->obj["xs"] = per_material([&](Material const& m, std::size_t, double e) {
            return atom_sum(m, [e](Element const& el) { return pair_xs(el, e); });
        });
    }
    // Positron annihilation (cross section computed on the fly)
    add_process("annihil", "annihilation", "positron", true,
                model_json("annihil-2-gamma", "eplusgg"));

    // Ionisation and bremsstrahlung for e- and e+
    for (bool electron : {true, false})
    {
        char const* pname = electron ? "electron" : "positron";
        // Delta rays are electrons: both leptons use the electron cut
        auto const& ecut = cut_electron;
        std::vector<std::vector<double>> dedx(nmat, std::vector<double>(npts));
        for (std::size_t mi = 0; mi < nmat; ++mi)
        {
            Material const& m = materials[mi];
            for (int i = 0; i < npts; ++i)
            {
                double const e = g.energy(i);
                double const kc = std::min(cut_gamma[mi], e);
                double brem = atom_sum(m, [&](Element const& el) {
                    return integrate_log(
                        [&](double k) { return brems_k_dsigma(el, m, e, k, enable_lpm); },
                        kc * 1e-6, kc, 6);
                });
                dedx[mi][i] = ioni_dedx(m, e, ecut[mi], electron) + brem;
            }
            // Below its low-energy maximum (a few keV) the Berger-Seltzer
            // form fails: continue below the peak as sqrt(E)
            int peak = 0;
            for (int i = 0; i < npts && g.energy(i) < 0.1; ++i)
                if (dedx[mi][i] > dedx[mi][peak])
                    peak = i;
            CELER_VALIDATE(dedx[mi][peak] > 0, << "no valid stopping power in " << m.label);
            for (int i = 0; i < peak; ++i)
                dedx[mi][i] = dedx[mi][peak] * std::sqrt(g.energy(i) / g.energy(peak));
            for (int i = peak + 1; i < npts; ++i)
                if (!(dedx[mi][i] > 0))
                    dedx[mi][i] = dedx[mi][i - 1];
        }
        JsonValue* ioni = add_process(electron ? "eIoni" : "eIoni+", "e_ioni", pname, true,
                                      model_json("ioni-moller-bhabha", "moller_bhabha"));
        ioni->obj["xs"] = per_material_from(
            [&](std::size_t mi) { return (electron ? 2 : 1) * ecut[mi]; },
            [&](Material const& m, std::size_t mi, double e) {
                return m.electron_density * ioni_xs(e, ecut[mi], electron);
            });
        JsonValue dg = JsonValue::array(), rg = JsonValue::array();
        for (std::size_t mi = 0; mi < nmat; ++mi)
        {
            dg.arr.push_back(grid_json(g, dedx[mi]));
            rg.arr.push_back(grid_json(g, range_table(g, dedx[mi])));
        }
        ioni->obj["dedx_grid"] = dg;
        ioni->obj["range_grid"] = rg;

        // One combined model for single-element materials; with compounds,
        // Seltzer-Berger below 1 GeV and relativistic above as two models
        // with element selection (the combined model rejects compounds,
        // PhysicsParams.cc:662-670)
        auto brems_micro = [&](Element const& el, Material const& m, std::size_t mi,
                               double e) {
            double const kc = cut_gamma[mi];
            if (e <= kc)
                return 0.0;
            return integrate_log([&](double k) { return brems_k_dsigma(el, m, e, k, enable_lpm) / k; },
                                 kc, e, 8);
        };
        JsonValue brem_models;
        if (any_compound)
        {
            JsonValue const cdf = element_cdf(
                [&](std::size_t mi) { return cut_gamma[mi]; }, brems_micro);
            brem_models = JsonValue::array();
            for (int k = 0; k < 2; ++k)
            {
                JsonValue mj = JsonValue::object();
                mj.obj["label"] = JsonValue::string(k == 0 ? "brems-sb" : "brems-rel");
                mj.obj["model"] = JsonValue::string(k == 0 ? "seltzer_berger"
                                                           : "relativistic_brem");
                mj.obj["low"] = JsonValue::number(k == 0 ? 0.0 : 1e3);
                mj.obj["element_cdf"] = cdf;
                brem_models.arr.push_back(std::move(mj));
            }
        }
        else
        {
            brem_models = model_json("brems-combined", "combined_brem");
        }
        JsonValue* brems = add_process(electron ? "eBrem" : "eBrem+", "e_brems", pname, true,
                                       std::move(brem_models));
        brems->obj["xs"] = per_material_from(
            [&](std::size_t mi) { return cut_gamma[mi]; },
            [&](Material const& m, std::size_t mi, double e) {
                double const kc = cut_gamma[mi];
                if (e <= kc)
                    return 0.0;
                return atom_sum(m, [&](Element const& el) {
                    return integrate_log(
                        [&](double k) { return brems_k_dsigma(el, m, e, k, enable_lpm) / k; }, kc, e, 8);
                });
            });
    }

    // Assemble
    JsonValue out = JsonValue::object();
    out.obj["particles"] = particles;
    JsonValue ej = JsonValue::array();
    for (auto const& el : elements)
    {
        JsonValue e = JsonValue::object();
        e.obj["z"] = JsonValue::number(el.z);
        e.obj["atomic_mass"] = JsonValue::number(el.atomic_mass);
        JsonValue sb = JsonValue::object();
        sb.obj["x"] = JsonValue::array(el.sb_x);
        sb.obj["y"] = JsonValue::array(el.sb_y);
        sb.obj["values"] = JsonValue::array(el.sb_values);
        e.obj["sb_table"] = sb;
        e.obj["livermore_pe"] = el.pe_data.is_null() ? pe_element(el) : el.pe_data;
        ej.arr.push_back(e);
    }
    out.obj["elements"] = ej;
    JsonValue mj = JsonValue::array();
    for (auto const& m : materials)
    {
        JsonValue o = JsonValue::object();
        o.obj["label"] = JsonValue::string(m.label);
        o.obj["number_density"] = JsonValue::number(m.number_density);
        o.obj["electron_density"] = JsonValue::number(m.electron_density);
        o.obj["density"] = JsonValue::number(m.density);
        o.obj["mean_exc_energy"] = JsonValue::number(m.mean_exc);
        o.obj["rad_length"] = JsonValue::number(m.rad_length);
        o.obj["zeff"] = JsonValue::number(m.zeff);
        JsonValue comps = JsonValue::array();
        for (auto const& c : m.comps)
        {
            JsonValue cj = JsonValue::object();
            cj.obj["element"] = JsonValue::number(double(c.first));
            cj.obj["fraction"] = JsonValue::number(c.second);
            comps.arr.push_back(cj);
        }
        o.obj["elements"] = comps;
        mj.arr.push_back(o);
    }
    out.obj["materials"] = mj;
    JsonValue cuts = JsonValue::object();
    for (auto const& pc : {std::make_pair("gamma", &cut_gamma),
                           std::make_pair("electron", &cut_electron),
                           std::make_pair("positron", &cut_positron)})
    {
        JsonValue arr = JsonValue::array();
        for (double c : *pc.second)
        {
            JsonValue o = JsonValue::object();
            o.obj["energy"] = JsonValue::number(c);
            arr.arr.push_back(o);
        }
        cuts.obj[pc.first] = arr;
    }
    out.obj["cutoffs"] = cuts;
    JsonValue options = JsonValue::object();
    options.obj["eloss_fluctuation"] = JsonValue::boolean(true);
    options.obj["lpm"] = JsonValue::boolean(enable_lpm);
    for (auto const& kv : opts.obj)
        if (kv.first != "bins_per_decade" && kv.first != "emin" && kv.first != "emax"
            && kv.first != "msc" && kv.first != "msc_step_limit_algorithm"
            && kv.first != "msc_range_factor")
            options.obj[kv.first] = kv.second;
    out.obj["options"] = options;

    // Urban MSC: transport cross sections scaled by E^2 per material
    if (opts.get_bool("msc", true))
    {
        JsonValue msc = JsonValue::object();
        msc.obj["model"] = JsonValue::string("urban");
        msc.obj["step_limit_algorithm"]
            = JsonValue::string(opts.get_str("msc_step_limit_algorithm", "safety"));
        msc.obj["range_factor"] = JsonValue::number(opts.get("msc_range_factor", 0.04));
        msc.obj["safety_factor"] = JsonValue::number(0.6);
        msc.obj["lambda_limit"] = JsonValue::number(0.1);
        JsonValue grids = per_material([&](Material const& m, std::size_t, double e) {
            return atom_sum(m, [e](Element const& el) { return msc_transport_xs(el, e); })
                   * e * e;
        });
        msc.obj["electron"] = grids;
        msc.obj["positron"] = grids;
        out.obj["msc"] = msc;
    }
    out.obj["processes"] = processes;
    return out;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
