//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/PhysicsParams.cc
//---------------------------------------------------------------------------//
#include "PhysicsParams.hh"

#include <algorithm>
#include <cstring>
#include <type_traits>
#include <cmath>
#include <map>

#include "global/ActionInterface.hh"
#include "PhysicsActions.hh"

namespace celeritas_amd
{
namespace
{
//---------------------------------------------------------------------------//
ModelClass model_from_string(std::string const& s)
{
    static char const* const names[] = {"klein_nishina", "livermore_pe",
                                        "bethe_heitler", "moller_bhabha",
                                        "seltzer_berger", "relativistic_brem",
                                        "eplusgg", "combined_brem"};
    for (int i = 0; i < static_cast<int>(ModelClass::size_); ++i)
    {
        if (s == names[i])
            return static_cast<ModelClass>(i);
    }
    throw RuntimeError("unsupported model '" + s + "'");
}

ProcessClass process_from_string(std::string const& s)
{
    if (s == "compton")
        return ProcessClass::compton;
    if (s == "photoelectric")
        return ProcessClass::photoelectric;
    if (s == "conversion")
        return ProcessClass::conversion;
    if (s == "e_ioni")
        return ProcessClass::ionization;
    if (s == "e_brems")
        return ProcessClass::bremsstrahlung;
    if (s == "annihilation")
        return ProcessClass::annihilation;
    return ProcessClass::other;
}

struct GridHost
{
    double front{0}, back{0}, delta{0};
    std::uint32_t prime{invalid_id};
    std::vector<double> values;
    explicit operator bool() const { return values.size() >= 2; }
};

//! Merge "lambda"/"lambda_prim" vectors (grid/ValueGridBuilder.cc:85-172)
GridHost grid_from_vectors(JsonValue const* lo, JsonValue const* hi)
{
    GridHost g;
    double emin, emax, eprime = 0;
    if (lo && hi)
    {
        auto lx = lo->at("x").as_doubles();
        auto ly = lo->at("y").as_doubles();
        auto hx = hi->at("x").as_doubles();
        auto hy = hi->at("y").as_doubles();
        CELER_VALIDATE(ly.size() == lx.size() && hy.size() == hx.size() && lx.size() >= 1,
                       << "inconsistent lambda table sizes");
        g.values.assign(ly.begin(), ly.end() - 1);
        g.values.insert(g.values.end(), hy.begin(), hy.end());
        emin = lx.front();
        eprime = hx.front();
        emax = hx.back();
    }
    else if (hi)
    {
        auto hx = hi->at("x").as_doubles();
        g.values = hi->at("y").as_doubles();
        emin = eprime = hx.front();
        emax = hx.back();
    }
    else
    {
        auto lx = lo->at("x").as_doubles();
        g.values = lo->at("y").as_doubles();
        emin = lx.front();
        emax = lx.back();
    }
    CELER_VALIDATE(g.values.size() >= 2 && emin > 0 && emax > emin,
                   << "invalid physics vector");
    g.front = std::log(emin);
    g.back = std::log(emax);
    g.delta = (g.back - g.front) / static_cast<double>(g.values.size() - 1);
    if (eprime > 0)
    {
        double const le = std::log(eprime);
        auto p = static_cast<std::uint32_t>((le - g.front) / g.delta);
        double const gp1 = g.front + g.delta * (p + 1);
        if (std::fabs(gp1 - le) <= 1e-12 * std::max(1.0, std::fabs(le)))
            ++p;
        g.prime = p;
    }
    return g;
}

GridHost grid_from_json(JsonValue const& j)
{
    GridHost g;
    g.values = j.at("values").as_doubles();
    CELER_VALIDATE(g.values.size() >= 2, << "grid needs at least two values");
    g.front = std::log(j.at("emin").as_double());
    g.back = std::log(j.at("emax").as_double());
    g.delta = (g.back - g.front) / static_cast<double>(g.values.size() - 1);
    long long p = j.get_int("prime", -1);
    g.prime = p < 0 ? invalid_id : static_cast<std::uint32_t>(p);
    return g;
}

std::vector<GridHost> grids_of(JsonValue const& proc,
                               char const* direct,
                               char const* lo_key,
                               char const* hi_key,
                               std::size_t num_mats)
{
    std::vector<GridHost> result(num_mats);
    if (proc.has(direct))
    {
        auto const& d = proc.at(direct);
        for (std::size_t m = 0; m < num_mats && m < d.size(); ++m)
        {
            if (d[m].type == JsonValue::Type::object)
                result[m] = grid_from_json(d[m]);
        }
        return result;
    }
    JsonValue const* lo = (lo_key && proc.has(lo_key)) ? &proc.at(lo_key) : nullptr;
    JsonValue const* hi = (hi_key && proc.has(hi_key)) ? &proc.at(hi_key) : nullptr;
    if (!lo && !hi)
        return result;
    for (std::size_t m = 0; m < num_mats; ++m)
        result[m] = grid_from_vectors(lo ? &(*lo)[m] : nullptr, hi ? &(*hi)[m] : nullptr);
    return result;
}

//---------------------------------------------------------------------------//
constexpr double alpha_fine = 7.2973525693e-3;  // CODATA 2018
constexpr double r_electron_cm = 2.8179403262e-13;
constexpr double avogadro = 6.02214076e23;
constexpr double amu_gram = 1.66053906660e-24;  // CODATA 2018 [g]

/*!
 * Riemann zeta(n) - 1 for integer n >= 3.
 *
 * Direct sum over 2 <= m < 64 (smallest terms first) plus the
 * Euler-Maclaurin remainder from m = 64: the integral, the half end term
 * and the B2, B4, B6 corrections; the next term is below 1e-18.
 */
double zeta_minus_one(int n)
{
    constexpr int big_m = 64;
    double sum = 0;
    for (int m = big_m - 1; m >= 2; --m)
        sum += std::pow(static_cast<double>(m), -n);
    double const x = big_m;
    double const xn = std::pow(x, -n);
    double const dn = n;
    double tail = x * xn / (dn - 1) + xn / 2 + dn * xn / (12 * x)
                  - dn * (dn + 1) * (dn + 2) * xn / (720 * x * x * x)
                  + dn * (dn + 1) * (dn + 2) * (dn + 3) * (dn + 4) * xn
                        / (30240 * x * x * x * x * x);
    return tail + sum;
}

/*!
 * Coulomb correction f(Z) of Davies, Bethe and Maximon (1954).
 *
 * f = a^2 sum_{n>=1} 1/(n (n^2 + a^2)) with a = alpha Z, evaluated through
 * its expansion a^2 [1/(1+a^2) + sum_k (-a^2)^k (zeta(2k+3) - 1)] to 16
 * terms (exact to double precision through Z = 92).
 */
double calc_coulomb_correction(int z)
{
    double const a2 = (alpha_fine * z) * (alpha_fine * z);
    double result = 1 / (1 + a2);
    double apow = 1;
    for (int k = 0; k < 16; ++k)
    {
        result += apow * zeta_minus_one(2 * k + 3);
        apow *= -a2;
    }
    return result * a2;
}

//! Inverse radiation length per unit density [cm^2/g] (Tsai, PDG review)
double calc_mass_radiation_coeff(int z, double atomic_mass)
{
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
    return 4 * alpha_fine * r_electron_cm * r_electron_cm / (atomic_mass * amu_gram)
           * (z * z * (lrad - calc_coulomb_correction(z)) + z * lrad_prime);
}

//! Value of a grid at point i, unscaled (XsCalculator::operator[])
double grid_point_xs(GridHost const& g, std::uint32_t i)
{
    double r = g.values[i];
    if (i >= g.prime)
        r /= std::exp(g.front + g.delta * i);
    return r;
}
//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
struct PhysicsParams::HostData
{
    std::vector<ParticlePhysics> particle;
    std::vector<XsGrid> grids;
    std::vector<double> grid_reals;  //!< XsGrid values and energy_max_xs
    std::vector<double> reals;  //!< SB and Livermore tables
    std::vector<double> mass;
    std::vector<double> charge;
    std::vector<double> cutoff_energy;
    std::vector<MaterialRecord> materials;
    std::vector<ElementComponent> components;
    std::vector<ElementRecord> elements;

    DeviceVector<char> d_hot;  //!< the tables the step kernels stage in LDS
    DeviceVector<double> d_reals;
    DeviceVector<double> d_cutoff;
    DeviceVector<ElementComponent> d_components;
    DeviceVector<ElementRecord> d_elements;
    std::vector<std::uint32_t> sb_argmax;
    DeviceVector<std::uint32_t> d_sb_argmax;
    std::vector<std::uint32_t> elem_cdf_index;
    DeviceVector<std::uint32_t> d_elem_cdf_index;
    std::vector<LivermoreElement> pe_elements;
    std::vector<LivermoreShell> pe_shells;
    std::vector<UrbanFluctuation> fluct;
    std::vector<UrbanMscMaterial> msc_material;
    std::vector<UrbanMscParMat> msc_par_mat;
    std::vector<std::uint32_t> model_class;  //!< per model action
    DeviceVector<LivermoreElement> d_pe_elements;
    DeviceVector<LivermoreShell> d_pe_shells;

    //! Append a tabulated function {x, y} (increasing x)
    GenericGrid add_generic(JsonValue const& j)
    {
        auto x = j.at("x").as_doubles();
        auto y = j.at("y").as_doubles();
        CELER_VALIDATE(x.size() >= 2 && x.size() == y.size()
                           && std::is_sorted(x.begin(), x.end()),
                       << "invalid tabulated function");
        GenericGrid g{};
        g.x = static_cast<std::uint32_t>(reals.size());
        reals.insert(reals.end(), x.begin(), x.end());
        g.y = static_cast<std::uint32_t>(reals.size());
        reals.insert(reals.end(), y.begin(), y.end());
        g.size = static_cast<std::uint32_t>(x.size());
        return g;
    }

    /*!
     * Append Livermore photoelectric data {thresh_lo, thresh_hi, xs_lo,
     * xs_hi, shells: [{binding_energy, param_lo, param_hi, x, y}]} with
     * shells in order of decreasing binding energy
     * (em/model/detail/LivermoreXsInserter.hh).
     */
    void append_pe(JsonValue const* j)
    {
        LivermoreElement el{};
        el.shell_begin = static_cast<std::uint32_t>(pe_shells.size());
        if (j)
        {
            el.thresh_lo = j->at("thresh_lo").as_double();
            el.thresh_hi = j->at("thresh_hi").as_double();
            CELER_VALIDATE(el.thresh_lo <= el.thresh_hi, << "bad Livermore thresholds");
            if (j->has("xs_lo"))
                el.xs_lo = add_generic(j->at("xs_lo"));
            el.xs_hi = add_generic(j->at("xs_hi"));
            for (auto const& sj : j->at("shells").arr)
            {
                LivermoreShell sh{};
                sh.binding_energy = sj.at("binding_energy").as_double();
                auto lo = sj.at("param_lo").as_doubles();
                auto hi = sj.at("param_hi").as_doubles();
                CELER_VALIDATE(lo.size() == 6 && hi.size() == 6,
                               << "Livermore fits need 6 parameters");
                std::copy(lo.begin(), lo.end(), sh.param[0]);
                std::copy(hi.begin(), hi.end(), sh.param[1]);
                CELER_VALIDATE(sh.binding_energy <= el.thresh_lo,
                               << "binding energy above thresh_lo");
                sh.xs = add_generic(sj);
                pe_shells.push_back(sh);
            }
            el.num_shells = static_cast<std::uint32_t>(pe_shells.size()) - el.shell_begin;
            CELER_VALIDATE(el.num_shells > 0, << "Livermore data without shells");
        }
        pe_elements.push_back(el);
    }

    //! Append grids (one per material); returns first index or invalid
    std::uint32_t add_grids(std::vector<GridHost> const& gs)
    {
        if (gs.empty() || !gs[0])
            return invalid_id;
        auto first = static_cast<std::uint32_t>(grids.size());
        for (auto const& g : gs)
        {
            XsGrid x{};
            x.loge_front = g.front;
            x.loge_back = g.back;
            x.loge_delta = g.delta;
            x.size = static_cast<std::uint32_t>(g.values.size());
            x.prime_index = g.prime;
            x.offset = static_cast<std::uint32_t>(grid_reals.size());
            grid_reals.insert(grid_reals.end(), g.values.begin(), g.values.end());
            grids.push_back(x);
        }
        return first;
    }

    /*!
     * Append an SB table {x: log(E/MeV), y: k/E, values: x-major} with the
     * first maximum of each x row (SeltzerBergerModel.cc:append_table).
     */
    void append_sb_table(JsonValue const& j, ElementRecord& r)
    {
        auto x = j.at("x").as_doubles();
        auto y = j.at("y").as_doubles();
        auto v = j.at("values").as_doubles();
        CELER_VALIDATE(x.size() >= 2 && y.size() >= 2 && v.size() == x.size() * y.size(),
                       << "inconsistent Seltzer-Berger table");
        CELER_VALIDATE(std::is_sorted(x.begin(), x.end())
                           && std::is_sorted(y.begin(), y.end()),
                       << "Seltzer-Berger grids must increase");
        r.sb_nx = static_cast<std::uint32_t>(x.size());
        r.sb_ny = static_cast<std::uint32_t>(y.size());
        r.sb_x = static_cast<std::uint32_t>(reals.size());
        reals.insert(reals.end(), x.begin(), x.end());
        r.sb_y = static_cast<std::uint32_t>(reals.size());
        reals.insert(reals.end(), y.begin(), y.end());
        r.sb_values = static_cast<std::uint32_t>(reals.size());
        reals.insert(reals.end(), v.begin(), v.end());
        r.sb_argmax = static_cast<std::uint32_t>(sb_argmax.size());
        for (std::size_t i = 0; i < x.size(); ++i)
        {
            auto first = v.begin() + i * y.size();
            sb_argmax.push_back(static_cast<std::uint32_t>(
                std::max_element(first, first + y.size()) - first));
        }
    }
};

//---------------------------------------------------------------------------//
PhysicsParams::PhysicsParams(JsonValue const& inp, ActionRegistry* reg)
    : host_(std::make_unique<HostData>())
{
    HostData& h = *host_;
    PhysicsScalars& sc = ref_.scalars;
    sc = PhysicsScalars{};
    sc.min_range = 0.1;
    sc.max_step_over_range = 0.2;
    sc.min_eprime_over_e = 0.8;
    sc.linear_loss_limit = 0.01;
    sc.lowest_electron_energy = 0.001;
    sc.fixed_step_limiter = 0;
    sc.secondary_stack_factor = 3;
    sc.discrete_action = sc.range_action = sc.integral_rejection_action = invalid_id8;
    sc.failure_action = sc.msc_action = sc.fixed_step_action = invalid_id8;

    // Particles
    if (inp.has("particles"))
    {
        for (auto const& p : inp.at("particles").arr)
        {
            particles_.push_back({p.at("label").as_string(),
                                  static_cast<int>(p.get_int("pdg", 0)),
                                  p.get("mass", 0.0),
                                  p.get("charge", 0.0)});
        }
    }
    for (auto const& p : particles_)
    {
        h.mass.push_back(p.mass);
        h.charge.push_back(p.charge);
    }
    ref_.gamma_id = this->find_particle(22);
    ref_.electron_id = this->find_particle(11);
    ref_.positron_id = this->find_particle(-11);

    // Elements and materials
    if (inp.has("elements"))
    {
        for (auto const& e : inp.at("elements").arr)
        {
            ElementRecord r{};
            r.z = e.get("z", 0.0);
            r.atomic_mass = e.get("atomic_mass", 0.0);
            r.cbrt_z = std::cbrt(r.z);
            r.cbrt_zzp = std::cbrt(r.z * (r.z + 1));
            r.log_z = r.z > 0 ? std::log(r.z) : 0;
            int const z = static_cast<int>(r.z);
            r.coulomb_correction = z > 0 ? calc_coulomb_correction(z) : 0;
            r.mass_radiation_coeff
                = z > 0 ? calc_mass_radiation_coeff(z, r.atomic_mass) : 0;
            if (z >= 1 && ref_.electron_id != invalid_id)
            {
                // Thomas-Fermi form factors; Dirac-Fock values for Z < 5
                // (RelativisticBremModel.cc:150-200)
                static double const dirac_fock[4][2] = {
                    {5.3104, 5.9173}, {4.7935, 5.6125}, {4.7402, 5.5377}, {4.7112, 5.4728}};
                double const me = particles_[ref_.electron_id].mass;
                double const ff_el = z < 5 ? dirac_fock[z - 1][0]
                                           : std::log(184.15) - r.log_z / 3;
                double const ff_inel = z < 5 ? dirac_fock[z - 1][1]
                                             : std::log(1194.0) - 2 * r.log_z / 3;
                double const fc = r.coulomb_correction;
                double const invz = 1.0 / z;
                r.rb_fz = r.log_z / 3 + fc;
                r.rb_factor1 = (ff_el - fc) + ff_inel * invz;
                r.rb_factor2 = (1 + invz) / 12;
                r.rb_gamma_factor = 100 * me / r.cbrt_z;
                r.rb_epsilon_factor = 100 * me / (r.cbrt_z * r.cbrt_z);
            }
            if (e.has("sb_table"))
                h.append_sb_table(e.at("sb_table"), r);
            h.append_pe(e.has("livermore_pe") ? &e.at("livermore_pe") : nullptr);
            h.elements.push_back(r);
        }
    }
    if (inp.has("materials"))
    {
        for (auto const& m : inp.at("materials").arr)
        {
            MaterialRecord r{};
            r.number_density = m.get("number_density", 0.0);
            r.electron_density = m.get("electron_density", 0.0);
            r.density = m.get("density", 0.0);
            r.mean_exc_energy = m.get("mean_exc_energy", 0.0);
            r.log_mean_exc_energy
                = r.mean_exc_energy > 0 ? std::log(r.mean_exc_energy) : 0;
            // Stored as exp(log I), as MaterialParams does
            if (r.mean_exc_energy > 0)
                r.mean_exc_energy = std::exp(r.log_mean_exc_energy);
            r.rad_length = m.get("rad_length", 0.0);
            r.zeff = m.get("zeff", 0.0);
            r.elem_begin = static_cast<std::uint32_t>(h.components.size());
            if (m.has("elements"))
            {
                for (auto const& c : m.at("elements").arr)
                {
                    ElementComponent ec{};
                    ec.element = static_cast<std::uint32_t>(c.get_int("element", 0));
                    ec.fraction = c.get("fraction", 1.0);
                    h.components.push_back(ec);
                }
            }
            r.num_elements = static_cast<std::uint32_t>(h.components.size()) - r.elem_begin;
            if (r.rad_length == 0 && r.number_density > 0 && r.num_elements > 0)
            {
                // Derived as MaterialParams does (MaterialParams.cc:496-519):
                // 1 / (sum_i w_i (1/X0)_i rho), rho = n sum_i w_i A_i amu
                double avg_amu = 0, rad_coeff = 0;
                for (std::uint32_t i = r.elem_begin; i < h.components.size(); ++i)
                {
                    ElementComponent const& ec = h.components[i];
                    CELER_VALIDATE(ec.element < h.elements.size(),
                                   << "material refers to an unknown element");
                    ElementRecord const& el = h.elements[ec.element];
                    avg_amu += ec.fraction * el.atomic_mass;
                    rad_coeff += ec.fraction * el.mass_radiation_coeff;
                }
                double const density = r.number_density * avg_amu * amu_gram;
                r.rad_length = 1 / (rad_coeff * density);
            }
            CELER_VALIDATE(r.num_elements <= static_cast<std::uint32_t>(max_elements),
                           << "too many elements in a material");
            h.materials.push_back(r);
        }
    }
    num_materials_ = static_cast<size_type>(std::max<std::size_t>(h.materials.size(), 1));

    // Options
    if (inp.has("options"))
    {
        auto const& o = inp.at("options");
        sc.secondary_stack_factor = o.get("secondary_stack_factor", sc.secondary_stack_factor);
        sc.min_range = o.get("min_range", sc.min_range);
        sc.max_step_over_range = o.get("max_step_over_range", sc.max_step_over_range);
        sc.min_eprime_over_e = o.get("min_eprime_over_e", sc.min_eprime_over_e);
        sc.linear_loss_limit = o.get("linear_loss_limit", sc.linear_loss_limit);
        sc.lowest_electron_energy
            = o.get("lowest_electron_energy", sc.lowest_electron_energy);
        sc.fixed_step_limiter = o.get("fixed_step_limiter", 0.0);
        sc.apply_post_interaction_cut = o.get_bool("apply_cuts", false);
        sc.enable_lpm = o.get_bool("lpm", false) ? 1 : 0;
    }

    // Cutoffs
    h.cutoff_energy.assign(particles_.size() * num_materials_, 0.0);
    if (inp.has("cutoffs"))
    {
        for (auto const& kv : inp.at("cutoffs").obj)
        {
            std::uint32_t pid = this->find_particle(kv.first);
            CELER_VALIDATE(pid != invalid_id, << "cutoff for unknown particle " << kv.first);
            for (std::size_t m = 0; m < num_materials_ && m < kv.second.size(); ++m)
                h.cutoff_energy[pid * num_materials_ + m] = kv.second[m].get("energy", 0.0);
        }
    }

    // Processes: actions first (PhysicsParams.cc:97-143 order)
    std::vector<JsonValue const*> procs;
    if (inp.has("processes"))
    {
        for (auto const& p : inp.at("processes").arr)
            procs.push_back(&p);
    }
    if (!procs.empty())
    {
        CELER_VALIDATE(reg, << "physics requires an action registry");
        reg->insert(std::make_shared<PreStepAction>(reg->next_id()));
        sc.msc_action = static_cast<std::uint8_t>(reg->next_id());
        reg->insert(std::make_shared<ImplicitAction>(
            reg->next_id(), "msc-range", "limit range due to multiple scattering"));
        sc.range_action = static_cast<std::uint8_t>(reg->next_id());
        sc.fixed_step_action = sc.range_action;
        reg->insert(std::make_shared<ImplicitAction>(
            reg->next_id(), "eloss-range", "limit range due to energy loss"));
        sc.discrete_action = static_cast<std::uint8_t>(reg->next_id());
        reg->insert(std::make_shared<DiscreteSelectAction>(reg->next_id()));
        sc.integral_rejection_action = static_cast<std::uint8_t>(reg->next_id());
        reg->insert(std::make_shared<ImplicitAction>(
            reg->next_id(), "physics-integral-rejected", "reject by integral cross section"));
    }

    h.particle.assign(particles_.size(), ParticlePhysics{});
    for (auto& pp : h.particle)
    {
        pp.eloss_process = -1;
        pp.dedx_grid = pp.range_grid = pp.inv_range_grid = pp.msc_grid = invalid_id;
    }
    std::uint32_t num_models = 0;
    std::map<std::string, std::pair<std::uint32_t, ModelClass>> model_actions;
    for (JsonValue const* pj : procs)
    {
        std::string const plabel = pj->get_str("particle", "");
        std::uint32_t pid = this->find_particle(plabel);
        CELER_VALIDATE(pid != invalid_id, << "process for unknown particle '" << plabel << "'");
        ParticlePhysics& pp = h.particle[pid];
        CELER_VALIDATE(pp.num_processes < static_cast<std::uint32_t>(max_processes),
                       << "too many processes for particle " << plabel);
        ProcessRecord& rec = pp.process[pp.num_processes];
        rec = ProcessRecord{};
        rec.process_class = static_cast<std::uint32_t>(
            process_from_string(pj->get_str("class", "")));
        rec.integral = pj->get_bool("integral", false) ? 1 : 0;
        if (rec.process_class == static_cast<std::uint32_t>(ProcessClass::annihilation))
        {
            // e+ annihilation applies at rest (EPlusAnnihilationProcess)
            pp.has_at_rest = 1;
        }
        if (rec.process_class == static_cast<std::uint32_t>(ProcessClass::photoelectric))
        {
            for (auto const& c : h.components)
                CELER_VALIDATE(h.pe_elements.at(c.element).num_shells > 0,
                               << "photoelectric process needs Livermore data for "
                                  "every element in use");
        }

        auto xs = grids_of(*pj, "xs", "lambda", "lambda_prim", num_materials_);
        rec.xs_grid = h.add_grids(xs);
        rec.energy_max_xs = invalid_id;
        if (rec.integral)
        {
            rec.energy_max_xs = static_cast<std::uint32_t>(h.grid_reals.size());
            for (std::size_t m = 0; m < num_materials_; ++m)
            {
                double e_max = 0, xs_max = 0;
                GridHost const& g = xs[m];
                if (g && rec.process_class != static_cast<std::uint32_t>(ProcessClass::annihilation))
                {
                    for (std::uint32_t i = 0; i < g.values.size(); ++i)
                    {
                        double v = grid_point_xs(g, i);
                        if (v > xs_max)
                        {
                            xs_max = v;
                            e_max = std::exp(g.front + g.delta * i);
                        }
                    }
                }
                h.grid_reals.push_back(e_max);
            }
        }
        auto dedx = grids_of(*pj, "dedx_grid", "dedx", nullptr, num_materials_);
        auto range = grids_of(*pj, "range_grid", "range", nullptr, num_materials_);
        if (!dedx.empty() && dedx[0] && range[0])
        {
            pp.eloss_process = static_cast<std::int32_t>(pp.num_processes);
            pp.dedx_grid = h.add_grids(dedx);
            pp.range_grid = h.add_grids(range);
        }

        auto const& models = pj->at("models");
        CELER_VALIDATE(models.size() >= 1 && models.size() <= static_cast<std::size_t>(max_models),
                       << "process must have 1-" << max_models << " models");
        rec.num_models = static_cast<std::uint32_t>(models.size());
        for (std::size_t k = 0; k < models.size(); ++k)
        {
            auto const& mj = models[k];
            ModelClass mc = model_from_string(mj.at("model").as_string());
            std::string const mlabel = mj.at("label").as_string();
            if (mc == ModelClass::seltzer_berger || mc == ModelClass::combined_brem)
            {
                for (auto const& c : h.components)
                    CELER_VALIDATE(h.elements.at(c.element).sb_nx >= 2,
                                   << "model '" << mlabel
                                   << "' needs a Seltzer-Berger table for every element in use");
            }
            rec.models[k].energy_low = mj.get("low", 0.0);
            rec.models[k].model = static_cast<std::uint32_t>(mc);
            rec.models[k].elem_cdf = invalid_id;
            if (mc == ModelClass::combined_brem)
            {
                // PhysicsParams.cc:662-670: no element selection for it
                for (std::size_t m = 0; m < num_materials_; ++m)
                    CELER_VALIDATE(h.materials[m].num_elements <= 1,
                                   << "model '" << mlabel
                                   << "' cannot be used with materials of more than one "
                                      "element");
            }
            if (mj.has("element_cdf"))
            {
                // Per material: one CDF grid per element for compounds
                auto const& cj = mj.at("element_cdf");
                CELER_VALIDATE(cj.size() == num_materials_,
                               << "element_cdf of '" << mlabel << "' needs one entry per material");
                rec.models[k].elem_cdf = static_cast<std::uint32_t>(h.elem_cdf_index.size());
                for (std::size_t m = 0; m < num_materials_; ++m)
                {
                    std::uint32_t first = invalid_id;
                    auto const& mjm = cj[m];
                    if (mjm.type == JsonValue::Type::object)
                    {
                        auto const& rows = mjm.at("cdf");
                        CELER_VALIDATE(rows.size() == h.materials[m].num_elements,
                                       << "element_cdf needs one grid per element");
                        std::vector<GridHost> gs;
                        for (auto const& row : rows.arr)
                        {
                            GridHost gh;
                            gh.values = row.as_doubles();
                            CELER_VALIDATE(gh.values.size() >= 2, << "short CDF grid");
                            gh.front = std::log(mjm.at("emin").as_double());
                            gh.back = std::log(mjm.at("emax").as_double());
                            gh.delta = (gh.back - gh.front)
                                       / static_cast<double>(gh.values.size() - 1);
                            gs.push_back(std::move(gh));
                        }
                        first = h.add_grids(gs);
                    }
                    h.elem_cdf_index.push_back(first);
                }
            }
            // One model (and action) may serve several particles, as the
            // reference's MollerBhabha and brems models serve e- and e+
            auto found = model_actions.find(mlabel);
            if (found != model_actions.end())
            {
                CELER_VALIDATE(found->second.second == mc,
                               << "model '" << mlabel << "' reused with another class");
                rec.models[k].action = found->second.first;
                continue;
            }
            rec.models[k].action = reg->next_id();
            if (num_models == 0)
                sc.first_model_action = static_cast<std::uint8_t>(reg->next_id());
            h.model_class.push_back(static_cast<std::uint32_t>(mc));
            model_actions[mlabel] = {reg->next_id(), mc};
            reg->insert(std::make_shared<ModelAction>(
                reg->next_id(), mlabel, static_cast<std::uint32_t>(mc)));
            ++num_models;
        }
        process_labels_.push_back(pj->get_str("label", ""));
        ++pp.num_processes;
    }
    sc.num_models = num_models;
    if (!procs.empty())
    {
        sc.failure_action = static_cast<std::uint8_t>(reg->next_id());
        reg->insert(std::make_shared<ImplicitAction>(
            reg->next_id(), "physics-failure", "mark a track that failed to sample an interaction"));
    }

    // Model constants
    if (ref_.electron_id != invalid_id)
    {
        ref_.model.electron_mass = particles_[ref_.electron_id].mass;
        ref_.model.inv_electron_mass = 1 / ref_.model.electron_mass;
    }

    // Energy-loss fluctuations (FluctuationParams.cc): two oscillators
    if (inp.has("options") && inp.at("options").get_bool("eloss_fluctuation", false))
    {
        CELER_VALIDATE(ref_.electron_id != invalid_id,
                       << "energy-loss fluctuations need electrons");
        ref_.fluct_enabled = 1;
        for (auto const& m : h.materials)
        {
            UrbanFluctuation f{};
            double const avg_z = m.number_density > 0 ? m.electron_density / m.number_density
                                                      : 0;
            f.oscillator_strength[1] = avg_z > 2 ? 2 / avg_z : 0;
            f.oscillator_strength[0] = 1 - f.oscillator_strength[1];
            f.binding_energy[1] = 1e-5 * (avg_z * avg_z);
            f.binding_energy[0] = std::pow(
                m.mean_exc_energy / std::pow(f.binding_energy[1], f.oscillator_strength[1]),
                1 / f.oscillator_strength[0]);
            f.log_binding_energy[1] = std::log(f.binding_energy[1]);
            f.log_binding_energy[0] = std::log(f.binding_energy[0]);
            h.fluct.push_back(f);
        }
    }

    // Urban multiple scattering (UrbanMscParams.cc)
    if (inp.has("msc"))
    {
        auto const& mj = inp.at("msc");
        CELER_VALIDATE(mj.get_str("model", "urban") == "urban",
                       << "only the Urban MSC model is available");
        std::string const alg = mj.get_str("step_limit_algorithm", "safety");
        CELER_VALIDATE(alg == "minimal" || alg == "safety" || alg == "safety_plus",
                       << "unknown MSC step limit algorithm '" << alg
                       << "' (minimal, safety or safety_plus)");
        CELER_VALIDATE(ref_.electron_id != invalid_id && ref_.positron_id != invalid_id,
                       << "MSC needs electrons and positrons");
        MscScalars& ms = ref_.msc;
        ms.step_limit_algorithm = alg == "minimal" ? 0 : (alg == "safety" ? 1 : 2);
        ms.range_factor = mj.get("range_factor", 0.04);
        ms.safety_factor = mj.get("safety_factor", 0.6);
        ms.lambda_limit = mj.get("lambda_limit", 0.1);
        CELER_VALIDATE(ms.range_factor > 0 && ms.range_factor < 1 && ms.safety_factor >= 0.1
                           && ms.lambda_limit > 0,
                       << "invalid MSC options");
        auto eg = grids_of(mj, "electron", nullptr, nullptr, num_materials_);
        auto pg = grids_of(mj, "positron", nullptr, nullptr, num_materials_);
        CELER_VALIDATE(!eg.empty() && eg[0] && !pg.empty() && pg[0],
                       << "MSC needs cross-section grids for e- and e+");
        std::vector<GridHost> interleaved;
        for (std::size_t m = 0; m < num_materials_; ++m)
        {
            CELER_VALIDATE(eg[m] && pg[m], << "missing MSC grid for material " << m);
            interleaved.push_back(eg[m]);
            interleaved.push_back(pg[m]);
        }
        ms.xs_grid = h.add_grids(interleaved);
        ms.low_energy_limit = std::exp(eg[0].front);
        ms.high_energy_limit = std::exp(eg[0].back);
        ms.enabled = 1;
        auto poly2 = [](double c0, double c1, double c2, double x) {
            return std::fma(x, std::fma(x, c2, c1), c0);
        };
        auto fastpow = [](double a, double b) { return std::exp(b * std::log(a)); };
        for (auto const& m : h.materials)
        {
            double const zeff = m.zeff;
            CELER_VALIDATE(zeff > 0, << "MSC needs the effective Z of every material");
            UrbanMscMaterial d{};
            d.stepmin_coeff[0] = 1e3 * 27.725 / (1 + 0.203 * zeff);
            d.stepmin_coeff[1] = 1e3 * 6.152 / (1 + 0.111 * zeff);
            double const z16 = fastpow(zeff, 1.0 / 6.0);
            double const fz = poly2(0.990395, -0.168386, 0.093286, z16);
            d.theta_coeff[0] = fz * (1 - 8.7780e-2 / zeff);
            d.theta_coeff[1] = fz * (4.0780e-2 + 1.7315e-4 * zeff);
            double const z13 = z16 * z16;
            d.tail_coeff[0] = poly2(2.3785, -4.1981e-1, 6.3100e-2, z13);
            d.tail_coeff[1] = poly2(4.7526e-1, 1.7694, -3.3885e-1, z13);
            d.tail_coeff[2] = poly2(2.3683e-1, -1.8111, 3.2774e-1, z13);
            d.tail_corr = poly2(1.7888e-2, 1.9659e-2, -2.6664e-3, z13);
            h.msc_material.push_back(d);
            static double const a_coeff[2] = {0.87, 0.70};
            static double const b_coeff[2] = {2.0 / 3, 1.0 / 2};
            for (int p = 0; p < 2; ++p)
            {
                UrbanMscParMat pm{};
                pm.scaled_zeff = a_coeff[p] * fastpow(zeff, b_coeff[p]);
                pm.d_over_r = 9.6280e-1 - 8.4848e-2 * std::sqrt(zeff) + 4.3769e-3 * zeff;
                h.msc_par_mat.push_back(pm);
            }
        }
    }

    // Upload. The small tables the step kernels read per track (particle
    // and process records, grid headers and values, material, MSC and
    // fluctuation records, masses and charges) share one contiguous region
    // that a kernel can copy into LDS and rebase into.
    std::vector<char> hot;
    auto place = [&hot](auto const& v) {
        std::size_t const off = (hot.size() + 15) / 16 * 16;
        std::size_t const n = v.size() * sizeof(v[0]);
        hot.resize(off + n);
        if (n)
            std::memcpy(hot.data() + off, v.data(), n);
        return off;
    };
    std::size_t const o_particle = place(h.particle);
    std::size_t const o_grids = place(h.grids);
    std::size_t const o_grid_reals = place(h.grid_reals);
    std::size_t const o_mass = place(h.mass);
    std::size_t const o_charge = place(h.charge);
    std::size_t const o_materials = place(h.materials);
    std::size_t const o_msc_material = place(h.msc_material);
    std::size_t const o_msc_par_mat = place(h.msc_par_mat);
    std::size_t const o_fluct = place(h.fluct);
    std::size_t const o_model_class = place(h.model_class);
    hot.resize((hot.size() + 15) / 16 * 16);
    h.d_hot = to_device(hot);
    char const* base = h.d_hot.data();
    auto at = [base](std::size_t off, auto const& v) {
        using T = std::remove_const_t<std::remove_reference_t<decltype(v[0])>>;
        return v.empty() ? static_cast<T const*>(nullptr)
                         : reinterpret_cast<T const*>(base + off);
    };
    ref_.hot_base = base;
    ref_.hot_bytes = static_cast<std::uint32_t>(hot.size());
    ref_.particle = at(o_particle, h.particle);
    ref_.grids = at(o_grids, h.grids);
    ref_.grid_reals = at(o_grid_reals, h.grid_reals);
    ref_.mass = at(o_mass, h.mass);
    ref_.charge = at(o_charge, h.charge);
    ref_.materials = at(o_materials, h.materials);
    ref_.msc_material = at(o_msc_material, h.msc_material);
    ref_.msc_par_mat = at(o_msc_par_mat, h.msc_par_mat);
    ref_.fluct = at(o_fluct, h.fluct);
    ref_.model_class = at(o_model_class, h.model_class);

    h.d_reals = to_device(h.reals);
    h.d_cutoff = to_device(h.cutoff_energy);
    h.d_components = to_device(h.components);
    h.d_elements = to_device(h.elements);
    h.d_sb_argmax = to_device(h.sb_argmax);
    ref_.sb_argmax = h.d_sb_argmax.data();
    h.d_elem_cdf_index = to_device(h.elem_cdf_index);
    ref_.elem_cdf_index = h.d_elem_cdf_index.data();
    h.d_pe_elements = to_device(h.pe_elements);
    h.d_pe_shells = to_device(h.pe_shells);
    ref_.pe_elements = h.d_pe_elements.data();
    ref_.pe_shells = h.d_pe_shells.data();
    ref_.num_particles = this->num_particles();
    ref_.num_materials = num_materials_;
    ref_.reals = h.d_reals.data();
    ref_.cutoff_energy = h.d_cutoff.data();
    ref_.components = h.d_components.data();
    ref_.elements = h.d_elements.data();
}

PhysicsParams::~PhysicsParams() = default;

std::uint32_t PhysicsParams::find_particle(int pdg) const
{
    for (std::size_t i = 0; i < particles_.size(); ++i)
    {
        if (particles_[i].pdg == pdg)
            return static_cast<std::uint32_t>(i);
    }
    return invalid_id;
}

std::uint32_t PhysicsParams::find_particle(std::string const& label) const
{
    for (std::size_t i = 0; i < particles_.size(); ++i)
    {
        if (particles_[i].label == label)
            return static_cast<std::uint32_t>(i);
    }
    return invalid_id;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
