//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see em.hh).
//---------------------------------------------------------------------------//
#include "em.hh"

#include <cmath>
#include <stdexcept>

#include "problem.hh"

namespace orc
{
namespace
{
constexpr double pi = 3.14159265358979323846;

double dot3(double const* a, double const* b)
{
    double r = 0;
    for (int i = 0; i < 3; ++i)
        r = std::fma(a[i], b[i], r);
    return r;
}

//---------------------------------------------------------------------------//
/*!
 * Klein-Nishina Compton scattering (KleinNishinaInteractor.hh:104-190).
 *
 * Sampling of epsilon from the 1/eps + eps mixture with rejection
 * (Geant4 Physics Reference Manual, section "Compton scattering").
 */
Interaction klein_nishina(State& s, Track& t)
{
    PhysicsData const& pd = *s.problem->physics;
    Interaction r;
    std::uint32_t start;
    SecondaryP* sec = allocate_secondaries(s, 1, &start);
    if (!sec)
    {
        r.action = Interaction::failed;
        return r;
    }
    double const inv_mass = 1 / pd.particles[pd.electron].mass;
    double const E = t.energy;
    double const k = E * inv_mass;
    double const eps0 = 1 / (1 + 2 * k);
    // Bernoulli(-log eps0, (1 - eps0^2)/2): p = a / (a + b)
    double const a = -std::log(eps0);
    double const b = 0.5 * (1 - eps0 * eps0);
    double const p_f1 = a / (a + b);
    double const logratio = std::log(1.0 * eps0);  // Reciprocal(eps0)
    double const f2_lo = eps0 * eps0;
    double const f2_delta = 1 - f2_lo;
    double eps, omc;
    while (true)
    {
        double eps_sq;
        if (canonical_double(t.rng) < p_f1)
        {
            eps = std::exp(logratio * canonical_double(t.rng));
            eps_sq = eps * eps;
        }
        else
        {
            eps_sq = std::fma(f2_delta, canonical_double(t.rng), f2_lo);
            eps = std::sqrt(eps_sq);
        }
        omc = (1 - eps) / (eps * k);
        double sin2 = omc * (2 - omc);
        double reject = eps * sin2 / (1 + eps_sq);
        if (!(canonical_double(t.rng) < reject))
            break;
    }
    r.energy = eps * E;
    double phi = std::fma(2 * pi, canonical_double(t.rng), 0.0);
    double local[3];
    from_spherical(1 - omc, phi, local);
    rotate(local, t.geo.dir, r.dir);
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
    // Momentum conservation (InteractionUtils.hh calc_exiting_direction)
    double v[3];
    for (int i = 0; i < 3; ++i)
        v[i] = t.geo.dir[i] * E - r.dir[i] * r.energy;
    make_unit(v);
    for (int i = 0; i < 3; ++i)
        sec->dir[i] = v[i];
    return r;
}
//---------------------------------------------------------------------------//
}  // namespace

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

Interaction sample_interaction(State& s, Track& t, ModelDef const& model)
{
    if (model.model == "klein_nishina")
        return klein_nishina(s, t);
    throw std::runtime_error("oracle: unsupported model " + model.model);
}

double hardwired_xs(State const&, ProcessDef const& p, std::uint32_t, double)
{
    throw std::runtime_error("oracle: no cross section for process " + p.label);
}

std::uint32_t sample_element(State& s, Track&, ProcessDef const&, std::size_t,
                             std::uint32_t mat)
{
    PhysicsData const& pd = *s.problem->physics;
    if (mat < pd.materials.size() && pd.materials[mat].elements.size() > 1)
        throw std::runtime_error("oracle: element selection not implemented");
    return 0;
}

//---------------------------------------------------------------------------//
}  // namespace orc
