//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/em.hh
//! EM interactors, hardwired cross sections and element selection.
//!
//! Interactors are templates over an engine returning uniform doubles so
//! the same code runs with the transport XORWOW stream and with the
//! std::mt19937 + std::generate_canonical stream of the reference's
//! interactor unit tests (test/celeritas/phys/InteractorHostTestBase.hh),
//! whose golden values pin them:
//! - KleinNishina.test.cc:89-146 (energies and angles, 4 samples)
//! - BetheHeitler.test.cc:90-150 (Cu, 100 MeV)
//! - MollerBhabha.test.cc:108-200
//! - EPlusGG.test.cc:82-240
//---------------------------------------------------------------------------//
#pragma once

#include <cstdint>
#include <random>
#include <stdexcept>

#include "physics.hh"

namespace orc
{
//---------------------------------------------------------------------------//
//! Result of an interactor (celeritas/phys/Interaction.hh)
struct Interaction
{
    enum Action
    {
        scattered,
        absorbed,
        unchanged,
        failed
    } action = scattered;
    double energy = 0;
    double dir[3] = {0, 0, 0};
    double energy_deposition = 0;
    std::uint32_t sec_start = 0;
    std::uint32_t num_secondaries = 0;
};

//! Uniform doubles from a track's XORWOW state
struct XorwowUniform
{
    XorwowState& s;
    double operator()() { return canonical_double(s); }
};

//! Uniform doubles from std::mt19937 as std::generate_canonical makes them,
//! counting the 32-bit draws (DiagnosticRngEngine)
struct MtUniform
{
    std::mt19937& mt;
    std::uint64_t count = 0;
    struct Counting
    {
        using result_type = std::mt19937::result_type;
        std::mt19937& mt;
        std::uint64_t& count;
        static constexpr result_type min() { return std::mt19937::min(); }
        static constexpr result_type max() { return std::mt19937::max(); }
        result_type operator()()
        {
            ++count;
            return mt();
        }
    };
    double operator()()
    {
        Counting c{mt, count};
        return std::generate_canonical<double, 53>(c);
    }
};

//! Uniform doubles replayed from a buffer (device parity tests)
struct ReplayUniform
{
    double const* u;
    std::uint32_t i = 0;
    std::uint32_t size = 0xffffffffu;  //!< uniforms available to this sample
    double operator()()
    {
        if (i >= size)
            throw std::runtime_error("oracle: replayed uniform stream exhausted");
        return u[i++];
    }
};

//! Inputs common to all interactors
struct InteractorInput
{
    PhysicsData const* pd = nullptr;
    std::uint32_t particle = none;
    double energy = 0;
    double dir[3] = {0, 0, 1};
    std::uint32_t material = 0;
    std::uint32_t element = 0;  // component index in the material
};

//! Secondary allocation on the state's stack (StackAllocator semantics)
SecondaryP* allocate_secondaries(State& s, std::uint32_t n, std::uint32_t* start);

//! Sample one interaction of a model (by model name)
template<class Engine>
Interaction sample_model(std::string const& model, InteractorInput const& in,
                         State& s, Engine& rng);

//! Sample the interaction of the given model for a track
Interaction sample_interaction(State& s, Track& t, ModelDef const& model);

//! Cross sections computed on the fly (LivermorePE, EPlusGG)
double hardwired_xs(State const& s, ProcessDef const& p, std::uint32_t mat, double e);

//! Sample an element component for multi-element materials
std::uint32_t sample_element(State& s, Track& t, ProcessDef const& p,
                             std::size_t model_index, std::uint32_t mat);

//! Element constants derived from Z (mat/detail/Utils.cc equivalents)
double coulomb_correction(int z);
double mass_radiation_coeff(int z, double atomic_mass);

//---------------------------------------------------------------------------//
// Bremsstrahlung (SeltzerBerger, RelativisticBrem, CombinedBrem)
//! Build an SB table with its per-x argmax (SeltzerBergerModel.cc:append_table)
SBTableO make_sb_table(Json const& j);
//! Relativistic brems element data (RelativisticBremModel.cc:150-185)
void set_rb_element_data(double electron_mass, ElementDef& el);
//! Relativistic DCS without LPM (RBDiffXsCalculator::dxsec_per_atom)
double rb_dxsec(ElementDef const& el, double total_energy, double gamma_energy);

//! LPM suppression functions (em/xs/LPMCalculator.hh)
struct LpmFunctions
{
    double xi = 0, g = 0, phi = 0;
};
//! Planted mistake of the LPM functions for the pin tests (0: none)
extern int g_lpm_mutation;
//! Characteristic LPM energy per unit radiation length [MeV/cm]
double lpm_constant();
//! xi, G and phi at epsilon = (lepton energy) / (photon energy)
LpmFunctions lpm_functions(ElementDef const& el, MaterialDef const& mat,
                           bool dielectric_suppression, double gamma_energy,
                           double epsilon);

//! Relativistic DCS of an incident lepton (RBDiffXsCalculator): with LPM
//! when enabled and the total energy exceeds the material's threshold
struct RbDxsec
{
    ElementDef const& el;
    MaterialDef const& mat;
    double total_energy;
    double density_corr = 0;
    bool enable_lpm = false;

    RbDxsec(ElementDef const& el, MaterialDef const& mat, bool lpm, double total);
    double operator()(double gamma_energy) const;
};
//! Positron DCS scaling (SBPositronXsCorrector)
double sb_positron_scale(int z, double mass, double kc, double inc_energy,
                         double exit_energy);
//! SB photon energy sampler (SBEnergySampler / SBEnergyDistribution);
//! corrector 0 = electron, 1 = positron (SBPositronXsCorrector), 2 = a
//! constant one-half; also reports the bounding max_xs
template<class Engine>
double sample_sb_energy(ElementDef const& el, double electron_density,
                        double inc_energy, double mass, double kc,
                        int corrector, Engine& rng, double* max_xs = nullptr);
//! Migdal constant 4 pi r_e lambdabar_e^2 [cm^3]
double migdal_constant();

//---------------------------------------------------------------------------//
// Livermore photoelectric (LivermorePEInteractor.hh, LivermorePEMicroXs-
// Calculator.hh); pinned by LivermorePE.test.cc:137-203 and :484-517
LivermoreO make_livermore(Json const& j);
//! Micro cross section [barn]
double livermore_pe_micro_xs(ElementDef const& el, double energy);
//! Macro cross section [1/cm] (MacroXsCalculator)
double livermore_pe_macro_xs(PhysicsData const& pd, std::uint32_t mat, double energy);

//---------------------------------------------------------------------------//
// Math helpers shared by interactors (corecel/math/ArrayUtils.hh)
void rotate(double const* dir, double const* rot, double* out);
void from_spherical(double costheta, double phi, double* out);
void make_unit(double* v);

//---------------------------------------------------------------------------//
}  // namespace orc
