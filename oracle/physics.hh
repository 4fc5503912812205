//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/physics.hh
//! Physics data and the per-step physics actions of the oracle, restated
//! from celeritas/phys/detail/{PreStep,DiscreteSelect,TrackingCut}Executor.hh,
//! celeritas/phys/PhysicsStepUtils.hh, celeritas/phys/PhysicsTrackView.hh,
//! celeritas/grid/XsCalculator.hh, celeritas/phys/InteractionApplier.hh,
//! celeritas/geo/detail/BoundaryExecutor.hh and the along-step appliers in
//! celeritas/global/alongstep/detail/*.hh.
//!
//! Pinned by: XsCalculator.test.cc:38-60 (closed-form grids) and the
//! SimpleCompton whole-loop anchor (Stepper.test.cc:194-225: 919 step
//! iterations, 53.8125 steps per primary, queue high-water mark {1, 6},
//! emptying step 3).
//---------------------------------------------------------------------------//
#pragma once

#include <string>
#include <vector>

#include "core.hh"
#include "json.hh"

namespace orc
{
struct ActionIds;

//---------------------------------------------------------------------------//
//! Uniform-in-log(E) grid; values at i >= prime are scaled by E
struct XsGridO
{
    double front = 0, back = 0, delta = 0;
    std::uint32_t size = 0;
    std::uint32_t prime = 0xffffffffu;
    std::vector<double> values;

    explicit operator bool() const { return size >= 2; }
    double calc(double energy) const;  // XsCalculator::operator()
    double at(std::uint32_t i) const;  // XsCalculator::operator[]
    double energy(std::uint32_t i) const;
};

struct ParticleDef
{
    std::string label;
    int pdg = 0;
    double mass = 0;
    double charge = 0;
};

//! Seltzer-Berger scaled DCS table: x = log(E/MeV), y = k/E, values x-major
struct SBTableO
{
    std::vector<double> x, y, values;
    std::vector<std::uint32_t> argmax;  // per x: first y index of the max

    explicit operator bool() const { return !x.empty(); }
};

//! Tabulated function on a nonuniform grid (GenericCalculator)
struct GenericO
{
    std::vector<double> x, y;
    explicit operator bool() const { return !x.empty(); }
    double operator()(double v) const;
};

//! Livermore photoelectric data (em/data/LivermorePEData.hh)
struct LivermoreO
{
    struct Shell
    {
        double binding = 0;
        double param[2][6] = {};
        GenericO xs;
    };
    double thresh_lo = 0, thresh_hi = 0;
    GenericO xs_lo, xs_hi;
    std::vector<Shell> shells;
    explicit operator bool() const { return !shells.empty(); }
};

struct ElementDef
{
    LivermoreO pe;
    int z = 0;
    double atomic_mass = 0;  // [amu]
    double cbrt_z = 0;
    double log_z = 0;
    double coulomb_correction = 0;
    double mass_radiation_coeff = 0;  // [cm^2/g]
    SBTableO sb;
    // Relativistic brems element data (RelativisticBremModel.cc:150-185)
    double rb_fz = 0, rb_factor1 = 0, rb_factor2 = 0;
    double rb_gamma_factor = 0, rb_epsilon_factor = 0;
};

struct MaterialDef
{
    std::string label;
    double number_density = 0;  // [1/cm^3]
    double electron_density = 0;  // [1/cm^3]
    double mean_exc_energy = 0;  // [MeV], exp(log I)
    double log_mean_exc_energy = 0;
    double rad_length = 0;  // [cm]
    double zeff = 0;
    std::vector<std::pair<std::uint32_t, double>> elements;  // (id, fraction)
};

//! Urban MSC data (em/data/UrbanMscData.hh, params/UrbanMscParams.cc)
//! Step-limit algorithms of Urban MSC (MscStepLimitAlgorithm)
enum class MscLimitAlg
{
    minimal = 0,
    safety = 1,
    safety_plus = 2,
};

struct MscDataO
{
    struct Material
    {
        double stepmin[2] = {0, 0};
        double theta[2] = {0, 0};
        double tail[3] = {0, 0, 0};
        double tail_corr = 0;
        double scaled_zeff[2] = {0, 0};  // e-, e+
        double d_over_r = 0;
    };
    bool enabled = false;
    double low = 0, high = 0;  // energy limits [MeV]
    double range_factor = 0.04, safety_factor = 0.6, lambda_limit = 0.1;
    MscLimitAlg limit_alg = MscLimitAlg::safety;
    std::vector<XsGridO> xs[2];  // [e-, e+][material]: sigma_tr * E^2
    std::vector<Material> mat;
};

//! Urban fluctuation oscillators (FluctuationParams.cc)
struct FluctO
{
    double binding[2] = {0, 0};
    double log_binding[2] = {0, 0};
    double strength[2] = {0, 0};
};

struct ModelDef
{
    std::string label;
    std::string model;  // klein_nishina, ...
    double low = 0, high = 0;
    std::uint32_t action = none;
};

struct ProcessDef
{
    std::string label;
    std::string cls;
    std::uint32_t particle = none;
    bool integral = false;
    std::vector<XsGridO> xs;  // per material
    std::vector<double> energy_max_xs;  // per material (integral)
    std::vector<XsGridO> dedx, range;  // per material (energy loss)
    std::vector<ModelDef> models;
    //! Element CDF grids [model][material][element] (multi-element only)
    std::vector<std::vector<std::vector<XsGridO>>> elem_cdf;
};

struct PhysicsData
{
    std::vector<ParticleDef> particles;
    std::vector<ElementDef> elements;
    std::vector<MaterialDef> materials;
    std::vector<ProcessDef> processes;
    std::vector<std::vector<std::uint32_t>> particle_processes;  // per particle
    std::vector<int> eloss_process;  // per particle: index into its list
    std::vector<std::string> model_labels;
    // cutoffs[particle][material]
    std::vector<std::vector<double>> cutoff_energy;
    // Options (PhysicsParamsOptions defaults)
    double secondary_stack_factor = 3;
    double min_range = 0.1;
    double max_step_over_range = 0.2;
    double min_eprime_over_e = 0.8;
    double linear_loss_limit = 0.01;
    double lowest_electron_energy = 0.001;
    double fixed_step_limiter = 0;
    bool apply_post_interaction_cut = false;
    //! LPM suppression in the relativistic brems and Bethe-Heitler models
    //! (GeantPhysicsOptions::lpm)
    bool enable_lpm = false;
    // Along-step models
    MscDataO msc;
    bool fluct = false;
    std::vector<FluctO> fluct_params;  // per material
    // Resolved action IDs
    std::uint32_t msc_action = none;
    std::uint32_t discrete_action = none, range_action = none;
    std::uint32_t integral_rejection_action = none, failure_action = none;
    std::uint32_t fixed_step_action = none;
    std::uint32_t gamma = none, electron = none, positron = none;

    void finalize_actions(ActionIds const& ids);
};

//! Parse the "physics" section (same schema as the product)
void build_physics(Json const& inp, PhysicsData& out);

//! Range and inverse range on a range grid (grid/RangeCalculator.hh,
//! grid/InverseRangeCalculator.hh)
double calc_range(XsGridO const& g, double energy);
//! Planted mistake of the grid calculators for the pin tests (0: none)
extern int g_grid_mutation;
double calc_inverse_range(XsGridO const& g, double range);

// Step actions (each loops over all slots, applying to matching tracks)
void pre_step(State& s);
void along_step_neutral(State& s);
void discrete_select(State& s);
void interact(State& s, std::uint32_t model_action);
void boundary(State& s);
void tracking_cut(State& s);
void along_step_charged(State& s);
//! SimpleCalo tally of per-volume energy deposition (user/detail/
//! SimpleCaloExecutor.hh:48-66, StepGatherExecutor.hh:41-153)
void calo_tally(State& s);

//! Steps of the full sequence in execution order
void step(State& s);

//---------------------------------------------------------------------------//
}  // namespace orc
