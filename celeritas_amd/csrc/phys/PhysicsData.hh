//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/PhysicsData.hh
//! Physics tables and per-slot physics state.
//!
//! Counterparts of PhysicsParamsData/PhysicsStateData
//! (celeritas/phys/PhysicsData.hh:403-503), MaterialParamsData
//! (celeritas/mat/MaterialData.hh), ParticleParamsData and CutoffParamsData.
//! The per-particle process list is stored inline (at most max_processes
//! processes per particle, max_models models per process) so that a track
//! reaches all of its physics descriptors with one 512-byte record load;
//! grids are small descriptors into one array of reals.
//---------------------------------------------------------------------------//
#pragma once

#include "corecel/Types.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
inline constexpr int max_processes = 4;  //!< per particle type
inline constexpr int max_models = 3;  //!< per process
inline constexpr int max_elements = 8;  //!< per material

//! Interaction models with device implementations
enum class ModelClass : std::uint32_t
{
    klein_nishina = 0,
    livermore_pe,
    bethe_heitler,
    moller_bhabha,
    seltzer_berger,
    relativistic_brem,
    eplusgg,
    combined_brem,
    size_
};

//! Process classes with special handling
enum class ProcessClass : std::uint32_t
{
    other = 0,
    compton,
    photoelectric,
    conversion,
    ionization,
    bremsstrahlung,
    annihilation,
};

//! Log-uniform grid with values (celeritas/grid/XsGridData.hh:28-50)
struct XsGrid
{
    double loge_front;  //!< log(E_min / MeV)
    double loge_back;  //!< log(E_max / MeV)
    double loge_delta;
    std::uint32_t size;  //!< number of points (0: no grid)
    std::uint32_t prime_index;  //!< values at i >= prime are scaled by E
    std::uint32_t offset;  //!< into the reals array
    std::uint32_t pad_;
};

//! One model of a process with its lower energy bound
struct ModelRange
{
    double energy_low;  //!< [MeV]
    std::uint32_t model;  //!< ModelClass
    std::uint32_t action;  //!< step action ID of the model
    //! Offset into elem_cdf_index (one entry per material: first XsGrid of
    //! the material's per-element CDFs, or invalid), or invalid
    std::uint32_t elem_cdf;
    std::uint32_t pad_;
};

//! One particle-process: tables indexed by material
struct ProcessRecord
{
    std::uint32_t process_class;  //!< ProcessClass
    std::uint32_t xs_grid;  //!< first XsGrid index (+material), or invalid
    std::uint32_t integral;  //!< 1 if the integral approach is used
    std::uint32_t energy_max_xs;  //!< reals offset (+material)
    std::uint32_t num_models;
    ModelRange models[max_models];
};

//! Per-particle physics
struct ParticlePhysics
{
    std::uint32_t num_processes;
    std::int32_t eloss_process;  //!< index of the process with dE/dx, or -1
    std::uint32_t dedx_grid;  //!< first XsGrid (+material), or invalid
    std::uint32_t range_grid;  //!< first XsGrid (+material), or invalid
    std::uint32_t inv_range_grid;  //!< first XsGrid (+material), or invalid
    std::uint32_t msc_grid;  //!< first XsGrid (+material), or invalid
    std::uint32_t has_at_rest;
    std::uint32_t pad_;
    ProcessRecord process[max_processes];
};

//---------------------------------------------------------------------------//
//! Element properties (celeritas/mat/ElementView.hh)
struct ElementRecord
{
    double z;
    double atomic_mass;  //!< [amu]
    double cbrt_z;
    double cbrt_zzp;  //!< (Z(Z+1))^(1/3)
    double log_z;
    double coulomb_correction;  //!< f(Z) of Davies-Bethe-Maximon
    double mass_radiation_coeff;  //!< 1/X0 per unit density [cm^2/g]
    // Relativistic brems (RelativisticBremModel.cc:150-185)
    double rb_fz;
    double rb_factor1;
    double rb_factor2;
    double rb_gamma_factor;
    double rb_epsilon_factor;
    // Seltzer-Berger table: offsets into reals (x = log E, y = k/E, values
    // x-major) and into sb_argmax; sb_nx == 0 when absent
    std::uint32_t sb_x;
    std::uint32_t sb_y;
    std::uint32_t sb_values;
    std::uint32_t sb_argmax;
    std::uint32_t sb_nx;
    std::uint32_t sb_ny;
};

//! Tabulated function on a nonuniform grid: offsets into reals
struct GenericGrid
{
    std::uint32_t x;
    std::uint32_t y;
    std::uint32_t size;  //!< 0: absent
    std::uint32_t pad_;
};

//! Livermore photoelectric subshell (em/data/LivermorePEData.hh)
struct LivermoreShell
{
    double binding_energy;  //!< [MeV]
    double param[2][6];  //!< cumulative 1/E fits below/above thresh_hi [b]
    GenericGrid xs;  //!< sigma * E^3 [b MeV^3] vs E [MeV]
};

//! Livermore photoelectric element data
struct LivermoreElement
{
    GenericGrid xs_lo;  //!< total sigma E^3 below the K edge
    GenericGrid xs_hi;  //!< total sigma E^3 from the K edge to thresh_lo
    double thresh_lo;  //!< [MeV]
    double thresh_hi;  //!< [MeV]
    std::uint32_t shell_begin;
    std::uint32_t num_shells;  //!< 0: no data
};

//! Urban MSC material coefficients (UrbanMscParams::calc_material_data)
struct UrbanMscMaterial
{
    double stepmin_coeff[2];
    double theta_coeff[2];
    double tail_coeff[3];
    double tail_corr;
};

//! Urban MSC particle-material data: [material][e-, e+]
struct UrbanMscParMat
{
    double scaled_zeff;  //!< a Z^b
    double d_over_r;  //!< maximum distance over range
};

//! Urban energy-loss fluctuation oscillators per material
//! (em/data/FluctuationData.hh, FluctuationParams.cc)
struct UrbanFluctuation
{
    double binding_energy[2];  //!< [MeV]
    double log_binding_energy[2];
    double oscillator_strength[2];
};

//! Multiple-scattering options and table index
struct MscScalars
{
    double low_energy_limit{0};  //!< [MeV]
    double high_energy_limit{0};  //!< [MeV]
    double range_factor{0.04};
    double safety_factor{0.6};
    double lambda_limit{0.1};  //!< [cm]
    std::uint32_t xs_grid{invalid_id};  //!< first XsGrid: [material][e-, e+]
    std::uint8_t enabled{0};
    //! MscStepLimitAlgorithm: 0 minimal, 1 safety, 2 safety_plus
    std::uint8_t step_limit_algorithm{1};
    std::uint8_t pad_[2]{};
};

//! Component of a material
struct ElementComponent
{
    std::uint32_t element;
    std::uint32_t pad_;
    double fraction;  //!< number fraction
};

//! Material properties (celeritas/mat/MaterialView.hh)
struct MaterialRecord
{
    double number_density;  //!< atoms / cm^3
    double electron_density;  //!< electrons / cm^3
    double density;  //!< g / cm^3
    double mean_exc_energy;  //!< [MeV]
    double log_mean_exc_energy;
    double rad_length;  //!< [cm]
    double zeff;
    std::uint32_t elem_begin;
    std::uint32_t num_elements;
};

//---------------------------------------------------------------------------//
//! Tracking and step-limit scalars (celeritas/phys/PhysicsData.hh:271-330)
struct PhysicsScalars
{
    double min_range;  //!< rho [cm]
    double max_step_over_range;  //!< alpha
    double min_eprime_over_e;  //!< xi
    double linear_loss_limit;
    double lowest_electron_energy;  //!< [MeV]
    double fixed_step_limiter;  //!< [cm] (0 = off)
    double secondary_stack_factor;
    std::uint32_t num_models;
    std::uint8_t apply_post_interaction_cut;
    std::uint8_t discrete_action;
    std::uint8_t range_action;
    std::uint8_t integral_rejection_action;
    std::uint8_t failure_action;
    std::uint8_t msc_action;
    std::uint8_t fixed_step_action;
    std::uint8_t first_model_action;  //!< model actions are consecutive
    //! LPM suppression in the relativistic brems and Bethe-Heitler models
    std::uint8_t enable_lpm;
};

//! Per-model constants (filled for the models in the problem)
struct ModelData
{
    double electron_mass;  //!< [MeV]
    double inv_electron_mass;
};

//---------------------------------------------------------------------------//
/*!
 * Immutable physics data (device pointers).
 */
struct PhysicsParamsRef
{
    PhysicsScalars scalars{};
    std::uint32_t num_particles{0};
    std::uint32_t num_materials{0};
    ParticlePhysics const* particle{nullptr};
    XsGrid const* grids{nullptr};
    double const* grid_reals{nullptr};  //!< XsGrid values, energy_max_xs
    double const* reals{nullptr};  //!< SB and Livermore tables
    // Particle properties
    double const* mass{nullptr};  //!< [MeV/c^2]
    double const* charge{nullptr};  //!< [e]
    std::uint32_t gamma_id{invalid_id};
    std::uint32_t electron_id{invalid_id};
    std::uint32_t positron_id{invalid_id};
    std::uint32_t pad_{0};
    // Cutoffs [particle * num_materials + material]
    double const* cutoff_energy{nullptr};
    // Materials
    MaterialRecord const* materials{nullptr};
    ElementComponent const* components{nullptr};
    //! Per model and material: first element-CDF grid (ModelRange::elem_cdf)
    std::uint32_t const* elem_cdf_index{nullptr};
    //! Model class of each model action (action - first_model_action)
    std::uint32_t const* model_class{nullptr};
    ElementRecord const* elements{nullptr};
    std::uint32_t const* sb_argmax{nullptr};  //!< per-x argmax of SB tables
    LivermoreElement const* pe_elements{nullptr};  //!< per element
    LivermoreShell const* pe_shells{nullptr};
    // Along-step models
    MscScalars msc{};
    UrbanMscMaterial const* msc_material{nullptr};  //!< per material
    UrbanMscParMat const* msc_par_mat{nullptr};  //!< [material][e-, e+]
    UrbanFluctuation const* fluct{nullptr};  //!< per material
    std::uint32_t fluct_enabled{0};
    ModelData model{};
    //! Contiguous region holding particle, grids, grid_reals, mass, charge,
    //! materials, msc_material, msc_par_mat, fluct and model_class
    //! (16-byte aligned)
    char const* hot_base{nullptr};
    std::uint32_t hot_bytes{0};
};

//---------------------------------------------------------------------------//
/*!
 * Per-slot physics state (SoA).
 *
 * Step-local fields (macro_xs, per_process_xs, energy deposition, MSC step)
 * are only written by the unfused per-action kernels; the fused step kernel
 * keeps them in registers.
 */
struct PhysicsStateRef
{
    double* interaction_mfp{nullptr};  //!< remaining MFPs (<=0: interact)
    double* macro_xs{nullptr};
    double* per_process_xs{nullptr};  //!< [max_processes][size]
    double* dedx_range{nullptr};
    double* energy_deposition{nullptr};
    std::uint32_t* sec_start{nullptr};  //!< first secondary on the stack
    std::uint32_t* sec_count{nullptr};
    std::uint8_t* element{nullptr};  //!< sampled element component
    // Urban MSC persistent range data (celeritas/em/data/UrbanMscData.hh)
    double* msc_range_init{nullptr};
    double* msc_range_factor{nullptr};
    double* msc_limit_min{nullptr};
    // MSC step data
    double* msc_true_path{nullptr};
    double* msc_geom_path{nullptr};
    double* msc_alpha{nullptr};
    std::uint8_t* msc_is_displaced{nullptr};
    std::uint8_t* msc_enabled{nullptr};
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
