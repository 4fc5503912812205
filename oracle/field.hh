//----------------------------------*-C++-*----------------------------------//
// ORACLE (test infrastructure only): CPU restatement of the reference's
// charged-particle propagation in a magnetic field. Used by tests/ and
// smoke() as the checker; never by the product.
//---------------------------------------------------------------------------//
//! \file oracle/field.hh
//---------------------------------------------------------------------------//
#pragma once

#include "geo.hh"

namespace orc
{
//---------------------------------------------------------------------------//
//! FieldDriverOptions defaults (field/FieldDriverOptions.hh:26-75), in cm
struct FieldOptions
{
    double minimum_step = 1.0e-6;  // 1e-5 mm
    double delta_chord = 0.025;  // 0.25 mm
    double delta_intersection = 1.0e-5;  // 1e-4 mm
    double epsilon_step = 1.0e-5;
    double epsilon_rel_max = 1.0e-3;
    double pgrow = -0.20;
    double pshrink = -0.25;
    double safety = 0.9;
    double max_stepping_increase = 5;
    double max_stepping_decrease = 0.1;
    int max_nsteps = 100;
    int max_substeps = 10;
    double initial_step_tol = 1e-6;
    double dchord_tol = 1e-6;  // 1e-5 mm
    double min_chord_shrink = 0.5;
};

//! Position [cm] and momentum [MeV/c]
struct OdeState
{
    double pos[3];
    double mom[3];
};

//! Dormand-Prince RK5(4) result (field/Types.hh FieldStepperResult)
struct StepperResult
{
    OdeState mid_state;
    OdeState end_state;
    OdeState err_state;
};

//! Curvature per unit field for unit charge and momentum: 1/(cm T) per MeV/c
inline constexpr double lorentz_per_cm_tesla = 2.99792458;

//! Uniform field equation of motion and its Dormand-Prince stepper
struct UniformFieldStepper
{
    double field[3];  //!< [T]
    double coeff;  //!< charge [e] * lorentz_per_cm_tesla
    mutable long count = 0;  //!< number of steps taken (for tests)

    OdeState rhs(OdeState const& y) const;
    StepperResult operator()(double step, OdeState const& beg) const;
};

//! Propagation result (geocel/Types.hh Propagation)
struct FieldPropagation
{
    double distance = 0;
    bool boundary = false;
    bool looping = false;
};

/*!
 * State kept by one propagator object: the ODE state and the driver's
 * chord-length memory (FieldPropagator::state_, FieldDriver::max_chord_).
 * Transport builds a fresh one per step; the reference's unit tests reuse
 * one propagator over several calls.
 */
struct FieldPropagatorState
{
    OdeState state;
    double max_chord;
};

//! Fresh propagator state from the geometry position and direction
FieldPropagatorState make_field_state(GeoState const& g, double momentum);

/*!
 * Propagate a charged track in a uniform field by up to `step`
 * (FieldPropagator.hh:150-331 with FieldDriver.hh): the geometry state is
 * moved along chords, the direction set from the final momentum.
 */
FieldPropagation field_propagate(Geometry const& geo,
                                 GeoState& g,
                                 UniformFieldStepper const& stepper,
                                 FieldOptions const& opts,
                                 FieldPropagatorState& ps,
                                 double step);

//! Propagation with a fresh propagator (as in the along-step action)
FieldPropagation field_propagate(Geometry const& geo,
                                 GeoState& g,
                                 UniformFieldStepper const& stepper,
                                 double momentum,
                                 FieldOptions const& opts,
                                 double step);

//---------------------------------------------------------------------------//
}  // namespace orc
