//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file field/FieldData.hh
//! Uniform-field parameters shared by host and device.
//---------------------------------------------------------------------------//
#pragma once

#include <cstdint>

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
/*!
 * Driver options (field/FieldDriverOptions.hh:26-75), lengths in cm.
 */
struct FieldDriverOptions
{
    double minimum_step = 1.0e-6;
    double delta_chord = 0.025;
    double delta_intersection = 1.0e-5;
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
    double dchord_tol = 1e-6;
    double min_chord_shrink = 0.5;
};

//! Looping-track thresholds (track/SimData.hh:43-57)
struct LoopingThreshold
{
    std::uint32_t max_subthreshold_steps = 10;
    std::uint32_t max_steps = 100;
    double threshold_energy = 250;  //!< [MeV]
};

//! Uniform field and its propagation options
struct FieldParamsRef
{
    double field[3] = {0, 0, 0};  //!< [T]
    FieldDriverOptions options;
    LoopingThreshold looping;
    bool enabled = false;
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
