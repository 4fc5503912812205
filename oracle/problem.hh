//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/problem.hh
//! Problem description shared by the oracle's actions: geometry, action IDs,
//! track-initialization parameters and (when present) physics data.
//---------------------------------------------------------------------------//
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "field.hh"
#include "geo.hh"
#include "json.hh"

namespace orc
{
struct PhysicsData;

//---------------------------------------------------------------------------//
struct ActionIds
{
    std::uint32_t pre_step = none, msc_range = none, eloss_range = none;
    std::uint32_t discrete_select = none, integral_rejected = none;
    std::uint32_t failure = none;
    std::vector<std::uint32_t> model;  // per model index
    std::uint32_t along_neutral = none, along_user = none;
    std::uint32_t initialize_tracks = none, propagation_limit = none;
    std::uint32_t boundary = none, tracking_cut = none;
    std::uint32_t extend_secondaries = none, extend_primaries = none;
    std::vector<std::string> labels;
};

struct Problem
{
    Geometry geo;
    std::vector<std::uint32_t> volume_to_material;
    std::vector<std::uint32_t> volume_to_detector;
    std::uint32_t num_detectors = 0;
    std::uint32_t seed = 12345;
    std::uint32_t init_capacity = 4096;
    std::uint32_t max_events = 4096;
    bool in_place = true;
    double secondary_stack_factor = 3;
    ActionIds actions;
    std::shared_ptr<PhysicsData const> physics;
    // Uniform field for the uniform-field along-step (SimData.hh:43-57
    // looping thresholds)
    bool field_enabled = false;
    double field[3] = {0, 0, 0};  //!< [T]
    FieldOptions field_options;
    std::uint32_t looping_max_subthreshold_steps = 10;
    std::uint32_t looping_max_steps = 100;
    double looping_threshold_energy = 250;  //!< [MeV]

    //! Build from the same JSON problem description as the product
    static std::shared_ptr<Problem> from_json(std::string const& text);
};

//---------------------------------------------------------------------------//
}  // namespace orc
