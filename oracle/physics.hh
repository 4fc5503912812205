//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/physics.hh
//! Physics data and per-step physics actions of the oracle.
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
struct ParticleDef
{
    std::string label;
    int pdg = 0;
    double mass = 0;
    double charge = 0;
};

struct ProcessDef
{
    std::uint32_t particle = none;
    std::string label;
};

struct PhysicsData
{
    std::vector<ParticleDef> particles;
    std::vector<ProcessDef> processes;
    std::vector<std::string> model_labels;
    double secondary_stack_factor = 3;
    // Action IDs resolved after registration
    std::uint32_t pre_step_action = none;

    void finalize_actions(ActionIds const& ids);
};

//! Parse the "physics" section (same schema as the product)
void build_physics(Json const& inp, PhysicsData& out);

//! PreStepExecutor (phys/detail/PreStepExecutor.hh:46-115)
void pre_step(State& s);

//---------------------------------------------------------------------------//
}  // namespace orc
