//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker by tests/, smoke() and the
// cpu_baseline leg of bench.py. Never linked into the product library.
//---------------------------------------------------------------------------//
//! \file oracle/em.hh
//! EM interactors, hardwired cross sections and element selection.
//---------------------------------------------------------------------------//
#pragma once

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

//! Secondary allocation on the state's stack (StackAllocator semantics)
SecondaryP* allocate_secondaries(State& s, std::uint32_t n, std::uint32_t* start);

//! Sample the interaction of the given model for a track
Interaction sample_interaction(State& s, Track& t, ModelDef const& model);

//! Cross sections computed on the fly (LivermorePE, EPlusGG)
double hardwired_xs(State const& s, ProcessDef const& p, std::uint32_t mat, double e);

//! Sample an element component for multi-element materials
std::uint32_t sample_element(State& s, Track& t, ProcessDef const& p,
                             std::size_t model_index, std::uint32_t mat);

//---------------------------------------------------------------------------//
// Math helpers shared by interactors (corecel/math/ArrayUtils.hh)
void rotate(double const* dir, double const* rot, double* out);
void from_spherical(double costheta, double phi, double* out);
void make_unit(double* v);

//---------------------------------------------------------------------------//
}  // namespace orc
