//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/TableGen.hh
//! Physics-table generator: builds the cross-section, stopping-power and
//! range tables of the standard EM processes from published formulas.
//---------------------------------------------------------------------------//
#pragma once

#include <string>

#include "corecel/Json.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
/*!
 * Expand a compact physics description into a full "physics" input.
 *
 * The reference imports these tables from Geant4 (PhysicsParams::build_xs,
 * celeritas/phys/PhysicsParams.cc:445-636; io/ImportPhysicsTable.hh). This
 * generator computes them instead, on the same layout (uniform in log E,
 * default 7 bins per decade from 100 eV to 100 TeV):
 *
 * - Compton: Klein-Nishina total cross section per electron (closed form).
 * - Pair production: the Bethe-Heitler DCS with the screening functions and
 *   Coulomb correction that the interactor samples from, integrated over
 *   the positron energy fraction.
 * - Photoelectric: per-element data in the Livermore layout (binding
 *   energies, subshell tables, cumulative 1/E polynomial fits), built from
 *   a synthetic shell model sigma_s(E) = w_s C (1 + E/m)^2 / E^3 above each
 *   edge, calibrated to one total cross section per element.
 * - Ionisation: Moller/Bhabha cross sections above the production cut
 *   (closed-form integrals of the DCS the interactor samples) and the
 *   Berger-Seltzer restricted stopping power (ICRU 37) with the Sternheimer
 *   density effect (Sternheimer-Peierls general parameters).
 * - Bremsstrahlung: Tsai's DCS (the relativistic model's) with dielectric
 *   suppression, tabulated as a Seltzer-Berger-format scaled DCS below
 *   1 GeV; cross section above the photon cut and restricted energy loss
 *   below it by Gauss-Legendre quadrature of the same DCS the sampler uses.
 * - Range: integral of 1/(dE/dx) with 2 E0/(dE/dx)(E0) below the grid.
 *
 * Absolute agreement with Geant4's tables is not claimed.
 */
JsonValue generate_physics(JsonValue const& spec);

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
