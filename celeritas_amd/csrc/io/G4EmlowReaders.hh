//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file io/G4EmlowReaders.hh
//! Readers for the G4EMLOW ASCII data the EM models consume.
//!
//! SeltzerBergerReader follows io/SeltzerBergerReader.cc (a "brZ" file per
//! element under G4LEDATA/brem_SB) and LivermorePEReader follows
//! io/LivermorePEReader.cc (the five pe-*-Z.dat files per element under
//! G4LEDATA/livermore/phot_epics2014). Both produce the JSON layout that
//! the physics tables use for an element ("sb_table", "livermore_pe").
//---------------------------------------------------------------------------//
#pragma once

#include <string>
#include <vector>

#include "corecel/Json.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! Scaled brems DCS of one element (ImportSBTable): x = log(E/MeV),
//! y = reduced photon energy k/E, values x-major [mb]
struct SBTable
{
    std::vector<double> x;
    std::vector<double> y;
    std::vector<double> values;
};

//---------------------------------------------------------------------------//
/*!
 * Read Seltzer-Berger tables from a directory of brZ files
 * (io/SeltzerBergerReader.cc:24-102).
 */
class SeltzerBergerReader
{
  public:
    //! Directory from $G4LEDATA/brem_SB (throws if G4LEDATA is unset)
    SeltzerBergerReader();
    //! Directory holding the brZ files
    explicit SeltzerBergerReader(std::string path);

    //! Table of element Z (1 <= Z <= 100)
    SBTable operator()(int z) const;

    std::string const& path() const { return path_; }

  private:
    std::string path_;
};

//---------------------------------------------------------------------------//
/*!
 * Read Livermore photoelectric data from a directory of pe-*-Z.dat files
 * (io/LivermorePEReader.cc:24-209): total cross sections above and below
 * the K edge, per-shell binding energies, low- and high-energy fit
 * parameters and tabulated subshell cross sections.
 *
 * The result is the element's "livermore_pe" object: xs_hi and xs_lo
 * ({"x", "y"}; xs_lo absent when its file is empty), thresh_lo, thresh_hi
 * and shells [{binding_energy, param_lo[6], param_hi[6], x, y}].
 */
class LivermorePEReader
{
  public:
    //! Directory from $G4LEDATA/livermore/phot_epics2014
    LivermorePEReader();
    //! Directory holding the pe-*-Z.dat files
    explicit LivermorePEReader(std::string path);

    JsonValue operator()(int z) const;

    std::string const& path() const { return path_; }

  private:
    std::string path_;
};

//---------------------------------------------------------------------------//
//! SB table as the element's "sb_table" JSON object
JsonValue to_json(SBTable const& t);

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
