//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/PhysicsParams.hh
//---------------------------------------------------------------------------//
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "corecel/DeviceRuntime.hh"
#include "corecel/Json.hh"
#include "PhysicsData.hh"

namespace celeritas_amd
{
class ActionRegistry;

//---------------------------------------------------------------------------//
/*!
 * Particle, material, cutoff and physics-table data.
 *
 * Same responsibilities as ParticleParams, MaterialParams, CutoffParams and
 * PhysicsParams (celeritas/phys/PhysicsParams.cc:83-190): registers the
 * physics actions (pre-step, msc-range, eloss-range, discrete-select,
 * integral-rejected, one per model, failure) in the reference's order and
 * builds flat cross-section tables.
 *
 * Input is the "physics" section of a problem description. A process gives
 * its macroscopic cross sections per material either as Geant4-style
 * "lambda" (unscaled, low energy) and "lambda_prim" (scaled by E) vectors
 * on log-spaced energies, or directly as "xs" grids {emin, emax, prime,
 * values}; energy-loss processes add "dedx"/"range" vectors or grids.
 */
class PhysicsParams
{
  public:
    //! Construct from the "physics" section of a problem description
    PhysicsParams(JsonValue const& inp, ActionRegistry* reg);
    ~PhysicsParams();

    PhysicsParamsRef const& device_ref() const { return ref_; }

    size_type num_particles() const { return static_cast<size_type>(particles_.size()); }
    size_type num_materials() const { return num_materials_; }
    std::uint32_t find_particle(int pdg) const;
    std::uint32_t find_particle(std::string const& label) const;
    double mass(std::uint32_t pid) const { return particles_.at(pid).mass; }
    double charge(std::uint32_t pid) const { return particles_.at(pid).charge; }
    std::vector<std::string> const& process_labels() const { return process_labels_; }
    bool has_processes() const { return !process_labels_.empty(); }

    struct Particle
    {
        std::string label;
        int pdg;
        double mass;
        double charge;
    };

    //! Host data block mirrored on device
    struct HostData;
    HostData const& host_data() const { return *host_; }

  private:
    std::vector<Particle> particles_;
    size_type num_materials_{0};
    std::vector<std::string> process_labels_;
    std::unique_ptr<HostData> host_;
    PhysicsParamsRef ref_;
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
