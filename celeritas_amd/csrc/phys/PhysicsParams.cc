//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file phys/PhysicsParams.cc
//---------------------------------------------------------------------------//
#include "PhysicsParams.hh"

#include "global/ActionInterface.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
struct PhysicsParams::HostData
{
    std::vector<double> mass;
    std::vector<double> charge;
    DeviceVector<double> d_mass;
    DeviceVector<double> d_charge;
};

//---------------------------------------------------------------------------//
PhysicsParams::PhysicsParams(JsonValue const& inp, ActionRegistry*)
    : host_(std::make_unique<HostData>())
{
    if (inp.has("particles"))
    {
        for (auto const& p : inp.at("particles").arr)
        {
            particles_.push_back({p.at("label").as_string(),
                                  static_cast<int>(p.get_int("pdg", 0)),
                                  p.get("mass", 0.0),
                                  p.get("charge", 0.0)});
        }
    }
    for (auto const& p : particles_)
    {
        host_->mass.push_back(p.mass);
        host_->charge.push_back(p.charge);
    }
    host_->d_mass = to_device(host_->mass);
    host_->d_charge = to_device(host_->charge);
    ref_.num_particles = this->num_particles();
    ref_.mass = host_->d_mass.data();
    ref_.charge = host_->d_charge.data();
    ref_.gamma_id = this->find_particle(22);
    ref_.electron_id = this->find_particle(11);
    ref_.positron_id = this->find_particle(-11);
}

PhysicsParams::~PhysicsParams() = default;

std::uint32_t PhysicsParams::find_particle(int pdg) const
{
    for (std::size_t i = 0; i < particles_.size(); ++i)
    {
        if (particles_[i].pdg == pdg)
            return static_cast<std::uint32_t>(i);
    }
    return invalid_id;
}

std::uint32_t PhysicsParams::find_particle(std::string const& label) const
{
    for (std::size_t i = 0; i < particles_.size(); ++i)
    {
        if (particles_[i].label == label)
            return static_cast<std::uint32_t>(i);
    }
    return invalid_id;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
