//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/CoreParams.hh
//---------------------------------------------------------------------------//
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "corecel/DeviceRuntime.hh"
#include "geo/GeoParams.hh"
#include "phys/PhysicsParams.hh"
#include "random/XorwowParams.hh"
#include "ActionInterface.hh"
#include "CoreData.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
/*!
 * Shared problem data: geometry, physics, RNG, track-init parameters and the
 * action registry.
 *
 * Mirrors CoreParams (celeritas/global/CoreParams.hh:68-172): the
 * constructor registers the always-present actions in the reference's order
 * (CoreParams.cc:139-215) after the physics actions, and device_ref()
 * returns the POD view every kernel receives by value.
 */
class CoreParams
{
  public:
    struct Input
    {
        std::shared_ptr<GeoParams const> geometry;
        std::vector<std::uint32_t> volume_to_material;
        std::vector<std::uint32_t> volume_to_detector;  //!< calo map (opt)
        std::shared_ptr<PhysicsParams const> physics;
        std::shared_ptr<XorwowParams const> rng;
        std::shared_ptr<ActionRegistry> action_reg;
        size_type init_capacity{4096};
        size_type max_events{4096};
        TrackOrder track_order{TrackOrder::none};
        PrecisionMode precision{PrecisionMode::fp64};
        //! Sample every discrete model in one launch (the first model
        //! action launches it; the others launch nothing)
        bool fuse_interactions{true};
        //! Run the track-disjoint actions of one step order (the along-step
        //! actions; the interactions, boundary and tracking cut) on
        //! concurrent streams
        bool concurrent_actions{true};
        //! Run the discrete selection inside the along-step kernels
        bool fuse_discrete_select{true};
        //! Along-step action for charged particles (label), created by caller
        std::string along_step_label;
        //! Uniform magnetic field (enabled for the uniform-field along-step)
        FieldParamsRef field;
    };

    explicit CoreParams(Input inp);

    //! Build everything from a JSON problem description
    static std::shared_ptr<CoreParams> from_json(std::string const& json);

    CoreParamsRef const& device_ref() const { return ref_; }
    std::shared_ptr<ActionRegistry> const& action_reg() const { return reg_; }
    GeoParams const& geometry() const { return *geo_; }
    PhysicsParams const& physics() const { return *phys_; }
    XorwowParams const& rng() const { return *rng_; }
    size_type init_capacity() const { return init_capacity_; }
    size_type max_events() const { return max_events_; }
    TrackOrder track_order() const { return ref_.scalars.track_order; }
    PrecisionMode precision() const { return ref_.scalars.precision; }
    bool fuse_interactions() const { return fuse_interactions_; }
    bool concurrent_actions() const { return concurrent_actions_; }
    std::vector<std::uint32_t> const& volume_to_material() const
    {
        return vol_to_mat_;
    }
    size_type num_detectors() const { return num_detectors_; }

  private:
    std::shared_ptr<GeoParams const> geo_;
    std::shared_ptr<PhysicsParams const> phys_;
    std::shared_ptr<XorwowParams const> rng_;
    std::shared_ptr<ActionRegistry> reg_;
    size_type init_capacity_;
    size_type max_events_;
    size_type num_detectors_{0};
    bool fuse_interactions_{true};
    bool concurrent_actions_{true};
    std::vector<std::uint32_t> vol_to_mat_;
    DeviceVector<std::uint32_t> d_vol_to_mat_;
    DeviceVector<std::uint32_t> d_vol_to_det_;
    DeviceVector<XorwowJumpTables> d_jump_;
    DeviceVector<std::uint8_t> d_particle_along_;
    CoreParamsRef ref_;
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
