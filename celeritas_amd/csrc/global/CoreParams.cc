//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/CoreParams.cc
//---------------------------------------------------------------------------//
#include "CoreParams.hh"

#include "phys/TableGen.hh"
#include "track/TrackActions.hh"
#include "CoreActions.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
CoreParams::CoreParams(Input inp)
    : geo_(std::move(inp.geometry))
    , phys_(std::move(inp.physics))
    , rng_(std::move(inp.rng))
    , reg_(std::move(inp.action_reg))
    , init_capacity_(inp.init_capacity)
    , max_events_(inp.max_events)
{
    CELER_VALIDATE(geo_ && phys_ && rng_ && reg_, << "incomplete core params input");
    CELER_VALIDATE(init_capacity_ > 0 && max_events_ > 0,
                   << "track initializer capacity and max events must be positive");
    vol_to_mat_ = inp.volume_to_material;
    vol_to_mat_.resize(geo_->num_volumes(), invalid_id);
    d_vol_to_mat_ = to_device(vol_to_mat_);

    auto det = inp.volume_to_detector;
    det.resize(geo_->num_volumes(), invalid_id);
    for (auto d : det)
    {
        if (d != invalid_id)
            num_detectors_ = std::max<size_type>(num_detectors_, d + 1);
    }
    d_vol_to_det_ = to_device(det);

    std::vector<XorwowJumpTables> jt{rng_->tables()};
    d_jump_ = to_device(jt);

    // Always-present actions (CoreParams.cc:139-215 order)
    CoreScalars& sc = ref_.scalars;
    sc.initialize_tracks_action = static_cast<std::uint8_t>(reg_->next_id());
    reg_->insert(std::make_shared<InitializeTracksAction>(reg_->next_id()));

    ActionId user_along = inp.along_step_label.empty()
                              ? invalid_id
                              : reg_->find_action(inp.along_step_label);
    ActionId neutral = reg_->find_action("along-step-neutral");
    if (neutral == invalid_id)
    {
        neutral = reg_->next_id();
        reg_->insert(std::make_shared<AlongStepNeutralAction>(neutral));
    }
    sc.along_step_neutral_action = static_cast<std::uint8_t>(neutral);
    sc.along_step_user_action = static_cast<std::uint8_t>(
        user_along == invalid_id ? neutral : user_along);

    sc.propagation_limit_action = static_cast<std::uint8_t>(reg_->next_id());
    reg_->insert(std::make_shared<ImplicitAction>(
        reg_->next_id(), "geo-propagation-limit", "pause due to propagation misses"));
    sc.boundary_action = static_cast<std::uint8_t>(reg_->next_id());
    reg_->insert(std::make_shared<BoundaryAction>(reg_->next_id()));
    sc.tracking_cut_action = static_cast<std::uint8_t>(reg_->next_id());
    reg_->insert(std::make_shared<TrackingCutAction>(reg_->next_id()));
    reg_->insert(std::make_shared<ExtendFromSecondariesAction>(reg_->next_id()));
    if (reg_->find_action(ExtendFromPrimariesAction::label_str()) == invalid_id)
    {
        reg_->insert(std::make_shared<ExtendFromPrimariesAction>(reg_->next_id()));
    }
    if (num_detectors_ > 0)
    {
        reg_->insert(std::make_shared<CaloTallyAction>(reg_->next_id()));
    }
    // Optional track-sorting actions (CoreParams.cc:259-283): keys are
    // action IDs (plus one for inactive slots), particle IDs or status
    {
        std::vector<TrackOrder> sorts;
        switch (inp.track_order)
        {
            case TrackOrder::reindex_status:
            case TrackOrder::reindex_particle_type:
            case TrackOrder::reindex_along_step_action:
            case TrackOrder::reindex_step_limit_action:
                sorts.push_back(inp.track_order);
                break;
            case TrackOrder::reindex_both_action:
                sorts.push_back(TrackOrder::reindex_step_limit_action);
                sorts.push_back(TrackOrder::reindex_along_step_action);
                break;
            default:
                break;
        }
        size_type const total_actions = reg_->num_actions() + sorts.size();
        for (TrackOrder t : sorts)
        {
            size_type keys = total_actions + 1;
            if (t == TrackOrder::reindex_status)
                keys = 2;
            else if (t == TrackOrder::reindex_particle_type)
                keys = phys_->num_particles() + 1;
            reg_->insert(std::make_shared<SortTracksAction>(reg_->next_id(), t, keys));
        }
    }
    sc.num_actions = reg_->num_actions();
    sc.track_order = inp.track_order;
    sc.precision = inp.precision;
    fuse_interactions_ = inp.fuse_interactions;
    concurrent_actions_ = inp.concurrent_actions;
    sc.fuse_discrete_select = inp.fuse_discrete_select ? 1 : 0;

    ref_.geo = geo_->device_ref();
    ref_.volume_to_material = d_vol_to_mat_.data();
    ref_.volume_to_detector = d_vol_to_det_.data();
    ref_.phys = phys_->device_ref();
    ref_.field = inp.field;
    ref_.rng_jump = d_jump_.data();
    ref_.rng_seed = rng_->seed();

    // The pre-step's along-step action of each particle type (by charge)
    std::vector<std::uint8_t> along(std::max<size_type>(phys_->num_particles(), 1),
                                    invalid_id8);
    for (size_type p = 0; p < phys_->num_particles(); ++p)
    {
        along[p] = phys_->charge(p) == 0 ? sc.along_step_neutral_action
                                         : sc.along_step_user_action;
    }
    d_particle_along_ = to_device(along);
    ref_.particle_along_action = d_particle_along_.data();
}

//---------------------------------------------------------------------------//
std::shared_ptr<CoreParams> CoreParams::from_json(std::string const& text)
{
    JsonValue const doc = parse_json(text);
    Input inp;
    inp.action_reg = std::make_shared<ActionRegistry>();

    JsonValue const& geo_doc = doc.at("geometry");
    inp.geometry = std::make_shared<GeoParams>(geo_input_from_orange_json(geo_doc));

    // Physics first: its actions take the lowest IDs as in the reference
    JsonValue empty;
    empty.type = JsonValue::Type::object;
    JsonValue const& phys_in = doc.has("physics") ? doc.at("physics") : empty;
    // A compact description ({"generate": ...}) is expanded into tables
    JsonValue generated;
    if (phys_in.has("generate"))
        generated = generate_physics(phys_in.at("generate"));
    JsonValue const& phys_doc = phys_in.has("generate") ? generated : phys_in;
    inp.physics = std::make_shared<PhysicsParams>(phys_doc, inp.action_reg.get());

    // Along-step action for charged particles
    std::string along = doc.get_str("along_step", "");
    if (!along.empty())
    {
        inp.along_step_label = make_along_step_action(along, *inp.physics, doc,
                                                      inp.action_reg.get());
    }
    if (along == "uniform_field")
    {
        // Field [T], driver options (lengths in cm) and looping thresholds
        CELER_VALIDATE(doc.has("field"), << "uniform_field along-step needs 'field'");
        JsonValue const& f = doc.at("field");
        auto b = f.at("value").as_doubles();
        CELER_VALIDATE(b.size() == 3, << "field value must have 3 components");
        FieldParamsRef& fp = inp.field;
        for (int i = 0; i < 3; ++i)
            fp.field[i] = b[i];
        fp.enabled = true;
        if (f.has("options"))
        {
            JsonValue const& o = f.at("options");
            FieldDriverOptions& d = fp.options;
            d.minimum_step = o.get("minimum_step", d.minimum_step);
            d.delta_chord = o.get("delta_chord", d.delta_chord);
            d.delta_intersection = o.get("delta_intersection", d.delta_intersection);
            d.epsilon_step = o.get("epsilon_step", d.epsilon_step);
            d.epsilon_rel_max = o.get("epsilon_rel_max", d.epsilon_rel_max);
            d.max_nsteps = static_cast<int>(o.get_int("max_nsteps", d.max_nsteps));
            d.max_substeps = static_cast<int>(o.get_int("max_substeps", d.max_substeps));
            CELER_VALIDATE(d.minimum_step > 0 && d.delta_chord > 0
                               && d.delta_intersection > 0 && d.max_substeps > 0
                               && d.max_nsteps > 0,
                           << "invalid field driver options");
        }
        if (f.has("looping"))
        {
            JsonValue const& l = f.at("looping");
            LoopingThreshold& lt = fp.looping;
            lt.max_subthreshold_steps = static_cast<std::uint32_t>(
                l.get_int("max_subthreshold_steps", lt.max_subthreshold_steps));
            lt.max_steps = static_cast<std::uint32_t>(l.get_int("max_steps", lt.max_steps));
            lt.threshold_energy = l.get("threshold_energy", lt.threshold_energy);
        }
    }

    // Volume -> material map: explicit, else from the ORANGE document
    if (doc.has("volume_to_material"))
    {
        for (auto v : doc.at("volume_to_material").as_ints())
            inp.volume_to_material.push_back(
                v < 0 ? invalid_id : static_cast<std::uint32_t>(v));
    }
    else if (geo_doc.has("materials"))
    {
        for (auto v : geo_doc.at("materials").at("cell_to_mat").as_ints())
            inp.volume_to_material.push_back(
                v < 0 ? invalid_id : static_cast<std::uint32_t>(v));
    }
    if (doc.has("volume_to_detector"))
    {
        for (auto v : doc.at("volume_to_detector").as_ints())
            inp.volume_to_detector.push_back(
                v < 0 ? invalid_id : static_cast<std::uint32_t>(v));
    }

    inp.rng = std::make_shared<XorwowParams>(
        static_cast<std::uint32_t>(doc.get_int("seed", 12345)));
    if (doc.has("track_init"))
    {
        auto const& ti = doc.at("track_init");
        inp.init_capacity = static_cast<size_type>(ti.get_int("capacity", 4096));
        inp.max_events = static_cast<size_type>(ti.get_int("max_events", 4096));
        std::string order = ti.get_str("track_order", "none");
        if (order == "none")
            inp.track_order = TrackOrder::none;
        else if (order == "init_charge")
            inp.track_order = TrackOrder::init_charge;
        else if (order == "reindex_status")
            inp.track_order = TrackOrder::reindex_status;
        else if (order == "reindex_particle_type")
            inp.track_order = TrackOrder::reindex_particle_type;
        else if (order == "reindex_step_limit_action")
            inp.track_order = TrackOrder::reindex_step_limit_action;
        else if (order == "reindex_along_step_action")
            inp.track_order = TrackOrder::reindex_along_step_action;
        else if (order == "reindex_both_action")
            inp.track_order = TrackOrder::reindex_both_action;
        else if (order == "reindex_shuffle")
            inp.track_order = TrackOrder::reindex_shuffle;
        else
            CELER_VALIDATE(false, << "unsupported track order '" << order << "'");
    }
    std::string prec = doc.get_str("precision", "fp64");
    CELER_VALIDATE(prec == "fp64" || prec == "mixed",
                   << "precision must be 'fp64' or 'mixed'");
    inp.precision = prec == "fp64" ? PrecisionMode::fp64 : PrecisionMode::mixed;
    inp.fuse_interactions = doc.get_bool("fuse_interactions", true);
    inp.concurrent_actions = doc.get_bool("concurrent_actions", true);
    inp.fuse_discrete_select = doc.get_bool("fuse_discrete_select", true);
    return std::make_shared<CoreParams>(std::move(inp));
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
