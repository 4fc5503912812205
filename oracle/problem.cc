//----------------------------------*-C++-*----------------------------------//
// ORACLE — CPU restatement used only as a checker (see problem.hh).
//---------------------------------------------------------------------------//
#include "problem.hh"

#include <stdexcept>

#include "physics.hh"

namespace orc
{
//---------------------------------------------------------------------------//
std::shared_ptr<Problem> Problem::from_json(std::string const& text)
{
    Json doc = parse(text);
    auto p = std::make_shared<Problem>();
    Json const& geo = doc("geometry");
    p->geo = Geometry::from_orange_json(geo);
    // Volume simple-safety flag: all faces are planes/quadrics we support
    for (auto& v : p->geo.volumes)
        v.simple_safety = !(v.flags & 2);

    if (auto v2m = doc.find("volume_to_material"))
    {
        for (double v : v2m->nums())
            p->volume_to_material.push_back(v < 0 ? none : static_cast<std::uint32_t>(v));
    }
    else if (auto mats = geo.find("materials"))
    {
        for (double v : (*mats)("cell_to_mat").nums())
            p->volume_to_material.push_back(v < 0 ? none : static_cast<std::uint32_t>(v));
    }
    p->volume_to_material.resize(p->geo.volumes.size(), none);
    if (auto v2d = doc.find("volume_to_detector"))
    {
        for (double v : v2d->nums())
        {
            auto d = v < 0 ? none : static_cast<std::uint32_t>(v);
            p->volume_to_detector.push_back(d);
            if (d != none)
                p->num_detectors = std::max(p->num_detectors, d + 1);
        }
    }
    p->volume_to_detector.resize(p->geo.volumes.size(), none);
    p->seed = static_cast<std::uint32_t>(doc.num("seed", 12345));
    if (auto ti = doc.find("track_init"))
    {
        p->init_capacity = static_cast<std::uint32_t>(ti->num("capacity", 4096));
        p->max_events = static_cast<std::uint32_t>(ti->num("max_events", 4096));
        p->in_place = ti->str("track_order", "none") != "init_charge";
    }

    // Action IDs in the reference's registration order: physics actions
    // first (PhysicsParams.cc:97-143), then the charged along-step action,
    // then CoreParams' own (CoreParams.cc:139-215)
    ActionIds& a = p->actions;
    auto add = [&a](std::string const& label) {
        a.labels.push_back(label);
        return static_cast<std::uint32_t>(a.labels.size() - 1);
    };
    Json empty;
    empty.kind = Json::obj_k;
    Json const& phys = doc.find("physics") ? doc("physics") : empty;
    auto pd = std::make_shared<PhysicsData>();
    build_physics(phys, *pd);
    if (!pd->processes.empty())
    {
        a.pre_step = add("pre-step");
        a.msc_range = add("msc-range");
        a.eloss_range = add("eloss-range");
        a.discrete_select = add("physics-discrete-select");
        a.integral_rejected = add("physics-integral-rejected");
        for (auto const& m : pd->model_labels)
            a.model.push_back(add(m));
        a.failure = add("physics-failure");
    }
    p->secondary_stack_factor = pd->secondary_stack_factor;
    std::string along = doc.str("along_step", "");
    if (along == "general_linear")
        a.along_user = add("along-step-general-linear");
    else if (along == "uniform_field")
    {
        a.along_user = add("along-step-uniform-msc");
        Json const* f = doc.find("field");
        if (!f)
            throw std::runtime_error("uniform_field along-step needs 'field'");
        auto b = f->find("value")->nums();
        if (b.size() != 3)
            throw std::runtime_error("field value must have 3 components");
        p->field_enabled = true;
        for (int i = 0; i < 3; ++i)
            p->field[i] = b[i];
        if (Json const* o = f->find("options"))
        {
            FieldOptions& d = p->field_options;
            d.minimum_step = o->num("minimum_step", d.minimum_step);
            d.delta_chord = o->num("delta_chord", d.delta_chord);
            d.delta_intersection = o->num("delta_intersection", d.delta_intersection);
            d.epsilon_step = o->num("epsilon_step", d.epsilon_step);
            d.epsilon_rel_max = o->num("epsilon_rel_max", d.epsilon_rel_max);
            d.max_nsteps = static_cast<int>(o->num("max_nsteps", d.max_nsteps));
            d.max_substeps = static_cast<int>(o->num("max_substeps", d.max_substeps));
        }
        if (Json const* l = f->find("looping"))
        {
            p->looping_max_subthreshold_steps = static_cast<std::uint32_t>(
                l->num("max_subthreshold_steps", p->looping_max_subthreshold_steps));
            p->looping_max_steps
                = static_cast<std::uint32_t>(l->num("max_steps", p->looping_max_steps));
            p->looping_threshold_energy
                = l->num("threshold_energy", p->looping_threshold_energy);
        }
    }
    else if (along == "neutral")
        a.along_user = a.along_neutral = add("along-step-neutral");
    a.initialize_tracks = add("initialize-tracks");
    if (a.along_neutral == none)
        a.along_neutral = add("along-step-neutral");
    if (a.along_user == none)
        a.along_user = a.along_neutral;
    a.propagation_limit = add("geo-propagation-limit");
    a.boundary = add("geo-boundary");
    a.tracking_cut = add("tracking-cut");
    a.extend_secondaries = add("extend-from-secondaries");
    a.extend_primaries = add("extend-from-primaries");
    pd->finalize_actions(a);
    p->physics = pd;
    return p;
}

//---------------------------------------------------------------------------//
}  // namespace orc
