//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file user/UserActions.cc
//---------------------------------------------------------------------------//
#include "UserActions.hh"

#include <algorithm>

#include "global/CoreParams.hh"
#include "global/CoreState.hh"
#include "UserKernels.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
// PARTICLE TALLY STORE
//---------------------------------------------------------------------------//
ParticleTallyStore::ParticleTallyStore(size_type num_particles, size_type num_bins)
    : num_particles_(num_particles), num_bins_(num_bins)
{
    CELER_VALIDATE(num_particles_ > 0 && num_bins_ > 0,
                   << "particle tally needs particles and bins");
    CELER_VALIDATE(static_cast<std::size_t>(num_particles_) * num_bins_ <= max_tally_bins,
                   << "particle tally of " << num_particles_ << " x " << num_bins_
                   << " bins exceeds " << max_tally_bins);
}

ParticleTallyRef ParticleTallyStore::state(std::uint32_t stream_id, hipStream_t stream)
{
    auto& counts = counts_[stream_id];
    std::size_t const n = static_cast<std::size_t>(num_particles_) * num_bins_;
    if (counts.size() != n)
    {
        counts.resize(n);
        CELER_HIP_CALL(hipMemsetAsync(counts.data(), 0, n * sizeof(unsigned long long),
                                      stream));
    }
    ParticleTallyRef r;
    r.num_particles = num_particles_;
    r.num_bins = num_bins_;
    r.counts = counts.data();
    return r;
}

std::vector<std::uint64_t> ParticleTallyStore::sum() const
{
    std::vector<std::uint64_t> result(static_cast<std::size_t>(num_particles_) * num_bins_,
                                      0);
    CELER_HIP_CALL(hipDeviceSynchronize());
    for (auto const& kv : counts_)
    {
        auto host = kv.second.copy_to_host();
        for (std::size_t i = 0; i < host.size() && i < result.size(); ++i)
            result[i] += host[i];
    }
    return result;
}

void ParticleTallyStore::clear()
{
    CELER_HIP_CALL(hipDeviceSynchronize());
    for (auto& kv : counts_)
    {
        if (kv.second.size() > 0)
            CELER_HIP_CALL(hipMemset(kv.second.data(), 0,
                                     kv.second.size() * sizeof(unsigned long long)));
    }
}

//---------------------------------------------------------------------------//
// ACTION DIAGNOSTIC
//---------------------------------------------------------------------------//
std::shared_ptr<ActionDiagnostic> ActionDiagnostic::make_and_insert(CoreParams const& core)
{
    ActionRegistry& reg = *core.action_reg();
    auto result = std::make_shared<ActionDiagnostic>(reg.next_id(),
                                                     core.physics().num_particles());
    reg.insert_mutable(result);
    return result;
}

ActionDiagnostic::ActionDiagnostic(ActionId id, size_type num_particles)
    : ConcreteAction(id, "action-diagnostic", "accumulate post-step action counters")
    , num_particles_(num_particles)
{
    CELER_VALIDATE(num_particles_ > 0, << "action diagnostic needs particles");
}

void ActionDiagnostic::begin_run(CoreParams const& params, CoreState& state)
{
    // Sized lazily: actions may be added after this one is constructed
    // (ActionDiagnostic.cc:67-90)
    if (!store_)
    {
        store_ = std::make_unique<ParticleTallyStore>(num_particles_,
                                                      params.action_reg()->num_actions());
    }
    (void)store_->state(state.stream_id(), state.stream());
}

void ActionDiagnostic::step(CoreParams const&, CoreState& state) const
{
    CELER_VALIDATE(store_, << "action diagnostic was not set up by begin_run");
    ParticleTallyRef const t = store_->state(state.stream_id(), state.stream());
    CELER_HIP_CALL(launch_action_tally(state.ref(), t, state.stream()));
}

std::vector<std::uint64_t> ActionDiagnostic::calc_actions() const
{
    if (!store_)
        return {};
    return store_->sum();
}

void ActionDiagnostic::clear()
{
    if (store_)
        store_->clear();
}

//---------------------------------------------------------------------------//
// STEP DIAGNOSTIC
//---------------------------------------------------------------------------//
std::shared_ptr<StepDiagnostic>
StepDiagnostic::make_and_insert(CoreParams const& core, size_type max_step_bin)
{
    ActionRegistry& reg = *core.action_reg();
    auto result = std::make_shared<StepDiagnostic>(
        reg.next_id(), core.physics().num_particles(), max_step_bin);
    reg.insert_mutable(result);
    return result;
}

StepDiagnostic::StepDiagnostic(ActionId id, size_type num_particles, size_type max_step_bin)
    : ConcreteAction(id, "step-diagnostic", "accumulate total step counters")
    , store_(num_particles, max_step_bin + 2)
{
    CELER_VALIDATE(max_step_bin > 0, << "nonpositive step diagnostic 'max' bin");
}

void StepDiagnostic::begin_run(CoreParams const&, CoreState& state)
{
    (void)store_.state(state.stream_id(), state.stream());
}

void StepDiagnostic::step(CoreParams const&, CoreState& state) const
{
    ParticleTallyRef const t = store_.state(state.stream_id(), state.stream());
    CELER_HIP_CALL(launch_step_tally(state.ref(), t, state.stream()));
}

std::vector<std::uint64_t> StepDiagnostic::calc_steps() const
{
    return store_.sum();
}

//---------------------------------------------------------------------------//
// STEP COLLECTOR
//---------------------------------------------------------------------------//
struct StepCollector::StreamData
{
    size_type size{0};
    DeviceVector<std::uint32_t> u32;  //!< all 32-bit columns
    DeviceVector<double> f64;  //!< all real columns
    DeviceVector<StepHitCounters> counters;
    StepCollectorRef ref;
};

namespace
{
//! Bump layout of several columns in one allocation
template<class T>
struct ColumnPlan
{
    std::size_t total{0};
    std::vector<std::pair<T**, std::size_t>> cols;
    void add(T** ptr, std::size_t n, bool selected)
    {
        if (!selected)
            return;
        cols.emplace_back(ptr, total);
        total += (n + 63) / 64 * 64;
    }
    void assign(T* base)
    {
        for (auto const& c : cols)
            *c.first = base + c.second;
    }
};
}  // namespace

std::shared_ptr<StepCollector> StepCollector::make_and_insert(CoreParams const& core, Input inp)
{
    ActionRegistry& reg = *core.action_reg();
    auto result = std::make_shared<StepCollector>(core, std::move(inp));
    reg.insert_mutable(std::make_shared<StepGatherAction>(reg.next_id(), step_pre, result));
    reg.insert(std::make_shared<StepGatherAction>(reg.next_id(), step_post, result));
    return result;
}

StepCollector::StepCollector(CoreParams const& core, Input inp) : inp_(std::move(inp))
{
    CELER_VALIDATE(inp_.hit_capacity > 0, << "step collector needs a hit capacity");
    if (!inp_.volume_to_detector.empty())
    {
        auto det = inp_.volume_to_detector;
        CELER_VALIDATE(det.size() <= core.geometry().num_volumes(),
                       << "detector map has " << det.size() << " entries for "
                       << core.geometry().num_volumes() << " volumes");
        det.resize(core.geometry().num_volumes(), invalid_id);
        d_detector_map_ = to_device(det);
    }
}

StepCollector::~StepCollector() = default;

void StepCollector::begin_run(CoreParams const&, CoreState& state)
{
    auto& sd = streams_[state.stream_id()];
    if (sd && sd->size == state.size())
        return;
    sd = std::make_unique<StreamData>();
    size_type const n = state.size();
    size_type const cap = inp_.hit_capacity;
    StepSelection const& sel = inp_.selection;
    StepCollectorRef& r = sd->ref;
    r.selection = sel;
    r.detector_map = d_detector_map_.size() ? d_detector_map_.data() : nullptr;
    r.nonzero_energy_deposition = inp_.nonzero_energy_deposition ? 1u : 0u;
    r.hits.capacity = cap;

    unsigned int const nblocks = (n + step_gather_block - 1) / step_gather_block;
    ColumnPlan<std::uint32_t> u;
    ColumnPlan<double> f;
    StepPointSelection const& pre = sel.points[step_pre];
    u.add(&r.slot_detector, n, true);
    u.add(&r.block_counts, nblocks, true);
    f.add(&r.pre.time, n, pre.time);
    for (int j = 0; j < 3; ++j)
    {
        f.add(&r.pre.pos[j], n, pre.pos);
        f.add(&r.pre.dir[j], n, pre.dir);
    }
    u.add(&r.pre.volume_id, n, pre.volume_id);
    f.add(&r.pre.energy, n, pre.energy);
    // Hit buffer columns
    u.add(&r.hits.detector, cap, true);
    u.add(&r.hits.track_id, cap, true);
    for (int p = 0; p < 2; ++p)
    {
        StepPointSelection const& ps = sel.points[p];
        StepPointHits& h = r.hits.points[p];
        f.add(&h.time, cap, ps.time);
        f.add(&h.pos, 3 * std::size_t(cap), ps.pos);
        f.add(&h.dir, 3 * std::size_t(cap), ps.dir);
        u.add(&h.volume_id, cap, ps.volume_id);
        f.add(&h.energy, cap, ps.energy);
    }
    u.add(&r.hits.event_id, cap, sel.event_id);
    u.add(&r.hits.parent_id, cap, sel.parent_id);
    u.add(&r.hits.track_step_count, cap, sel.track_step_count);
    u.add(&r.hits.action_id, cap, sel.action_id);
    f.add(&r.hits.step_length, cap, sel.step_length);
    u.add(&r.hits.particle, cap, sel.particle);
    f.add(&r.hits.energy_deposition, cap, sel.energy_deposition);

    sd->u32.resize(u.total);
    sd->f64.resize(f.total);
    u.assign(sd->u32.data());
    if (f.total)
        f.assign(sd->f64.data());
    sd->counters.resize(1);
    r.counters = sd->counters.data();
    CELER_HIP_CALL(launch_step_hits_clear(r.counters, state.stream()));
    sd->size = n;
}

StepCollectorRef const* StepCollector::stream_ref(std::uint32_t stream_id) const
{
    auto iter = streams_.find(stream_id);
    if (iter == streams_.end())
        return nullptr;
    return &iter->second->ref;
}

DetectorStepOutput StepCollector::read(std::uint32_t stream_id)
{
    DetectorStepOutput out;
    auto iter = streams_.find(stream_id);
    if (iter == streams_.end())
        return out;
    StepCollectorRef const& r = iter->second->ref;
    CELER_HIP_CALL(hipDeviceSynchronize());
    StepHitCounters c{};
    CELER_HIP_CALL(hipMemcpy(&c, r.counters, sizeof(c), hipMemcpyDeviceToHost));
    std::size_t const n = c.num_hits;
    auto copy = [n](auto* dst, auto const* src, std::size_t width) {
        if (!src)
            return;
        dst->resize(n * width);
        if (n)
            CELER_HIP_CALL(hipMemcpy(dst->data(), src, n * width * sizeof((*dst)[0]),
                                     hipMemcpyDeviceToHost));
    };
    copy(&out.detector, r.hits.detector, 1);
    copy(&out.track_id, r.hits.track_id, 1);
    for (int p = 0; p < 2; ++p)
    {
        StepPointHits const& h = r.hits.points[p];
        copy(&out.points[p].time, h.time, 1);
        copy(&out.points[p].pos, h.pos, 3);
        copy(&out.points[p].dir, h.dir, 3);
        copy(&out.points[p].volume_id, h.volume_id, 1);
        copy(&out.points[p].energy, h.energy, 1);
    }
    copy(&out.event_id, r.hits.event_id, 1);
    copy(&out.parent_id, r.hits.parent_id, 1);
    copy(&out.track_step_count, r.hits.track_step_count, 1);
    copy(&out.action_id, r.hits.action_id, 1);
    copy(&out.step_length, r.hits.step_length, 1);
    copy(&out.particle, r.hits.particle, 1);
    copy(&out.energy_deposition, r.hits.energy_deposition, 1);
    out.overflow_steps = c.overflow;
    out.dropped_hits = c.dropped;
    CELER_HIP_CALL(launch_step_hits_clear(r.counters, nullptr));
    CELER_HIP_CALL(hipDeviceSynchronize());
    return out;
}

//---------------------------------------------------------------------------//
StepGatherAction::StepGatherAction(ActionId id,
                                   StepPoint point,
                                   std::shared_ptr<StepCollector> owner)
    : ConcreteAction(id,
                     point == step_pre ? "step-gather-pre" : "step-gather-post",
                     point == step_pre ? "gather pre-step steps/hits"
                                       : "gather post-step steps/hits")
    , point_(point)
    , owner_(std::move(owner))
{
    CELER_VALIDATE(owner_, << "step gather action needs its collector");
}

void StepGatherAction::begin_run(CoreParams const& params, CoreState& state)
{
    owner_->begin_run(params, state);
}

void StepGatherAction::step(CoreParams const& params, CoreState& state) const
{
    StepCollectorRef const* r = owner_->stream_ref(state.stream_id());
    CELER_VALIDATE(r, << "step collector was not set up for stream " << state.stream_id());
    if (point_ == step_pre)
    {
        CELER_HIP_CALL(launch_step_gather_pre(params.device_ref(), state.ref(), *r,
                                              state.stream()));
    }
    else
    {
        CELER_HIP_CALL(launch_step_gather_post(params.device_ref(), state.ref(), *r,
                                               state.stream()));
    }
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
