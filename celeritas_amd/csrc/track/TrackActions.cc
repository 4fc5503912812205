//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file track/TrackActions.cc
//---------------------------------------------------------------------------//
#include "TrackActions.hh"

#include <algorithm>
#include <unordered_map>

#include "global/CoreParams.hh"
#include "global/CoreState.hh"
#include "SortKernels.hh"
#include "TrackKernels.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
ExtendFromPrimariesAction::ExtendFromPrimariesAction(ActionId id)
    : ConcreteAction(id, label_str(), "create track initializers from primaries")
{
}

void ExtendFromPrimariesAction::insert(CoreParams const& params,
                                       CoreState& state,
                                       Primary const* host_primaries,
                                       size_type count) const
{
    // Event-local ranks make track IDs reproduce the serial host order
    std::vector<std::uint32_t> ranks(2 * std::size_t(count), 0);
    {
        std::unordered_map<std::uint32_t, std::uint32_t> seen;
        std::unordered_map<std::uint32_t, size_type> last;
        for (size_type i = 0; i < count; ++i)
        {
            ranks[i] = seen[host_primaries[i].event_id]++;
            last[host_primaries[i].event_id] = i;
        }
        for (auto const& kv : last)
            ranks[count + kv.second] = 1;
    }
    Primary* buf = state.primary_buffer(count);
    std::uint32_t* dranks = state.primary_rank_buffer(2 * count);
    CELER_HIP_CALL(hipMemcpyAsync(buf,
                                  host_primaries,
                                  count * sizeof(Primary),
                                  hipMemcpyHostToDevice,
                                  state.stream()));
    CELER_HIP_CALL(hipMemcpyAsync(dranks,
                                  ranks.data(),
                                  ranks.size() * sizeof(std::uint32_t),
                                  hipMemcpyHostToDevice,
                                  state.stream()));
    CELER_HIP_CALL(hipStreamSynchronize(state.stream()));
    this->insert_device(params, state, buf, count);
    pending_.back().ranks = dranks;
}

void ExtendFromPrimariesAction::insert_device(CoreParams const& params,
                                              CoreState& state,
                                              Primary const* device_primaries,
                                              size_type count) const
{
    for (auto const& p : pending_)
    {
        CELER_VALIDATE(p.state != &state,
                       << "multiple consecutive primary insertions are not "
                          "implemented");
    }
    size_type const num_init = state.counters().num_initializers;
    CELER_VALIDATE(count + num_init <= params.init_capacity(),
                   << "insufficient initializer capacity ("
                   << params.init_capacity() << ") with size (" << num_init
                   << ") for primaries (" << count << ")");
    pending_.push_back({&state, device_primaries, nullptr, count});
}

void ExtendFromPrimariesAction::step(CoreParams const&, CoreState& state) const
{
    Primary const* data = nullptr;
    std::uint32_t const* ranks = nullptr;
    size_type count = 0;
    auto iter = std::find_if(pending_.begin(), pending_.end(), [&](Pending const& p) {
        return p.state == &state;
    });
    if (iter != pending_.end())
    {
        data = iter->data;
        ranks = iter->ranks;
        count = iter->count;
        pending_.erase(iter);
    }
    CELER_HIP_CALL(launch_extend_from_primaries(state.ref(), data, ranks, count, state.stream()));
}

//---------------------------------------------------------------------------//
void InitializeTracksAction::step(CoreParams const& params, CoreState& state) const
{
    CELER_HIP_CALL(
        launch_initialize_tracks(params.device_ref(), state.ref(), state.stream()));
}

//---------------------------------------------------------------------------//
void ExtendFromSecondariesAction::step(CoreParams const& params,
                                       CoreState& state) const
{
    CELER_HIP_CALL(launch_extend_from_secondaries(
        params.device_ref(), state.ref(), state.stream()));
}

//---------------------------------------------------------------------------//
//---------------------------------------------------------------------------//
namespace
{
char const* sort_label(TrackOrder order)
{
    switch (order)
    {
        case TrackOrder::reindex_status:
            return "sort-tracks-status";
        case TrackOrder::reindex_along_step_action:
            return "sort-tracks-along-step";
        case TrackOrder::reindex_step_limit_action:
            return "sort-tracks-post-step";
        case TrackOrder::reindex_particle_type:
            return "sort-tracks-start";
        default:
            CELER_VALIDATE(false, << "track order does not sort tracks");
    }
    return nullptr;
}
}  // namespace

SortTracksAction::SortTracksAction(ActionId id, TrackOrder order, size_type num_keys)
    : ConcreteAction(id, sort_label(order), "sort tracks states")
    , track_order_(order)
    , num_keys_(num_keys)
{
    CELER_VALIDATE(num_keys_ >= 1 && num_keys_ <= max_sort_keys,
                   << "cannot sort tracks by " << num_keys_ << " keys (at most "
                   << max_sort_keys << ")");
    switch (order)
    {
        case TrackOrder::reindex_status:
            action_order_ = StepActionOrder::sort_start;
            key_ = SortKey::status;
            break;
        case TrackOrder::reindex_particle_type:
            action_order_ = StepActionOrder::sort_start;
            key_ = SortKey::particle;
            break;
        case TrackOrder::reindex_along_step_action:
            // Sorted before the pre-step by the action it will choose (the
            // permutation the reference's sort after it gives), so that the
            // pre-step too runs over waves of one along-step action
            action_order_ = StepActionOrder::sort_start;
            key_ = SortKey::along_action_of_particle;
            break;
        case TrackOrder::reindex_step_limit_action:
            action_order_ = StepActionOrder::sort_pre_post;
            key_ = SortKey::post_action;
            break;
        default:
            CELER_VALIDATE(false, << "track order does not sort tracks");
    }
}

void SortTracksAction::step(CoreParams const& params, CoreState& state) const
{
    CELER_HIP_CALL(launch_sort_tracks(state.ref(), key_, num_keys_, state.stream(),
                                      params.device_ref().particle_along_action));
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
