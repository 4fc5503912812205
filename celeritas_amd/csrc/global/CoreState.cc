//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/CoreState.cc
//---------------------------------------------------------------------------//
#include "CoreState.hh"

#include <algorithm>
#include <functional>
#include <numeric>
#include <random>

#include "random/XorwowParams.hh"
#include "track/SortKernels.hh"
#include "track/TrackKernels.hh"
#include "CoreParams.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
CoreState::CoreState(CoreParams const& params,
                     std::uint32_t stream_id,
                     size_type num_slots)
{
    require_device();
    CELER_VALIDATE(num_slots > 0, << "number of track slots must be positive");
    CELER_VALIDATE(num_slots <= max_track_slots,
                   << "number of track slots " << num_slots << " exceeds the maximum "
                   << max_track_slots);
    ref_.size = num_slots;
    ref_.stream_id = stream_id;
    num_detectors_ = params.num_detectors();

    std::size_t const n = num_slots;
    auto& a = arena_;
    // Plan all slices: (offset, assigner) pairs applied after allocation
    std::vector<std::pair<std::size_t, std::function<void(std::size_t)>>> plan;
    auto add = [&](auto** ptr, std::size_t count) {
        using T = std::remove_pointer_t<std::remove_reference_t<decltype(*ptr)>>;
        std::size_t off = a.reserve(count * sizeof(T));
        plan.emplace_back(off, [this, ptr](std::size_t o) {
            *ptr = arena_.get<T>(o);
        });
    };
    auto& s = ref_;
    add(&s.sim.track_id, n);
    add(&s.sim.parent_id, n);
    add(&s.sim.event_id, n);
    add(&s.sim.num_steps, n);
    add(&s.sim.time, n);
    add(&s.sim.status, n);
    add(&s.sim.step_length, n);
    add(&s.sim.post_step_action, n);
    add(&s.sim.along_step_action, n);
    add(&s.sim.num_looping_steps, n);
    add(&s.particle.particle_id, n);
    add(&s.particle.energy, n);
    for (int k = 0; k < 3; ++k)
    {
        add(&s.geo.pos[k], n);
        add(&s.geo.dir[k], n);
    }
    add(&s.geo.volume, n);
    add(&s.geo.surface, n);
    add(&s.geo.sense, n);
    add(&s.geo.boundary, n);
    add(&s.geo.failed, n);
    add(&s.geo.next_step, n);
    add(&s.geo.next_surface, n);
    add(&s.geo.next_sense, n);
    add(&s.phys.interaction_mfp, n);
    add(&s.phys.macro_xs, n);
    add(&s.phys.per_process_xs, n * max_processes);
    add(&s.phys.dedx_range, n);
    add(&s.phys.energy_deposition, n);
    add(&s.phys.sec_start, n);
    add(&s.phys.sec_count, n);
    add(&s.phys.element, n);
    add(&s.phys.msc_range_init, n);
    add(&s.phys.msc_range_factor, n);
    add(&s.phys.msc_limit_min, n);
    add(&s.phys.msc_true_path, n);
    add(&s.phys.msc_geom_path, n);
    add(&s.phys.msc_alpha, n);
    add(&s.phys.msc_is_displaced, n);
    add(&s.phys.msc_enabled, n);
    for (int k = 0; k < 5; ++k)
        add(&s.rng.xs[k], n);
    add(&s.rng.weyl, n);
    s.rng.size = num_slots;
    add(&s.init.vacancies, n);
    add(&s.init.secondary_counts, n + 1);
    add(&s.init.indices, n);
    s.init.capacity = params.init_capacity();
    s.init.max_events = params.max_events();
    add(&s.init.track_counters, s.init.max_events);
    add(&s.init.event_live, s.init.max_events);
    add(&s.init.event_edep, s.init.max_events);
    add(&s.init.completed, s.init.max_events);
    add(&s.init.initializers, s.init.capacity);
    add(&s.init.block_scratch, extend_block_scratch_size(num_slots) + 64);
    double const stack_factor
        = params.physics().device_ref().scalars.secondary_stack_factor;
    s.secondaries.capacity = static_cast<size_type>(
        std::max(1.0, stack_factor) * static_cast<double>(num_slots));
    add(&s.secondaries.storage, s.secondaries.capacity);
    add(&s.counters, 1);
    add(&s.pre_volume, n);
    add(&s.track_slots, n);
    add(&s.action_offsets, max_sort_keys + 1);
    add(&s.sort_counts, sort_counts_size(num_slots));
    add(&s.calo_edep, std::max<size_type>(num_detectors_, 1));
    add(&s.calo_partial,
        num_detectors_ <= max_lds_detectors
            ? calo_tally_blocks * std::max<size_type>(num_detectors_, 1)
            : 1);
    a.allocate();
    for (auto& p : plan)
        p.second(p.first);

    CELER_HIP_CALL(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    CELER_HIP_CALL(hipHostMalloc(reinterpret_cast<void**>(&host_counters_),
                                 sizeof(CoreCounters),
                                 hipHostMallocDefault));
    *host_counters_ = CoreCounters{};

    if (params.track_order() == TrackOrder::reindex_shuffle)
    {
        // Same permutation as the reference's host shuffle
        // (TrackSlotUtils.cc:21-31): mt19937 seeded with the slot count
        shuffled_slots_.resize(n);
        std::iota(shuffled_slots_.begin(), shuffled_slots_.end(), 0u);
        std::mt19937 g{static_cast<unsigned int>(n)};
        std::shuffle(shuffled_slots_.begin(), shuffled_slots_.end(), g);
    }

    // RNG: Mersenne-twister-seeded states, transposed to SoA
    initial_rng_.resize(6 * n);
    {
        std::vector<std::uint32_t> aos(6 * n);
        initialize_xorwow_states(params.rng().seed(), stream_id, num_slots, aos.data());
        for (std::size_t i = 0; i < n; ++i)
            for (int k = 0; k < 6; ++k)
                initial_rng_[k * n + i] = aos[6 * i + k];
    }
    this->reset();
}

//---------------------------------------------------------------------------//
CoreState::~CoreState()
{
    if (stream_)
    {
        (void)hipStreamSynchronize(stream_);
        (void)hipStreamDestroy(stream_);
    }
    if (host_counters_)
        (void)hipHostFree(host_counters_);
}

//---------------------------------------------------------------------------//
void CoreState::reset()
{
    auto const& s = ref_;
    std::size_t const n = s.size;
    CELER_HIP_CALL(hipMemsetAsync(arena_.data(), 0, arena_.size(), stream_));
    for (int k = 0; k < 5; ++k)
        CELER_HIP_CALL(hipMemcpyAsync(s.rng.xs[k], initial_rng_.data() + k * n,
                                      n * 4, hipMemcpyHostToDevice, stream_));
    CELER_HIP_CALL(hipMemcpyAsync(s.rng.weyl, initial_rng_.data() + 5 * n,
                                  n * 4, hipMemcpyHostToDevice, stream_));
    CELER_HIP_CALL(launch_iota_u32(s.init.vacancies, s.size, stream_));
    if (shuffled_slots_.empty())
    {
        CELER_HIP_CALL(launch_iota_u32(s.track_slots, s.size, stream_));
    }
    else
    {
        CELER_HIP_CALL(hipMemcpyAsync(s.track_slots, shuffled_slots_.data(), n * 4,
                                      hipMemcpyHostToDevice, stream_));
    }
    CELER_HIP_CALL(launch_fill_u32(s.geo.surface, s.size, invalid_id, stream_));
    CELER_HIP_CALL(launch_fill_u32(s.geo.next_surface, s.size, invalid_id, stream_));
    CELER_HIP_CALL(launch_fill_u32(s.sim.track_id, s.size, invalid_id, stream_));
    CELER_HIP_CALL(launch_fill_u32(s.sim.parent_id, s.size, invalid_id, stream_));
    CELER_HIP_CALL(launch_fill_u32(s.sim.event_id, s.size, invalid_id, stream_));
    CoreCounters c{};
    c.num_vacancies = s.size;
    *host_counters_ = c;
    CELER_HIP_CALL(hipMemcpyAsync(s.counters, host_counters_, sizeof(CoreCounters),
                                  hipMemcpyHostToDevice, stream_));
    CELER_HIP_CALL(hipStreamSynchronize(stream_));
}

//---------------------------------------------------------------------------//
void CoreState::enqueue_counter_readback()
{
    CELER_HIP_CALL(hipMemcpyAsync(host_counters_, ref_.counters, sizeof(CoreCounters),
                                  hipMemcpyDeviceToHost, stream_));
}

CoreCounters const& CoreState::sync_counters()
{
    this->enqueue_counter_readback();
    CELER_HIP_CALL(hipStreamSynchronize(stream_));
    return *host_counters_;
}

//---------------------------------------------------------------------------//
Primary* CoreState::primary_buffer(size_type count)
{
    if (primaries_.size() < count)
        primaries_.resize(std::max<std::size_t>(count, 2 * primaries_.size()));
    return primaries_.data();
}

std::uint32_t* CoreState::primary_rank_buffer(size_type count)
{
    if (primary_ranks_.size() < count)
        primary_ranks_.resize(std::max<std::size_t>(count, 2 * primary_ranks_.size()));
    return primary_ranks_.data();
}

//---------------------------------------------------------------------------//
void CoreState::enable_diagnostics(bool enable)
{
    CELER_HIP_CALL(hipStreamSynchronize(stream_));
    if (enable)
    {
        if (diag_.size() == 0)
            diag_.resize(num_diag_counters * diag_spread);
        CELER_HIP_CALL(hipMemsetAsync(diag_.data(), 0,
                                      sizeof(unsigned long long) * diag_.size(), stream_));
        ref_.diag = diag_.data();
    }
    else
    {
        ref_.diag = nullptr;
    }
}

std::vector<std::uint64_t> CoreState::read_diagnostics(bool reset)
{
    std::vector<std::uint64_t> result;
    if (!ref_.diag)
        return result;
    std::vector<unsigned long long> tmp(num_diag_counters * diag_spread);
    CELER_HIP_CALL(hipMemcpyAsync(tmp.data(), ref_.diag, sizeof(unsigned long long) * tmp.size(),
                                  hipMemcpyDeviceToHost, stream_));
    CELER_HIP_CALL(hipStreamSynchronize(stream_));
    if (reset)
        CELER_HIP_CALL(hipMemsetAsync(ref_.diag, 0, sizeof(unsigned long long) * tmp.size(),
                                      stream_));
    result.assign(num_diag_counters, 0);
    for (unsigned int c = 0; c < num_diag_counters; ++c)
    {
        for (unsigned int i = 0; i < diag_spread; ++i)
            result[c] += tmp[c * diag_spread + i];
    }
    return result;
}

//---------------------------------------------------------------------------//
void CoreState::block_size(size_type bs)
{
    CELER_VALIDATE(bs == 64 || bs == 128 || bs == 256,
                   << "step kernel workgroup size must be 64, 128 or 256 (got " << bs << ")");
    ref_.block_size = bs;
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
