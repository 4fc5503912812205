//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file global/CoreState.hh
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <vector>

#include "corecel/DeviceRuntime.hh"
#include "CoreData.hh"

namespace celeritas_amd
{
class CoreParams;

//---------------------------------------------------------------------------//
/*!
 * Per-stream mutable state: every per-slot field in one HBM arena.
 *
 * Mirrors CoreState<device> (celeritas/global/CoreState.cc:28-71):
 * constructed from params, a stream ID and a number of track slots; slots
 * start inactive with all of them vacant, RNG states are filled from the
 * seeded Mersenne twister stream, and the counters live on the device with
 * a pinned host mirror updated by sync_counters().
 */
class CoreState
{
  public:
    CoreState(CoreParams const& params, std::uint32_t stream_id, size_type num_slots);
    ~CoreState();
    CoreState(CoreState const&) = delete;
    CoreState& operator=(CoreState const&) = delete;

    CoreStateRef const& ref() const { return ref_; }
    size_type size() const { return ref_.size; }
    std::uint32_t stream_id() const { return ref_.stream_id; }
    //! Stream the step kernels are launched on: the state's own, or a side
    //! stream while the stepper runs an action concurrently with others
    hipStream_t stream() const { return launch_stream_ ? launch_stream_ : stream_; }
    //! The state's own stream
    hipStream_t own_stream() const { return stream_; }
    //! Launch on a side stream (null: back to the state's own)
    void launch_stream(hipStream_t s) { launch_stream_ = s; }

    //! Copy device counters to the host mirror (synchronizes the stream)
    CoreCounters const& sync_counters();
    //! Last synchronized counters
    CoreCounters const& counters() const { return *host_counters_; }
    //! Workgroup size of the step kernels: 64, 128 or 256
    void block_size(size_type bs);
    size_type block_size() const { return ref_.block_size; }
    //! Maintain the per-event live-track and deposition tallies (set before
    //! a step is captured in a graph: launches take the state by value)
    void event_tally(bool on) { ref_.init.event_tally = on ? 1u : 0u; }
    //! Attach a serving queue the device admits events from at the start of
    //! each step (set before a step is captured in a graph)
    void serve_queue(ServeQueueRef const& q) { ref_.serve = q; }
    bool event_tally() const { return ref_.init.event_tally != 0; }
    //! Enqueue an async copy of the counters to the pinned mirror
    void enqueue_counter_readback();

    //! Reset to an empty state (all slots vacant, counters cleared)
    void reset();

    //! Scratch buffer for primaries (grown on demand)
    Primary* primary_buffer(size_type count);
    //! Scratch buffer for primary event ranks (grown on demand)
    std::uint32_t* primary_rank_buffer(size_type count);

    bool warming_up() const { return warming_up_; }
    void warming_up(bool w) { warming_up_ = w; }

    std::size_t memory_bytes() const { return arena_.size(); }
    size_type num_detectors() const { return num_detectors_; }

    //! Turn the per-kernel work counters on (zeroed) or off
    void enable_diagnostics(bool enable);
    //! Read the work counters (synchronizes the stream); empty if off
    std::vector<std::uint64_t> read_diagnostics(bool reset);

  private:
    CoreStateRef ref_;
    DeviceArena arena_;
    hipStream_t stream_{nullptr};
    hipStream_t launch_stream_{nullptr};
    CoreCounters* host_counters_{nullptr};
    DeviceVector<Primary> primaries_;
    DeviceVector<std::uint32_t> primary_ranks_;
    DeviceVector<unsigned long long> diag_;
    size_type num_detectors_{0};
    bool warming_up_{false};
    std::vector<std::uint32_t> initial_rng_;
    std::vector<std::uint32_t> shuffled_slots_;  //!< reindex_shuffle order
};

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
