//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/KernelRegistry.cc
//---------------------------------------------------------------------------//
#include "KernelRegistry.hh"

#include <hip/hip_runtime_api.h>

#include "DeviceRuntime.hh"
#include "Json.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
KernelRegistry& KernelRegistry::instance()
{
    static KernelRegistry registry;
    return registry;
}

/*!
 * Query the runtime for a kernel the first time it is launched with a given
 * workgroup size and dynamic LDS (make_kernel_attributes in
 * corecel/sys/KernelAttributes.hh:58-121).
 */
//! The events are left to the process's end: the runtime may already be
//! torn down when a static registry is destroyed
KernelRegistry::~KernelRegistry() = default;

std::size_t KernelRegistry::record(char const* name,
                                   void const* func,
                                   unsigned int block_size,
                                   unsigned int dynamic_lds)
{
    std::lock_guard<std::mutex> lock(mutex_);
    for (std::size_t i = 0; i < keys_.size(); ++i)
    {
        auto const& k = keys_[i];
        if (k.func == func && k.block_size == block_size && k.dynamic_lds == dynamic_lds)
            return i;
    }
    KernelAttributes a;
    a.name = name;
    a.block_size = block_size;
    a.dynamic_lds = dynamic_lds;

    hipFuncAttributes fa{};
    CELER_HIP_CALL(hipFuncGetAttributes(&fa, func));
    a.num_regs = fa.numRegs;
    a.local_mem = fa.localSizeBytes;
    a.static_lds = fa.sharedSizeBytes;
    a.const_mem = fa.constSizeBytes;
    a.max_threads_per_block = fa.maxThreadsPerBlock;

    int blocks = 0;
    CELER_HIP_CALL(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &blocks, func, static_cast<int>(block_size), dynamic_lds));
    int device = 0;
    CELER_HIP_CALL(hipGetDevice(&device));
    int threads_per_cu = 0;
    CELER_HIP_CALL(hipDeviceGetAttribute(
        &threads_per_cu, hipDeviceAttributeMaxThreadsPerMultiProcessor, device));
    int simds_per_cu = 4;
    a.max_blocks_per_cu = blocks;
    // Wave64 everywhere: a CU's four SIMDs share its resident waves
    a.waves_per_simd = static_cast<double>(blocks) * block_size / 64.0 / simds_per_cu;
    a.occupancy = threads_per_cu > 0
                      ? static_cast<double>(blocks) * block_size / threads_per_cu
                      : 0.0;

    keys_.push_back({func, block_size, dynamic_lds});
    entries_.push_back(std::move(a));
    launches_.push_back(0);
    seconds_.push_back(0);
    return entries_.size() - 1;
}

void KernelRegistry::timing(bool on)
{
    std::lock_guard<std::mutex> lock(mutex_);
    timing_ = on;
}

std::pair<hipEvent_t, hipEvent_t>
KernelRegistry::timed_launch(std::size_t entry, hipStream_t stream)
{
    std::lock_guard<std::mutex> lock(mutex_);
    if (!timing_)
        return {nullptr, nullptr};
    hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
    CELER_HIP_CALL(hipStreamIsCapturing(stream, &capture));
    if (capture != hipStreamCaptureStatusNone)
        return {nullptr, nullptr};
    if (pending_.size() >= max_pending_timed_launches)
        this->drain_locked();
    hipEvent_t ev[2];
    for (auto& e : ev)
    {
        if (free_events_.empty())
        {
            CELER_HIP_CALL(hipEventCreate(&e));
        }
        else
        {
            e = free_events_.back();
            free_events_.pop_back();
        }
    }
    pending_.push_back({entry, ev[0], ev[1]});
    return {ev[0], ev[1]};
}

//! Wait for the pending timed launches and add their times
void KernelRegistry::drain_locked()
{
    for (auto const& p : pending_)
    {
        CELER_HIP_CALL(hipEventSynchronize(p.stop));
        float ms = 0;
        CELER_HIP_CALL(hipEventElapsedTime(&ms, p.start, p.stop));
        seconds_[p.entry] += 1e-3 * ms;
        launches_[p.entry] += 1;
        free_events_.push_back(p.start);
        free_events_.push_back(p.stop);
    }
    pending_.clear();
}

std::vector<KernelTime> KernelRegistry::times(bool reset)
{
    std::lock_guard<std::mutex> lock(mutex_);
    this->drain_locked();
    std::vector<KernelTime> result;
    for (std::size_t i = 0; i < entries_.size(); ++i)
    {
        if (launches_[i] == 0)
            continue;
        KernelTime t;
        t.name = entries_[i].name;
        t.block_size = entries_[i].block_size;
        t.dynamic_lds = entries_[i].dynamic_lds;
        t.launches = launches_[i];
        t.seconds = seconds_[i];
        result.push_back(std::move(t));
        if (reset)
        {
            launches_[i] = 0;
            seconds_[i] = 0;
        }
    }
    return result;
}

std::string KernelRegistry::times_json(bool reset)
{
    JsonValue arr = JsonValue::array();
    for (auto const& t : this->times(reset))
    {
        JsonValue e = JsonValue::object();
        e.obj["name"] = JsonValue::string(t.name);
        e.obj["block_size"] = JsonValue::number(t.block_size);
        e.obj["dynamic_lds"] = JsonValue::number(t.dynamic_lds);
        e.obj["launches"] = JsonValue::number(static_cast<double>(t.launches));
        e.obj["seconds"] = JsonValue::number(t.seconds);
        arr.arr.push_back(std::move(e));
    }
    return dump_json(arr);
}

std::vector<KernelAttributes> KernelRegistry::entries() const
{
    std::lock_guard<std::mutex> lock(mutex_);
    return entries_;
}

std::string KernelRegistry::to_json() const
{
    JsonValue arr = JsonValue::array();
    for (auto const& a : this->entries())
    {
        JsonValue e = JsonValue::object();
        e.obj["name"] = JsonValue::string(a.name);
        e.obj["block_size"] = JsonValue::number(a.block_size);
        e.obj["dynamic_lds"] = JsonValue::number(a.dynamic_lds);
        e.obj["num_regs"] = JsonValue::number(a.num_regs);
        e.obj["local_mem"] = JsonValue::number(static_cast<double>(a.local_mem));
        e.obj["static_lds"] = JsonValue::number(static_cast<double>(a.static_lds));
        if (a.const_mem != static_cast<std::size_t>(-1))  // -1: not reported
            e.obj["const_mem"] = JsonValue::number(static_cast<double>(a.const_mem));
        e.obj["max_threads_per_block"] = JsonValue::number(a.max_threads_per_block);
        e.obj["max_blocks_per_cu"] = JsonValue::number(a.max_blocks_per_cu);
        e.obj["waves_per_simd"] = JsonValue::number(a.waves_per_simd);
        e.obj["occupancy"] = JsonValue::number(a.occupancy);
        arr.arr.push_back(std::move(e));
    }
    return dump_json(arr);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
