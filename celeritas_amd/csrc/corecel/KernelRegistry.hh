//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/KernelRegistry.hh
//! Attributes of the kernels launched by the library.
//!
//! The counterpart of corecel/sys/KernelAttributes.hh:33-121 and
//! KernelRegistry: each launcher records the kernel it is about to launch
//! with its workgroup size and dynamic LDS, and the registry keeps what the
//! runtime reports for it (registers, scratch, static LDS) together with the
//! occupancy at that launch shape, so that register pressure and
//! occupancy can be read from a running program instead of from profiler
//! output.
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
//! What the runtime reports for one kernel at one launch shape
struct KernelAttributes
{
    std::string name;  //!< label given by the launcher
    unsigned int block_size{0};  //!< threads per workgroup
    unsigned int dynamic_lds{0};  //!< dynamic LDS per workgroup [B]
    int num_regs{0};  //!< vector registers per lane
    std::size_t local_mem{0};  //!< scratch per lane [B]
    std::size_t static_lds{0};  //!< static LDS per workgroup [B]
    std::size_t const_mem{0};  //!< constant memory [B] (-1: not reported)
    int max_threads_per_block{0};  //!< largest workgroup the kernel allows
    int max_blocks_per_cu{0};  //!< resident workgroups per CU at this shape
    double waves_per_simd{0};  //!< resident waves per SIMD at this shape
    double occupancy{0};  //!< resident waves / the CU's maximum
};

//! Device time of the timed launches of one kernel at one shape
struct KernelTime
{
    std::string name;
    unsigned int block_size{0};
    unsigned int dynamic_lds{0};
    unsigned long long launches{0};
    double seconds{0};  //!< sum of start-to-stop times of the launches
};

//---------------------------------------------------------------------------//
/*!
 * Process-wide record of launched kernels, one entry per (kernel, shape).
 *
 * With timing on, each launch of a recorded kernel carries a start and a
 * stop event that the runtime sets when the dispatch itself begins and
 * ends (hipExtLaunchKernel), so the times hold no gaps between launches
 * and no work of other packets; launches inside a stream capture are not
 * timed.
 */
class KernelRegistry
{
  public:
    static KernelRegistry& instance();
    ~KernelRegistry();

    //! Record a kernel about to be launched (first time per shape only);
    //! returns its entry index
    std::size_t record(char const* name, void const* func, unsigned int block_size,
                       unsigned int dynamic_lds);

    //! Copy of all entries in order of first launch
    std::vector<KernelAttributes> entries() const;

    //! Entries as a JSON array
    std::string to_json() const;

    //! Time the launches that follow (on) or stop timing (off)
    void timing(bool on);
    bool timing() const { return timing_; }
    //! Events for one timed launch of an entry on a stream (null pair when
    //! timing is off or the stream is capturing)
    std::pair<hipEvent_t, hipEvent_t> timed_launch(std::size_t entry, hipStream_t stream);
    //! Accumulated times per kernel and shape (waits for pending launches)
    std::vector<KernelTime> times(bool reset);
    //! Times as a JSON array
    std::string times_json(bool reset);

  private:
    struct Key
    {
        void const* func;
        unsigned int block_size;
        unsigned int dynamic_lds;
    };
    struct Pending
    {
        std::size_t entry;
        hipEvent_t start;
        hipEvent_t stop;
    };
    mutable std::mutex mutex_;
    std::vector<Key> keys_;
    std::vector<KernelAttributes> entries_;
    bool timing_{false};
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> free_events_;
    std::vector<unsigned long long> launches_;
    std::vector<double> seconds_;

    void drain_locked();
};

//! Record a kernel function pointer with its launch shape
#define CELER_RECORD_KERNEL(NAME, KERNEL, BLOCK, LDS)   \
    ::celeritas_amd::KernelRegistry::instance().record( \
        NAME, reinterpret_cast<void const*>(KERNEL), BLOCK, LDS)

//! Largest number of timed launches held before their events are read
inline constexpr std::size_t max_pending_timed_launches = 4096;

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
