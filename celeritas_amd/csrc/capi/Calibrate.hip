//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file capi/Calibrate.hip
//! Plain device copy of a known size, the reference point for converting
//! hardware memory counters (FETCH_SIZE, WRITE_SIZE) into bytes.
//---------------------------------------------------------------------------//
#include "celeritas_amd.h"

#include <hip/hip_runtime.h>

namespace
{
constexpr int block_size = 256;

__global__ void __launch_bounds__(block_size)
    calibration_copy_kernel(double2 const* __restrict__ in,
                            double2* __restrict__ out,
                            std::size_t n)
{
    std::size_t const i = std::size_t(blockIdx.x) * block_size + threadIdx.x;
    if (i < n)
        out[i] = in[i];
}
}  // namespace

/*!
 * Copy a device buffer of the given size (rounded down to 16 B) reps times
 * with a one-element-per-thread kernel; returns the mean kernel time in
 * milliseconds through *ms, or a negative hipError_t value.
 */
extern "C" int celer_calibration_copy(size_t bytes, uint32_t reps, float* ms)
{
    std::size_t const n = bytes / sizeof(double2);
    if (n == 0 || reps == 0)
        return -static_cast<int>(hipErrorInvalidValue);
    double2* in = nullptr;
    double2* out = nullptr;
    hipEvent_t start = nullptr, stop = nullptr;
    hipError_t err = hipMalloc(&in, n * sizeof(double2));
    if (err == hipSuccess)
        err = hipMalloc(&out, n * sizeof(double2));
    if (err == hipSuccess)
        err = hipMemset(in, 0, n * sizeof(double2));
    if (err == hipSuccess)
        err = hipEventCreate(&start);
    if (err == hipSuccess)
        err = hipEventCreate(&stop);
    if (err == hipSuccess)
    {
        unsigned int const nb = static_cast<unsigned int>((n + block_size - 1) / block_size);
        (void)hipEventRecord(start, nullptr);
        for (uint32_t r = 0; r < reps; ++r)
            calibration_copy_kernel<<<nb, block_size>>>(in, out, n);
        (void)hipEventRecord(stop, nullptr);
        err = hipGetLastError();
        if (err == hipSuccess)
            err = hipEventSynchronize(stop);
        float total = 0;
        if (err == hipSuccess)
            err = hipEventElapsedTime(&total, start, stop);
        if (ms)
            *ms = total / reps;
    }
    if (start)
        (void)hipEventDestroy(start);
    if (stop)
        (void)hipEventDestroy(stop);
    (void)hipFree(in);
    (void)hipFree(out);
    return err == hipSuccess ? 0 : -static_cast<int>(err);
}
