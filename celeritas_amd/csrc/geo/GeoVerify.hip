//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file geo/GeoVerify.hip
//! Device entry that locates points (the track initialization's volume
//! search, orange/univ/SimpleUnitTracker.hh:161-190 through the BIH), for
//! parity with the CPU restatement's exhaustive search.
//---------------------------------------------------------------------------//
#include <hip/hip_runtime.h>

#include "global/CoreData.hh"
#include "OrangeLiteTracker.hh"

namespace celeritas_amd
{
namespace
{
//! One point (pos[3], dir[3]) per thread; writes the volume or invalid_id
__global__ void __launch_bounds__(256)
    verify_locate_kernel(GeoParamsRef const geo,
                         double const* __restrict__ points,
                         std::uint32_t const n,
                         std::uint32_t* __restrict__ out)
{
    std::uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    geo::GeoTrack g;
    geo::initialize(geo, g, points + 6 * std::size_t(t), points + 6 * std::size_t(t) + 3);
    out[t] = g.failed ? invalid_id : g.volume;
}
}  // namespace

//---------------------------------------------------------------------------//
/*!
 * Locate n points (6 doubles each: pos [cm], dir); writes the volume ID of
 * each, or 0xffffffff where initialization fails (on a surface, outside
 * every volume with no background).
 */
extern "C" int celer_verify_geo_locate(const void* params_ref,
                                       const double* points,
                                       uint32_t n,
                                       uint32_t* out,
                                       void* stream)
{
    auto const& p = *static_cast<CoreParamsRef const*>(params_ref);
    auto st = static_cast<hipStream_t>(stream);
    if (n == 0)
        return 0;
    double* d_p = nullptr;
    std::uint32_t* d_o = nullptr;
    hipError_t err = hipSuccess;
    auto check = [&](hipError_t e) {
        if (err == hipSuccess)
            err = e;
    };
    check(hipMalloc(&d_p, sizeof(double) * 6 * std::size_t(n)));
    check(hipMalloc(&d_o, sizeof(std::uint32_t) * n));
    if (err == hipSuccess)
    {
        check(hipMemcpyAsync(d_p, points, sizeof(double) * 6 * std::size_t(n),
                             hipMemcpyHostToDevice, st));
        verify_locate_kernel<<<(n + 255) / 256, 256, 0, st>>>(p.geo, d_p, n, d_o);
        check(hipGetLastError());
        check(hipMemcpyAsync(out, d_o, sizeof(std::uint32_t) * n, hipMemcpyDeviceToHost, st));
        check(hipStreamSynchronize(st));
    }
    (void)hipFree(d_p);
    (void)hipFree(d_o);
    return static_cast<int>(err);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
