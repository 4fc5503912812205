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

/*!
 * Trace one ray per thread (pos[3], dir[3]) through the geometry until it
 * leaves the world or fails: the volume, the distance to the next boundary
 * (OrangeTrackView::find_next_step with no step limit) and the safety at
 * the segment's midpoint for each segment, then move_to_boundary and
 * cross_boundary (orange/OrangeTrackView.hh:555-775).
 */
__global__ void __launch_bounds__(256)
    verify_trace_kernel(GeoParamsRef const geo,
                        double const* __restrict__ rays,
                        std::uint32_t const n,
                        std::uint32_t const max_segments,
                        std::uint32_t* __restrict__ counts,
                        std::uint32_t* __restrict__ vols,
                        double* __restrict__ dists,
                        double* __restrict__ safeties)
{
    std::uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n)
        return;
    geo::GeoTrack g;
    geo::initialize(geo, g, rays + 6 * std::size_t(t), rays + 6 * std::size_t(t) + 3);
    std::size_t const base = std::size_t(t) * max_segments;
    std::uint32_t k = 0;
    while (k < max_segments && !geo::is_outside(geo, g) && !g.failed)
    {
        geo::Propagation const prop = geo::find_next_step(geo, g, HUGE_VAL);
        vols[base + k] = g.volume;
        dists[base + k] = prop.distance;
        geo::GeoTrack mid = g;
        geo::move_internal(mid, prop.distance / 2);
        safeties[base + k] = geo::find_safety(geo, mid);
        ++k;
        if (!prop.boundary)
            break;
        geo::move_to_boundary(g);
        geo::cross_boundary(geo, g);
    }
    counts[t] = k;
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

/*!
 * Trace n rays (6 doubles each: pos [cm], dir); per ray up to max_segments
 * (volume, distance, midpoint safety) triples, ray-major, and the number
 * of segments written.
 */
extern "C" int celer_verify_geo_trace(const void* params_ref,
                                      const double* rays,
                                      uint32_t n,
                                      uint32_t max_segments,
                                      uint32_t* counts,
                                      uint32_t* volumes,
                                      double* distances,
                                      double* safeties,
                                      void* stream)
{
    auto const& p = *static_cast<CoreParamsRef const*>(params_ref);
    auto st = static_cast<hipStream_t>(stream);
    if (n == 0 || max_segments == 0)
        return 0;
    std::size_t const m = std::size_t(n) * max_segments;
    double* d_r = nullptr;
    std::uint32_t* d_c = nullptr;
    std::uint32_t* d_v = nullptr;
    double* d_d = nullptr;
    double* d_s = nullptr;
    hipError_t err = hipSuccess;
    auto check = [&](hipError_t e) {
        if (err == hipSuccess)
            err = e;
    };
    check(hipMalloc(&d_r, sizeof(double) * 6 * std::size_t(n)));
    check(hipMalloc(&d_c, sizeof(std::uint32_t) * n));
    check(hipMalloc(&d_v, sizeof(std::uint32_t) * m));
    check(hipMalloc(&d_d, sizeof(double) * m));
    check(hipMalloc(&d_s, sizeof(double) * m));
    if (err == hipSuccess)
    {
        check(hipMemcpyAsync(d_r, rays, sizeof(double) * 6 * std::size_t(n),
                             hipMemcpyHostToDevice, st));
        verify_trace_kernel<<<(n + 255) / 256, 256, 0, st>>>(p.geo, d_r, n, max_segments,
                                                             d_c, d_v, d_d, d_s);
        check(hipGetLastError());
        check(hipMemcpyAsync(counts, d_c, sizeof(std::uint32_t) * n, hipMemcpyDeviceToHost, st));
        check(hipMemcpyAsync(volumes, d_v, sizeof(std::uint32_t) * m, hipMemcpyDeviceToHost, st));
        check(hipMemcpyAsync(distances, d_d, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        check(hipMemcpyAsync(safeties, d_s, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        check(hipStreamSynchronize(st));
    }
    (void)hipFree(d_r);
    (void)hipFree(d_c);
    (void)hipFree(d_v);
    (void)hipFree(d_d);
    (void)hipFree(d_s);
    return static_cast<int>(err);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
