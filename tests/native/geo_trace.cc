// Ray traces through the library's ORANGE tracker (geo/OrangeLiteTracker.hh)
// on the CPU: the same functions the device's verify_trace_kernel calls
// (initialize, find_next_step, find_safety at the segment midpoint,
// move_to_boundary, cross_boundary), on params built by the library's host
// code with host memory standing in for device memory.
//
// Usage: geo_trace problem.json rays.bin max_segments > out.bin
//   rays.bin: n x 6 doubles (pos, dir); out.bin per ray: u32 count, then
//   max_segments u32 volumes, doubles distances, doubles safeties
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "geo/OrangeLiteTracker.hh"
#include "global/CoreParams.hh"

using namespace celeritas_amd;

int main(int argc, char** argv)
{
    if (argc != 4)
    {
        std::fprintf(stderr, "usage: %s problem.json rays.bin max_segments\n", argv[0]);
        return 2;
    }
    std::ifstream f(argv[1]);
    std::stringstream ss;
    ss << f.rdbuf();
    auto params = CoreParams::from_json(ss.str());
    GeoParamsRef const& geo = params->device_ref().geo;

    std::ifstream rf(argv[2], std::ios::binary);
    std::vector<char> raw((std::istreambuf_iterator<char>(rf)), std::istreambuf_iterator<char>());
    std::size_t const n = raw.size() / (6 * sizeof(double));
    std::vector<double> rays(6 * n);
    std::memcpy(rays.data(), raw.data(), rays.size() * sizeof(double));
    std::uint32_t const m = static_cast<std::uint32_t>(std::stoul(argv[3]));

    std::vector<std::uint32_t> vols(m);
    std::vector<double> dists(m), safes(m);
    for (std::size_t t = 0; t < n; ++t)
    {
        std::fill(vols.begin(), vols.end(), 0u);
        std::fill(dists.begin(), dists.end(), 0.0);
        std::fill(safes.begin(), safes.end(), 0.0);
        geo::GeoTrack g;
        geo::initialize(geo, g, &rays[6 * t], &rays[6 * t + 3]);
        std::uint32_t k = 0;
        while (k < m && !geo::is_outside(geo, g) && !g.failed)
        {
            geo::Propagation const prop = geo::find_next_step(geo, g, HUGE_VAL);
            vols[k] = g.volume;
            dists[k] = prop.distance;
            geo::GeoTrack mid = g;
            geo::move_internal(mid, prop.distance / 2);
            safes[k] = geo::find_safety(geo, mid);
            ++k;
            if (!prop.boundary)
                break;
            geo::move_to_boundary(g);
            geo::cross_boundary(geo, g);
        }
        std::fwrite(&k, sizeof(k), 1, stdout);
        std::fwrite(vols.data(), sizeof(std::uint32_t), m, stdout);
        std::fwrite(dists.data(), sizeof(double), m, stdout);
        std::fwrite(safes.data(), sizeof(double), m, stdout);
    }
    return 0;
}
