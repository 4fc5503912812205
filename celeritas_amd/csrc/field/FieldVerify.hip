//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file field/FieldVerify.hip
//! Device entry that runs sequences of uniform-field propagations, for the
//! reference's FieldPropagator unit cases (test/celeritas/field/
//! FieldPropagator.test.cc) and for parity with the CPU restatement.
//---------------------------------------------------------------------------//
#include <hip/hip_runtime.h>

#include "global/CoreData.hh"
#include "FieldPropagator.hh"

namespace celeritas_amd
{
namespace
{
constexpr int record_size = 10;

/*!
 * Each thread propagates one track through its own request list through a
 * single propagator (a request of 0 crosses the boundary and starts a new
 * propagator, -1 only starts a new one).
 */
__global__ void __launch_bounds__(64)
    verify_field_kernel(CoreParamsRef const params,
                        FieldDriverOptions const opts,
                        double const* __restrict__ tracks,
                        std::uint32_t const ntracks,
                        double const* __restrict__ steps,
                        std::uint32_t const nsteps,
                        double* __restrict__ out)
{
    std::uint32_t const t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntracks)
        return;
    // Track record: pos[3], dir[3], momentum, charge, field[3]
    double const* in = tracks + 11 * t;
    geo::GeoTrack g;
    geo::initialize(params.geo, g, in, in + 3);
    double const momentum = in[6];
    field::UniformFieldStepper const stepper{{in[8], in[9], in[10]},
                                             in[7] * field::lorentz_per_cm_tesla};
    field::PropagatorState ps = field::make_propagator_state(g, momentum);
    for (std::uint32_t k = 0; k < nsteps; ++k)
    {
        double const req = steps[std::size_t(t) * nsteps + k];
        field::FieldPropagation r{0, false, false};
        if (req > 0)
        {
            r = field::propagate(params.geo, g, stepper, opts, ps, req);
        }
        else
        {
            if (req == 0)
                geo::cross_boundary(params.geo, g);
            ps = field::make_propagator_state(g, momentum);
        }
        double* o = out + (std::size_t(t) * nsteps + k) * record_size;
        o[0] = r.distance;
        o[1] = r.boundary;
        o[2] = r.looping;
        for (int i = 0; i < 3; ++i)
        {
            o[3 + i] = g.pos[i];
            o[6 + i] = g.dir[i];
        }
        o[9] = g.volume;
    }
}
}  // namespace

//---------------------------------------------------------------------------//
/*!
 * Propagate ntracks tracks (11 doubles each: pos, dir, momentum [MeV/c],
 * charge [e], field [T]) through nsteps requests each; writes 10 doubles
 * per request: distance, boundary, looping, pos, dir, volume.
 */
extern "C" int celer_verify_field_propagate(const void* params_ref,
                                            int max_substeps,
                                            const double* tracks,
                                            uint32_t ntracks,
                                            const double* steps,
                                            uint32_t nsteps,
                                            double* out,
                                            void* stream)
{
    auto const& p = *static_cast<CoreParamsRef const*>(params_ref);
    auto st = static_cast<hipStream_t>(stream);
    FieldDriverOptions opts = p.field.options;
    if (max_substeps > 0)
        opts.max_substeps = max_substeps;
    double *d_t = nullptr, *d_s = nullptr, *d_o = nullptr;
    std::size_t const nreq = std::size_t(ntracks) * nsteps;
    hipError_t err = hipSuccess;
    auto check = [&](hipError_t e) {
        if (err == hipSuccess)
            err = e;
    };
    check(hipMalloc(&d_t, sizeof(double) * 11 * ntracks));
    check(hipMalloc(&d_s, sizeof(double) * nreq));
    check(hipMalloc(&d_o, sizeof(double) * record_size * nreq));
    if (err == hipSuccess && ntracks > 0 && nsteps > 0)
    {
        check(hipMemcpyAsync(d_t, tracks, sizeof(double) * 11 * ntracks,
                             hipMemcpyHostToDevice, st));
        check(hipMemcpyAsync(d_s, steps, sizeof(double) * nreq, hipMemcpyHostToDevice, st));
        unsigned int const nb = (ntracks + 63) / 64;
        verify_field_kernel<<<nb, 64, 0, st>>>(p, opts, d_t, ntracks, d_s, nsteps, d_o);
        check(hipGetLastError());
        check(hipMemcpyAsync(out, d_o, sizeof(double) * record_size * nreq,
                             hipMemcpyDeviceToHost, st));
        check(hipStreamSynchronize(st));
    }
    (void)hipFree(d_t);
    (void)hipFree(d_s);
    (void)hipFree(d_o);
    return static_cast<int>(err);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
