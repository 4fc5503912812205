//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file em/InteractorVerify.hip
//! Runs a device interactor on replayed uniform streams.
//!
//! Counterpart of the reference's per-interactor device test wrappers
//! (test/celeritas/*/*.test.cu): one thread per sample, each with its own
//! stretch of a buffer of uniform doubles, writing the outgoing state so it
//! can be compared with a CPU implementation fed the same stream.
//---------------------------------------------------------------------------//
#include "celeritas_amd.h"

#include "global/CoreParams.hh"
#include "global/CoreState.hh"
#include "Interactors.hh"

namespace celeritas_amd
{
namespace
{
constexpr int block_size = 256;
constexpr int record_size = 17;

__global__ void __launch_bounds__(block_size)
    verify_interactor_kernel(CoreParamsRef const params,
                             CoreStateRef const state,
                             std::uint32_t const model,
                             std::uint32_t const particle,
                             double const* __restrict__ energies,
                             double const* __restrict__ dirs,
                             std::uint32_t const material,
                             std::uint32_t const element,
                             std::uint32_t const n,
                             double const* __restrict__ uniforms,
                             std::uint32_t const stride,
                             double* __restrict__ out)
{
    std::uint32_t const t = blockIdx.x * block_size + threadIdx.x;
    if (t >= n)
        return;
    ReplayEngine rng{uniforms + std::size_t(t) * stride, 0};
    double dir[3] = {dirs[3 * t], dirs[3 * t + 1], dirs[3 * t + 2]};
    InteractionContext ctx{params.phys, state, 0, material, element, particle};
    Interaction r = sample_interaction(static_cast<ModelClass>(model), ctx,
                                       energies[t], dir, rng);
    double* o = out + std::size_t(record_size) * t;
    o[0] = static_cast<double>(r.action);
    o[1] = r.energy;
    o[2] = r.dir[0];
    o[3] = r.dir[1];
    o[4] = r.dir[2];
    o[5] = r.energy_deposition;
    o[6] = r.num_secondaries;
    for (std::uint32_t j = 0; j < 2; ++j)
    {
        double* so = o + 7 + 5 * j;
        bool const ok = j < r.num_secondaries && r.action != Interaction::failed;
        if (ok)
        {
            Secondary const& sec = state.secondaries.storage[r.sec_start + j];
            so[0] = sec.particle_id == invalid_id ? -1.0 : double(sec.particle_id);
            so[1] = sec.energy;
            so[2] = sec.dir[0];
            so[3] = sec.dir[1];
            so[4] = sec.dir[2];
        }
        else
        {
            so[0] = -1;
            so[1] = so[2] = so[3] = so[4] = 0;
        }
    }
}
}  // namespace
}  // namespace celeritas_amd

using namespace celeritas_amd;

//---------------------------------------------------------------------------//
/*!
 * Run one interactor model over n samples with replayed uniforms.
 *
 * The state supplies the secondary stack (its counter is reset first).
 */
extern "C" int celer_verify_interactor(const void* params_ref,
                                       const void* state_ref,
                                       uint32_t model,
                                       uint32_t particle,
                                       const double* energies,
                                       const double* dirs,
                                       uint32_t material,
                                       uint32_t element,
                                       uint32_t n,
                                       const double* uniforms,
                                       uint32_t stride,
                                       double* out,
                                       void* stream)
{
    auto const& p = *static_cast<CoreParamsRef const*>(params_ref);
    auto const& s = *static_cast<CoreStateRef const*>(state_ref);
    auto st = static_cast<hipStream_t>(stream);
    double *d_e = nullptr, *d_dir = nullptr, *d_u = nullptr, *d_out = nullptr;
    hipError_t err = hipSuccess;
    auto check = [&](hipError_t e) {
        if (err == hipSuccess)
            err = e;
    };
    check(hipMalloc(&d_e, sizeof(double) * n));
    check(hipMalloc(&d_dir, sizeof(double) * 3 * n));
    check(hipMalloc(&d_u, sizeof(double) * std::size_t(stride) * n));
    check(hipMalloc(&d_out, sizeof(double) * record_size * n));
    if (err == hipSuccess)
    {
        std::uint32_t zero = 0;
        check(hipMemcpyAsync(&s.counters->secondary_stack_size, &zero, 4,
                             hipMemcpyHostToDevice, st));
        check(hipMemcpyAsync(d_e, energies, sizeof(double) * n, hipMemcpyHostToDevice, st));
        check(hipMemcpyAsync(d_dir, dirs, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
        check(hipMemcpyAsync(d_u, uniforms, sizeof(double) * std::size_t(stride) * n,
                             hipMemcpyHostToDevice, st));
        unsigned int const nb = (n + block_size - 1) / block_size;
        verify_interactor_kernel<<<nb, block_size, 0, st>>>(
            p, s, model, particle, d_e, d_dir, material, element, n, d_u, stride, d_out);
        check(hipGetLastError());
        check(hipMemcpyAsync(out, d_out, sizeof(double) * record_size * n,
                             hipMemcpyDeviceToHost, st));
        check(hipStreamSynchronize(st));
    }
    (void)hipFree(d_e);
    (void)hipFree(d_dir);
    (void)hipFree(d_u);
    (void)hipFree(d_out);
    return static_cast<int>(err);
}
