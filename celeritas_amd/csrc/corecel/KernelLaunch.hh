//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/KernelLaunch.hh
//! Kernel launch that records the kernel's attributes and, when the
//! registry is timing, attaches start and stop events to the dispatch.
//---------------------------------------------------------------------------//
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "KernelRegistry.hh"

namespace celeritas_amd
{
//---------------------------------------------------------------------------//
/*!
 * Launch a kernel under a registry name.
 *
 * The arguments are converted to the kernel's parameter types. A timed
 * launch goes through hipExtLaunchKernelGGL, whose events the runtime sets
 * at the dispatch's own start and end.
 */
template<class... Params, class... Args>
void launch_kernel(char const* name,
                   void (*kernel)(Params...),
                   unsigned int grid,
                   unsigned int block,
                   unsigned int lds,
                   hipStream_t stream,
                   Args const&... args)
{
    static_assert(sizeof...(Params) == sizeof...(Args), "kernel argument count");
    auto& registry = KernelRegistry::instance();
    std::size_t const entry
        = registry.record(name, reinterpret_cast<void const*>(kernel), block, lds);
    if (registry.timing())
    {
        auto const [start, stop] = registry.timed_launch(entry, stream);
        if (start)
        {
            hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), lds, stream, start, stop,
                                  0u, static_cast<Params>(args)...);
            return;
        }
    }
    kernel<<<grid, block, lds, stream>>>(static_cast<Params>(args)...);
}

//---------------------------------------------------------------------------//
}  // namespace celeritas_amd
