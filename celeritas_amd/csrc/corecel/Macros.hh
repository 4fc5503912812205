//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/Macros.hh
//! Function annotations shared by host code and gfx950 kernels.
//---------------------------------------------------------------------------//
#pragma once

#if defined(__HIPCC__)
#    include <hip/hip_runtime.h>
#    define CELER_FUNCTION inline __host__ __device__
#    define CELER_FORCEINLINE_FUNCTION __host__ __device__ __forceinline__
#    define CELER_DEVICE_FUNCTION __device__
#else
#    define CELER_FUNCTION inline
#    define CELER_FORCEINLINE_FUNCTION inline __attribute__((always_inline))
#    define CELER_DEVICE_FUNCTION
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#    define CELER_DEVICE_COMPILE 1
#else
#    define CELER_DEVICE_COMPILE 0
#endif

#define CELER_LIKELY(x) __builtin_expect(!!(x), 1)
#define CELER_UNLIKELY(x) __builtin_expect(!!(x), 0)
