//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file corecel/Precision.hh
//! Floating-point policy of the physics kernels.
//!
//! The physics headers (cross-section interpolation, interactors, MSC and
//! energy-loss fluctuations) are compiled twice: once in FP64 (the
//! reference's behaviour) and once in the mixed mode, selected by defining
//! CELER_MIXED_PRECISION before any include. In the mixed mode the
//! transcendental functions of sampling, interpolation and MSC (log, exp,
//! pow, sqrt, sin/cos, cbrt, log1p, expm1) are evaluated in FP32 with the
//! hardware v_log/v_exp/v_sqrt/v_sin/v_cos paths, and uniform draws are
//! 24-bit floats from one XORWOW word; positions, directions, safety,
//! field integration, energies and energy deposition stay FP64. Each
//! compilation's functions live in their own inline namespace (fp64 or
//! mixed), so the two sets never merge at link time, and kernel launchers
//! get a _mixed suffix.
//---------------------------------------------------------------------------//
#pragma once

#include <cmath>
#include <cstdint>

#include "Macros.hh"
#include "random/XorwowEngine.hh"

#if defined(CELER_MIXED_PRECISION) && CELER_MIXED_PRECISION
#    define CELER_PREC_NS mixed
#    define CELER_PREC_FN(NAME) NAME##_mixed
#    define CELER_PREC_IS_MIXED 1
#else
#    define CELER_PREC_NS fp64
#    define CELER_PREC_FN(NAME) NAME
#    define CELER_PREC_IS_MIXED 0
#endif

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
//---------------------------------------------------------------------------//
//! Type of uniform draws in the physics kernels
#if CELER_PREC_IS_MIXED
using canonical_real = float;
#else
using canonical_real = double;
#endif

/*!
 * Uniform draw for the physics kernels.
 *
 * FP64: the reference's 53-bit canonical double on [0, 1). Mixed: 24 bits of
 * one 32-bit word at the centre of their bin, (k + 1/2) 2^-24, so the draw
 * lies strictly inside (0, 1): a float product x 2^-32 can round to exactly
 * 1 (and be exactly 0), and the samplers' logarithms and inverse functions
 * would then step outside their physical ranges (a photon energy above the
 * electron's kinetic energy, an infinite free path).
 */
template<class Engine>
CELER_FORCEINLINE_FUNCTION canonical_real draw_canonical(Engine& rng)
{
#if CELER_PREC_IS_MIXED
    std::uint32_t const k = rng() >> 8;
    return (static_cast<float>(k) + 0.5f) * 5.9604644775390625e-8f;  // 2^-24
#else
    return generate_canonical<double>(rng);
#endif
}

namespace mth
{
#if CELER_PREC_IS_MIXED
// FP32 evaluation of double arguments: the fast hardware paths on the
// device; the host build (used only to exercise the mixed samplers on the
// CPU) evaluates the same functions with the FP32 C library
#    if CELER_DEVICE_COMPILE
#        define CELER_MTH_F32(DEVICE, HOST) DEVICE
#    else
#        define CELER_MTH_F32(DEVICE, HOST) HOST
#    endif
CELER_FORCEINLINE_FUNCTION double log(double x)
{
    return static_cast<double>(CELER_MTH_F32(__logf, ::logf)(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double exp(double x)
{
    return static_cast<double>(CELER_MTH_F32(__expf, ::expf)(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double sqrt(double x)
{
    return static_cast<double>(CELER_MTH_F32(__fsqrt_rn, ::sqrtf)(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double pow(double x, double y)
{
    return static_cast<double>(
        CELER_MTH_F32(__powf, ::powf)(static_cast<float>(x), static_cast<float>(y)));
}
CELER_FORCEINLINE_FUNCTION double sin(double x)
{
    return static_cast<double>(CELER_MTH_F32(__sinf, ::sinf)(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double cos(double x)
{
    return static_cast<double>(CELER_MTH_F32(__cosf, ::cosf)(static_cast<float>(x)));
}
#    undef CELER_MTH_F32
CELER_FORCEINLINE_FUNCTION double cbrt(double x)
{
    return static_cast<double>(::cbrtf(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double log1p(double x)
{
    return static_cast<double>(::log1pf(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double expm1(double x)
{
    return static_cast<double>(::expm1f(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double atan(double x)
{
    return static_cast<double>(::atanf(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double tan(double x)
{
    return static_cast<double>(::tanf(static_cast<float>(x)));
}
#else
using std::atan;
using std::cbrt;
using std::cos;
using std::exp;
using std::expm1;
using std::log;
using std::log1p;
using std::pow;
using std::sin;
using std::sqrt;
using std::tan;
#endif
}  // namespace mth

//---------------------------------------------------------------------------//
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
