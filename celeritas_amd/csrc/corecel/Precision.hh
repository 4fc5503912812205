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
//! correctly-handled single-precision library functions (the device's OCML
//! logf/expf/powf/sinf/cosf, the host's C library: the same results to an
//! ulp or two and the same special cases), and uniform draws are 23-bit
//! floats from one XORWOW word; positions, directions, safety, field
//! integration, energies and energy deposition stay FP64. Each
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
 * FP64: the reference's 53-bit canonical double on [0, 1). Mixed: the top
 * 23 bits k of one 32-bit word at the centre of their bin, (k + 1/2) 2^-23,
 * which is exact in FP32 and lies in [2^-24, 1 - 2^-24]: the draw is never
 * 0 or 1, so the samplers' logarithms and inverse functions stay inside
 * their physical ranges. (With 24 bits, (2^24 - 1) + 1/2 rounds to 2^24
 * and the draw would be exactly 1.)
 */
template<class Engine>
CELER_FORCEINLINE_FUNCTION canonical_real draw_canonical(Engine& rng)
{
#if CELER_PREC_IS_MIXED
    std::uint32_t const k = rng() >> 9;
    return (static_cast<float>(k) + 0.5f) * 1.1920928955078125e-7f;  // 2^-23
#else
    return generate_canonical<double>(rng);
#endif
}

//! Iterations after which a rejection or counting loop abandons its sample
inline constexpr unsigned int max_sampling_iterations = 1u << 16;

/*!
 * Count one more iteration of a sampling loop; false once the budget is
 * spent, with the engine marked exhausted so that the caller fails the
 * track instead of returning the abandoned sample. Every loop of the
 * samplers ends within tens of iterations on physical inputs, so the
 * budget is only reached by inputs outside a sampler's domain (such as a
 * pair-production sample below threshold, whose Tsai-Urban loop can never
 * accept).
 */
template<class Engine>
CELER_FORCEINLINE_FUNCTION bool within_budget(unsigned int& iter, Engine& rng)
{
    if (CELER_LIKELY(++iter < max_sampling_iterations))
        return true;
    rng.exhausted = true;
    return false;
}

namespace mth
{
#if CELER_PREC_IS_MIXED
// FP32 evaluation of double arguments with the single-precision library
// functions (the same on the device and the host)
CELER_FORCEINLINE_FUNCTION double log(double x)
{
    return static_cast<double>(::logf(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double exp(double x)
{
    return static_cast<double>(::expf(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double sqrt(double x)
{
    return static_cast<double>(::sqrtf(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double pow(double x, double y)
{
    return static_cast<double>(::powf(static_cast<float>(x), static_cast<float>(y)));
}
CELER_FORCEINLINE_FUNCTION double sin(double x)
{
    return static_cast<double>(::sinf(static_cast<float>(x)));
}
CELER_FORCEINLINE_FUNCTION double cos(double x)
{
    return static_cast<double>(::cosf(static_cast<float>(x)));
}
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
CELER_FORCEINLINE_FUNCTION double tanh(double x)
{
    return static_cast<double>(::tanhf(static_cast<float>(x)));
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
using std::tanh;
#endif
}  // namespace mth

//---------------------------------------------------------------------------//
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
