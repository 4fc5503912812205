//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file em/InteractorCommon.hh
//! Interaction result, secondary allocation and direction helpers.
//!
//! Direction math follows corecel/math/ArrayUtils.hh (rotate,
//! from_spherical, make_unit_vector with fma dot products) so device and
//! host results agree to the last few ulps.
//---------------------------------------------------------------------------//
#pragma once

#include <cmath>

#include "corecel/Precision.hh"
#include "global/CoreData.hh"

namespace celeritas_amd
{
inline namespace CELER_PREC_NS
{
//---------------------------------------------------------------------------//
//! Result of sampling an interaction (celeritas/phys/Interaction.hh)
struct Interaction
{
    enum Action : std::uint8_t
    {
        scattered,
        absorbed,
        unchanged,
        failed
    };
    Action action{scattered};
    double energy{0};
    double dir[3]{0, 0, 0};
    double energy_deposition{0};
    std::uint32_t sec_start{0};
    std::uint32_t num_secondaries{0};
};

//---------------------------------------------------------------------------//
CELER_FORCEINLINE_FUNCTION double dot3(double const* a, double const* b)
{
    double r = a[0] * b[0];
    r = std::fma(a[1], b[1], r);
    return std::fma(a[2], b[2], r);
}

CELER_FORCEINLINE_FUNCTION void make_unit(double* v)
{
    double const scale = 1 / mth::sqrt(dot3(v, v));
    v[0] *= scale;
    v[1] *= scale;
    v[2] *= scale;
}

/*!
 * sqrt(1 - c^2) for a cosine c. The mixed build evaluates c from FP32
 * transcendentals, so |c| can exceed 1 by an ulp; it returns 0 there (and
 * for a NaN) instead of a NaN that would spread into the track state.
 */
CELER_FORCEINLINE_FUNCTION double sin_of_cos(double c)
{
    double const s2 = 1 - c * c;
#if CELER_PREC_IS_MIXED
    return s2 > 0 ? mth::sqrt(s2) : 0.0;
#else
    return mth::sqrt(s2);
#endif
}

CELER_FORCEINLINE_FUNCTION void
from_spherical(double costheta, double phi, double* out)
{
#if CELER_PREC_IS_MIXED
    costheta = costheta > 1 ? 1.0 : (costheta < -1 ? -1.0 : costheta);
#endif
    double const sintheta = sin_of_cos(costheta);
    double s, c;
#if CELER_DEVICE_COMPILE
    sincos(phi, &s, &c);
#else
    s = mth::sin(phi);
    c = mth::cos(phi);
#endif
    out[0] = sintheta * c;
    out[1] = sintheta * s;
    out[2] = costheta;
}

//! Rotate a direction given in the frame whose z axis is rot
CELER_FUNCTION void rotate(double const* dir, double const* rot, double* out)
{
    double const sintheta = sin_of_cos(rot[2]);
    double cosphi, sinphi;
    if (sintheta >= 0.005)
    {
        double const inv = 1 / sintheta;
        cosphi = rot[0] * inv;
        sinphi = rot[1] * inv;
    }
    else if (sintheta > 0)
    {
        cosphi = rot[0] / mth::sqrt(rot[0] * rot[0] + rot[1] * rot[1]);
        sinphi = sin_of_cos(cosphi);
    }
    else
    {
        cosphi = 1;
        sinphi = 0;
    }
    double const a = rot[2] * dir[0] + sintheta * dir[2];
    out[0] = a * cosphi - sinphi * dir[1];
    out[1] = a * sinphi + cosphi * dir[1];
    out[2] = -sintheta * dir[0] + rot[2] * dir[2];
    make_unit(out);
}

//---------------------------------------------------------------------------//
/*!
 * Allocate secondaries on the shared stack (StackAllocator semantics).
 *
 * Returns the start index or invalid_id when the stack is exhausted; on
 * failure the size is restored so a smaller request may still succeed.
 */
CELER_FUNCTION std::uint32_t allocate_secondaries(CoreStateRef const& s,
                                                  std::uint32_t n)
{
#if CELER_DEVICE_COMPILE
    std::uint32_t const start
        = atomicAdd(&s.counters->secondary_stack_size, n);
    if (start + n > s.secondaries.capacity)
    {
        atomicSub(&s.counters->secondary_stack_size, n);
        return invalid_id;
    }
#else
    // Host (single-threaded) build of the samplers
    std::uint32_t const start = s.counters->secondary_stack_size;
    if (start + n > s.secondaries.capacity)
        return invalid_id;
    s.counters->secondary_stack_size = start + n;
#endif
    for (std::uint32_t i = 0; i < n; ++i)
    {
        Secondary& sec = s.secondaries.storage[start + i];
        sec.particle_id = invalid_id;
        sec.energy = 0;
        sec.dir[0] = sec.dir[1] = sec.dir[2] = 0;
    }
    return start;
}

//---------------------------------------------------------------------------//
}  // namespace CELER_PREC_NS
}  // namespace celeritas_amd
