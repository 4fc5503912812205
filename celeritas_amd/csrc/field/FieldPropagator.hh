//----------------------------------*-C++-*----------------------------------//
// celeritas_amd: MI355X-native per-step transport loop
//---------------------------------------------------------------------------//
//! \file field/FieldPropagator.hh
//! Charged-track propagation in a uniform magnetic field (device code).
//!
//! Dormand-Prince RK5(4)7M integration of the Lorentz equation with
//! Shampine's midpoint (field/DormandPrinceStepper.hh:102-222,
//! field/MagFieldEquation.hh), the adaptive chord/accuracy driver
//! (field/FieldDriver.hh:182-330) and the substep loop that intersects each
//! chord with the geometry (field/FieldPropagator.hh:150-331). Everything
//! stays in registers: the ODE state is 6 doubles, the seven stages are
//! consumed as they are produced, and a propagation is one thread's loop, so
//! the kernel's cost is its FP64 FMA chain (dense VALU work, no tables).
//! Sums use fused multiply-adds in the reference's order (its axpy and
//! dot_product are fma chains) so host and device round alike.
//---------------------------------------------------------------------------//
#pragma once

#include <cmath>

#include "corecel/Macros.hh"
#include "geo/OrangeLiteTracker.hh"
#include "FieldData.hh"

namespace celeritas_amd
{
namespace field
{
//! Curvature per cm for q = 1 e, p = 1 MeV/c and B = 1 T (c / 1e8)
inline constexpr double lorentz_per_cm_tesla = 2.99792458;

//---------------------------------------------------------------------------//
struct OdeState
{
    double pos[3];
    double mom[3];
};

struct StepperResult
{
    OdeState mid;
    OdeState end;
    OdeState err;
};

struct DriverResult
{
    double step;
    OdeState state;
};

CELER_FORCEINLINE_FUNCTION double dot3(double const* a, double const* b)
{
    return std::fma(a[2], b[2], std::fma(a[1], b[1], a[0] * b[0]));
}

CELER_FORCEINLINE_FUNCTION void axpy(double a, OdeState const& k, OdeState* y)
{
    for (int i = 0; i < 3; ++i)
    {
        y->pos[i] = std::fma(a, k.pos[i], y->pos[i]);
        y->mom[i] = std::fma(a, k.mom[i], y->mom[i]);
    }
}

//---------------------------------------------------------------------------//
/*!
 * Uniform-field equation of motion integrated by Dormand-Prince RK5(4).
 */
struct UniformFieldStepper
{
    double b[3];  //!< [T]
    double coeff;  //!< charge [e] * lorentz_per_cm_tesla

    CELER_FUNCTION OdeState rhs(OdeState const& y) const
    {
        double const inv = 1 / std::sqrt(dot3(y.mom, y.mom));
        OdeState r;
        for (int i = 0; i < 3; ++i)
            r.pos[i] = inv * y.mom[i];
        double const f = coeff * inv;
        r.mom[0] = f * (y.mom[1] * b[2] - y.mom[2] * b[1]);
        r.mom[1] = f * (y.mom[2] * b[0] - y.mom[0] * b[2]);
        r.mom[2] = f * (y.mom[0] * b[1] - y.mom[1] * b[0]);
        return r;
    }

    CELER_FUNCTION StepperResult operator()(double h, OdeState const& beg) const
    {
        // Butcher tableau of RK5(4)7M, fifth-order weights (row 6) and
        // fifth-minus-fourth-order weights; Shampine's midpoint weights
        constexpr double a11 = 0.2;
        constexpr double a21 = 0.075, a22 = 0.225;
        constexpr double a31 = 44 / 45., a32 = -56 / 15., a33 = 32 / 9.;
        constexpr double a41 = 19372 / 6561., a42 = -25360 / 2187.,
                         a43 = 64448 / 6561., a44 = -212 / 729.;
        constexpr double a51 = 9017 / 3168., a52 = -355 / 33., a53 = 46732 / 5247.,
                         a54 = 49 / 176., a55 = -5103 / 18656.;
        constexpr double a61 = 35 / 384., a63 = 500 / 1113., a64 = 125 / 192.,
                         a65 = -2187 / 6784., a66 = 11 / 84.;
        constexpr double d71 = a61 - 5179 / 57600., d73 = a63 - 7571 / 16695.,
                         d74 = a64 - 393 / 640., d75 = a65 + 92097 / 339200.,
                         d76 = a66 - 187 / 2100., d77 = -1 / 40.;
        constexpr double c71 = 6025192743. / 30085553152.,
                         c73 = 51252292925. / 65400821598.,
                         c74 = -2691868925. / 45128329728.,
                         c75 = 187940372067. / 1594534317056.,
                         c76 = -1776094331. / 19743644256.,
                         c77 = 11237099. / 235043384.;

        OdeState const k1 = rhs(beg);
        OdeState s = beg;
        axpy(a11 * h, k1, &s);
        OdeState const k2 = rhs(s);
        s = beg;
        axpy(a21 * h, k1, &s);
        axpy(a22 * h, k2, &s);
        OdeState const k3 = rhs(s);
        s = beg;
        axpy(a31 * h, k1, &s);
        axpy(a32 * h, k2, &s);
        axpy(a33 * h, k3, &s);
        OdeState const k4 = rhs(s);
        s = beg;
        axpy(a41 * h, k1, &s);
        axpy(a42 * h, k2, &s);
        axpy(a43 * h, k3, &s);
        axpy(a44 * h, k4, &s);
        OdeState const k5 = rhs(s);
        s = beg;
        axpy(a51 * h, k1, &s);
        axpy(a52 * h, k2, &s);
        axpy(a53 * h, k3, &s);
        axpy(a54 * h, k4, &s);
        axpy(a55 * h, k5, &s);
        OdeState const k6 = rhs(s);

        StepperResult r;
        r.end = beg;
        axpy(a61 * h, k1, &r.end);
        axpy(a63 * h, k3, &r.end);
        axpy(a64 * h, k4, &r.end);
        axpy(a65 * h, k5, &r.end);
        axpy(a66 * h, k6, &r.end);
        OdeState const k7 = rhs(r.end);

        r.err = OdeState{{0, 0, 0}, {0, 0, 0}};
        axpy(d71 * h, k1, &r.err);
        axpy(d73 * h, k3, &r.err);
        axpy(d74 * h, k4, &r.err);
        axpy(d75 * h, k5, &r.err);
        axpy(d76 * h, k6, &r.err);
        axpy(d77 * h, k7, &r.err);

        double const hh = h / 2;
        r.mid = beg;
        axpy(c71 * hh, k1, &r.mid);
        axpy(c73 * hh, k3, &r.mid);
        axpy(c74 * hh, k4, &r.mid);
        axpy(c75 * hh, k5, &r.mid);
        axpy(c76 * hh, k6, &r.mid);
        axpy(c77 * hh, k7, &r.mid);
        return r;
    }
};

//---------------------------------------------------------------------------//
//! Square of the step error relative to the step and starting momentum
CELER_FUNCTION double rel_err_sq(OdeState const& err, double step, double const* mom)
{
    double const p = dot3(err.pos, err.pos) / (step * step);
    double const m = dot3(err.mom, err.mom) / dot3(mom, mom);
    return p > m ? p : m;
}

//! Sagitta of the chord beg->end through mid
CELER_FUNCTION double
distance_chord(double const* beg, double const* mid, double const* end)
{
    double bm[3], be[3];
    for (int i = 0; i < 3; ++i)
    {
        bm[i] = mid[i] - beg[i];
        be[i] = end[i] - beg[i];
    }
    double const c[3] = {be[1] * bm[2] - be[2] * bm[1],
                         be[2] * bm[0] - be[0] * bm[2],
                         be[0] * bm[1] - be[1] * bm[0]};
    return std::sqrt(dot3(c, c) / dot3(be, be));
}

//---------------------------------------------------------------------------//
/*!
 * Adaptive step driver; one per propagation (its chord memory resets).
 */
class FieldDriver
{
  public:
    CELER_FUNCTION FieldDriver(FieldDriverOptions const& o,
                               UniformFieldStepper const& s,
                               double* max_chord)
        : o_(o), step_(s), max_chord_(*max_chord)
    {
    }

    CELER_FUNCTION DriverResult advance(double step, OdeState const& state)
    {
        if (step <= o_.minimum_step)
            return {step, step_(step, state).end};
        // Chord search: shrink until the sagitta is within delta_chord
        double h = step < max_chord_ ? step : max_chord_;
        StepperResult r;
        int remaining = o_.max_nsteps;
        bool ok = false;
        do
        {
            r = step_(h, state);
            double const dchord = distance_chord(state.pos, r.mid.pos, r.end.pos);
            if (dchord > o_.delta_chord + o_.dchord_tol)
            {
                double const scale = std::sqrt(o_.delta_chord / dchord);
                h *= scale > o_.min_chord_shrink ? scale : o_.min_chord_shrink;
            }
            else
            {
                ok = true;
            }
        } while (!ok && --remaining > 0);
        DriverResult out{h, r.end};
        double const err_sq = rel_err_sq(r.err, h, state.mom) / this->eps_sq();
        if (out.step < step)
            max_chord_ = out.step * (1 / o_.min_chord_shrink);
        if (err_sq > 1)
        {
            double const next = step * this->new_step_scale(err_sq);
            out = this->accurate_advance(out.step, state, next);
        }
        return out;
    }

  private:
    FieldDriverOptions const& o_;
    UniformFieldStepper const& step_;
    double& max_chord_;

    CELER_FORCEINLINE_FUNCTION double eps_sq() const
    {
        return o_.epsilon_rel_max * o_.epsilon_rel_max;
    }

    CELER_FUNCTION double new_step_scale(double err_sq) const
    {
        double const e = 0.5 * (err_sq > 1 ? o_.pshrink : o_.pgrow);
        return o_.safety * std::exp(e * std::log(err_sq));
    }

    //! Integrate [0, step] in accurate substeps (FieldDriver::accurate_advance)
    CELER_FUNCTION DriverResult accurate_advance(double step,
                                                 OdeState const& state,
                                                 double hinitial) const
    {
        double const end_length = step;
        double h = (hinitial > o_.initial_step_tol * step && hinitial < step) ? hinitial
                                                                              : step;
        double const h_threshold = o_.epsilon_step * step;
        DriverResult out{0, state};
        double length = 0;
        int remaining = o_.max_nsteps;
        bool ok = false;
        do
        {
            double proposed;
            out = this->integrate_step(h, out.state, &proposed);
            length += out.step;
            if (h < h_threshold || length >= end_length)
            {
                ok = true;
            }
            else
            {
                double const lo = proposed > o_.minimum_step ? proposed : o_.minimum_step;
                double const rest = end_length - length;
                h = lo < rest ? lo : rest;
            }
        } while (!ok && --remaining > 0);
        out.step = length < step ? length : step;
        return out;
    }

    CELER_FUNCTION DriverResult integrate_step(double step,
                                               OdeState const& state,
                                               double* proposed) const
    {
        if (step > o_.minimum_step)
        {
            // One good step: shrink until the error is within tolerance
            StepperResult r;
            double err_sq;
            unsigned int remaining = o_.max_nsteps;
            bool ok = false;
            do
            {
                r = step_(step, state);
                err_sq = rel_err_sq(r.err, step, state.mom) / this->eps_sq();
                if (err_sq > 1)
                {
                    double const s = this->new_step_scale(err_sq);
                    step *= s > o_.max_stepping_decrease ? s : o_.max_stepping_decrease;
                }
                else
                {
                    ok = true;
                }
            } while (!ok && --remaining > 0);
            double const s = this->new_step_scale(err_sq);
            *proposed = step * (s < o_.max_stepping_increase ? s : o_.max_stepping_increase);
            return {step, r.end};
        }
        StepperResult r = step_(step, state);
        double const err_sq = rel_err_sq(r.err, step, state.mom) / this->eps_sq();
        *proposed = step * this->new_step_scale(err_sq);
        return {step, r.end};
    }
};

//---------------------------------------------------------------------------//
struct FieldPropagation
{
    double distance;
    bool boundary;
    bool looping;
};

//! Whether the linear intercept lands within tol of the substep's end
CELER_FORCEINLINE_FUNCTION bool is_intercept_close(double const* pos,
                                                   double const* dir,
                                                   double distance,
                                                   double const* target,
                                                   double tol)
{
    double d2 = 0;
    for (int i = 0; i < 3; ++i)
    {
        double const d = pos[i] - target[i] + distance * dir[i];
        d2 += d * d;
    }
    return d2 <= tol * tol;
}

//! What one propagator object keeps between calls: the ODE state and the
//! driver's chord memory (transport uses a fresh one per step)
struct PropagatorState
{
    OdeState state;
    double max_chord;
};

CELER_FUNCTION PropagatorState make_propagator_state(geo::GeoTrack const& g,
                                                     double momentum)
{
    PropagatorState ps;
    for (int i = 0; i < 3; ++i)
    {
        ps.state.pos[i] = g.pos[i];
        ps.state.mom[i] = momentum * g.dir[i];
    }
    ps.max_chord = HUGE_VAL;
    return ps;
}

/*!
 * Move a charged track up to `step` along its curved path, stopping at
 * the first boundary (FieldPropagator::operator(), FieldPropagator.hh:
 * 150-331). The geometry direction is set from the final momentum.
 */
CELER_FUNCTION FieldPropagation propagate(GeoParamsRef const& geo,
                                          geo::GeoTrack& g,
                                          UniformFieldStepper const& stepper,
                                          FieldDriverOptions const& o,
                                          PropagatorState& ps,
                                          double step)
{
    FieldPropagation result{0, g.is_on_boundary(), false};
    FieldDriver driver(o, stepper, &ps.max_chord);
    OdeState& state = ps.state;
    double const bump = o.delta_intersection * 0.1;
    double remaining = step;
    int substeps = o.max_substeps;
    do
    {
        DriverResult const sub = driver.advance(remaining, state);
        double cdir[3];
        for (int i = 0; i < 3; ++i)
            cdir[i] = sub.state.pos[i] - state.pos[i];
        double const chord = std::sqrt(dot3(cdir, cdir));
        for (int i = 0; i < 3; ++i)
            cdir[i] /= chord;
        if (chord >= o.minimum_step)
            geo::set_dir(geo, g, cdir);
        geo::Propagation const lin
            = geo::find_next_step(geo, g, chord + o.delta_intersection);
        double const update = sub.step * lin.distance / chord;
        if (!lin.boundary)
        {
            // Accept the substep inside the volume
            state = sub.state;
            result.boundary = false;
            result.distance += sub.step;
            remaining = step - result.distance;
            geo::move_internal(g, state.pos);
            --substeps;
        }
        else if (result.boundary && lin.distance < bump)
        {
            // Heading back through the surface just crossed: halve
            remaining = sub.step / 2;
        }
        else if (update <= o.minimum_step
                 || is_intercept_close(state.pos, cdir, lin.distance, sub.state.pos,
                                       o.delta_intersection)
                 || chord == 0)
        {
            // Close enough: stop on the boundary (or at the substep end)
            result.boundary = (lin.distance <= chord
                               || result.distance + update <= step || chord == 0);
            if (!result.boundary)
            {
                for (int i = 0; i < 3; ++i)
                    state.pos[i] = sub.state.pos[i];
                geo::move_internal(g, sub.state.pos);
            }
            result.distance += update < sub.step ? update : sub.step;
            for (int i = 0; i < 3; ++i)
                state.mom[i] = sub.state.mom[i];
            remaining = 0;
        }
        else
        {
            // Intercept too far from the substep end: retry shorter
            remaining = update;
        }
    } while (remaining > o.minimum_step && substeps > 0);

    if (substeps == 0 && result.distance < step)
    {
        result.looping = true;
    }
    else if (result.distance > 0)
    {
        if (result.boundary)
        {
            geo::move_to_boundary(g);
            for (int i = 0; i < 3; ++i)
                state.pos[i] = g.pos[i];
        }
        else if (result.distance < step)
        {
            result.distance = step;
        }
    }

    double const inv = 1 / std::sqrt(dot3(state.mom, state.mom));
    double dir[3];
    for (int i = 0; i < 3; ++i)
        dir[i] = state.mom[i] * inv;
    geo::set_dir(geo, g, dir);

    if (result.distance == 0)
    {
        // Stuck: bump along the new direction
        result.distance = bump < step ? bump : step;
        result.boundary = false;
        for (int i = 0; i < 3; ++i)
            state.pos[i] = std::fma(result.distance, dir[i], state.pos[i]);
        geo::move_internal(g, state.pos);
    }
    return result;
}

//! Propagation by a fresh propagator (the along-step case)
CELER_FUNCTION FieldPropagation propagate(GeoParamsRef const& geo,
                                          geo::GeoTrack& g,
                                          UniformFieldStepper const& stepper,
                                          FieldDriverOptions const& o,
                                          double momentum,
                                          double step)
{
    PropagatorState ps = make_propagator_state(g, momentum);
    return propagate(geo, g, stepper, o, ps, step);
}

//---------------------------------------------------------------------------//
}  // namespace field
}  // namespace celeritas_amd
