//----------------------------------*-C++-*----------------------------------//
// ORACLE (test infrastructure only): uniform-field propagation.
//
// Follows field/DormandPrinceStepper.hh:102-222 (Dormand & Prince 1980
// RK5(4)7M tableau, Shampine's dense-output midpoint),
// field/MagFieldEquation.hh (dx/ds = p/|p|, dp/ds = (q/|p|) p x B),
// field/FieldDriver.hh:182-330 (chord search, accurate advance with the
// Geant4-style step control), field/detail/FieldUtils.hh (chord, sagitta,
// relative error, intercept test) and field/FieldPropagator.hh:150-331
// (substep loop against geometry boundaries). The sums are evaluated in
// the reference's order so that rounding matches.
//---------------------------------------------------------------------------//
#include "field.hh"

#include <algorithm>
#include <cmath>
#include <limits>

namespace orc
{
namespace
{
//---------------------------------------------------------------------------//
//! Dot product as a chain of fused multiply-adds (ArrayUtils.hh:90-99)
inline double dot3(double const* a, double const* b)
{
    return std::fma(a[2], b[2], std::fma(a[1], b[1], a[0] * b[0]));
}

inline void cross3(double const* a, double const* b, double* out)
{
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
}

//! y += a * k for both halves of the ODE state
inline void axpy(double a, OdeState const& k, OdeState* y)
{
    for (int i = 0; i < 3; ++i)
    {
        y->pos[i] = std::fma(a, k.pos[i], y->pos[i]);
        y->mom[i] = std::fma(a, k.mom[i], y->mom[i]);
    }
}

inline double fastpow(double a, double b)
{
    return std::exp(b * std::log(a));
}

//! Square of the truncation error relative to step and momentum
double rel_err_sq(OdeState const& err, double step, double const* mom)
{
    double errpos2 = dot3(err.pos, err.pos);
    double errvel2 = dot3(err.mom, err.mom);
    errpos2 /= step * step;
    errvel2 /= dot3(mom, mom);
    return std::max(errpos2, errvel2);
}

//! Sagitta of the chord from beg to end through mid
double distance_chord(double const* beg, double const* mid, double const* end)
{
    double bm[3], be[3], c[3];
    for (int i = 0; i < 3; ++i)
    {
        bm[i] = mid[i] - beg[i];
        be[i] = end[i] - beg[i];
    }
    cross3(be, bm, c);
    return std::sqrt(dot3(c, c) / dot3(be, be));
}

bool is_intercept_close(double const* pos, double const* dir, double distance,
                        double const* target, double tol)
{
    double d2 = 0;
    for (int i = 0; i < 3; ++i)
    {
        double const d = pos[i] - target[i] + distance * dir[i];
        d2 += d * d;
    }
    return d2 <= tol * tol;
}

struct DriverResult
{
    double step;
    OdeState state;
};

//---------------------------------------------------------------------------//
//! Adaptive driver (one per propagation call: the chord memory resets)
class Driver
{
  public:
    Driver(FieldOptions const& o, UniformFieldStepper const& s, double* max_chord)
        : o_(o), step_(s), max_chord_(*max_chord)
    {
    }

    DriverResult advance(double step, OdeState const& state)
    {
        if (step <= o_.minimum_step)
            return {step, step_(step, state).end_state};
        Chord out = this->find_next_chord(std::min(step, max_chord_), state);
        if (out.end.step < step)
            max_chord_ = out.end.step * (1 / o_.min_chord_shrink);
        if (out.err_sq > 1)
        {
            double const next = step * this->new_step_scale(out.err_sq);
            out.end = this->accurate_advance(out.end.step, state, next);
        }
        return out.end;
    }

  private:
    struct Chord
    {
        DriverResult end;
        double err_sq;
    };
    struct Integration
    {
        DriverResult end;
        double proposed;
    };

    FieldOptions const& o_;
    UniformFieldStepper const& step_;
    double& max_chord_;

    double eps_sq() const { return o_.epsilon_rel_max * o_.epsilon_rel_max; }

    double new_step_scale(double err_sq) const
    {
        return o_.safety * fastpow(err_sq, 0.5 * (err_sq > 1 ? o_.pshrink : o_.pgrow));
    }

    Chord find_next_chord(double step, OdeState const& state) const
    {
        StepperResult r;
        int remaining = o_.max_nsteps;
        bool ok = false;
        do
        {
            r = step_(step, state);
            double const dchord
                = distance_chord(state.pos, r.mid_state.pos, r.end_state.pos);
            if (dchord > o_.delta_chord + o_.dchord_tol)
                step *= std::max(std::sqrt(o_.delta_chord / dchord), o_.min_chord_shrink);
            else
                ok = true;
        } while (!ok && --remaining > 0);
        Chord out;
        out.end.step = step;
        out.end.state = r.end_state;
        out.err_sq = rel_err_sq(r.err_state, step, state.mom) / eps_sq();
        return out;
    }

    DriverResult
    accurate_advance(double step, OdeState const& state, double hinitial) const
    {
        double const end_length = step;
        double h = (hinitial > o_.initial_step_tol * step && hinitial < step) ? hinitial
                                                                              : step;
        double const h_threshold = o_.epsilon_step * step;
        Integration out;
        out.end.state = state;
        double length = 0;
        int remaining = o_.max_nsteps;
        bool ok = false;
        do
        {
            out = this->integrate_step(h, out.end.state);
            length += out.end.step;
            if (h < h_threshold || length >= end_length)
                ok = true;
            else
                h = std::min(std::max(out.proposed, o_.minimum_step), end_length - length);
        } while (!ok && --remaining > 0);
        out.end.step = std::min(length, step);
        return out.end;
    }

    Integration integrate_step(double step, OdeState const& state) const
    {
        if (step > o_.minimum_step)
            return this->one_good_step(step, state);
        StepperResult r = step_(step, state);
        Integration out;
        out.end.state = r.end_state;
        out.end.step = step;
        double const err_sq = rel_err_sq(r.err_state, step, state.mom) / eps_sq();
        out.proposed = step * this->new_step_scale(err_sq);
        return out;
    }

    Integration one_good_step(double step, OdeState const& state) const
    {
        StepperResult r;
        double err_sq;
        unsigned int remaining = o_.max_nsteps;
        bool ok = false;
        do
        {
            r = step_(step, state);
            err_sq = rel_err_sq(r.err_state, step, state.mom) / eps_sq();
            if (err_sq > 1)
                step *= std::max(this->new_step_scale(err_sq), o_.max_stepping_decrease);
            else
                ok = true;
        } while (!ok && --remaining > 0);
        Integration out;
        out.end.state = r.end_state;
        out.end.step = step;
        out.proposed = step * std::min(this->new_step_scale(err_sq), o_.max_stepping_increase);
        return out;
    }
};

//---------------------------------------------------------------------------//
}  // namespace

//---------------------------------------------------------------------------//
OdeState UniformFieldStepper::rhs(OdeState const& y) const
{
    double const inv = 1 / std::sqrt(dot3(y.mom, y.mom));
    OdeState r;
    for (int i = 0; i < 3; ++i)
        r.pos[i] = inv * y.mom[i];
    double c[3];
    cross3(y.mom, field, c);
    double const f = coeff * inv;
    for (int i = 0; i < 3; ++i)
        r.mom[i] = f * c[i];
    return r;
}

StepperResult UniformFieldStepper::operator()(double h, OdeState const& beg) const
{
    ++count;
    // Dormand-Prince RK5(4)7M: a_ij, the fifth-order weights (row 6) and
    // the differences to the embedded fourth-order weights
    constexpr double a11 = 0.2;
    constexpr double a21 = 0.075, a22 = 0.225;
    constexpr double a31 = 44 / 45., a32 = -56 / 15., a33 = 32 / 9.;
    constexpr double a41 = 19372 / 6561., a42 = -25360 / 2187., a43 = 64448 / 6561.,
                     a44 = -212 / 729.;
    constexpr double a51 = 9017 / 3168., a52 = -355 / 33., a53 = 46732 / 5247.,
                     a54 = 49 / 176., a55 = -5103 / 18656.;
    constexpr double a61 = 35 / 384., a63 = 500 / 1113., a64 = 125 / 192.,
                     a65 = -2187 / 6784., a66 = 11 / 84.;
    constexpr double d71 = a61 - 5179 / 57600., d73 = a63 - 7571 / 16695.,
                     d74 = a64 - 393 / 640., d75 = a65 + 92097 / 339200.,
                     d76 = a66 - 187 / 2100., d77 = -1 / 40.;
    // Shampine's midpoint weights
    constexpr double c71 = 6025192743. / 30085553152., c73 = 51252292925. / 65400821598.,
                     c74 = -2691868925. / 45128329728.,
                     c75 = 187940372067. / 1594534317056.,
                     c76 = -1776094331. / 19743644256., c77 = 11237099. / 235043384.;

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
    r.end_state = beg;
    axpy(a61 * h, k1, &r.end_state);
    axpy(a63 * h, k3, &r.end_state);
    axpy(a64 * h, k4, &r.end_state);
    axpy(a65 * h, k5, &r.end_state);
    axpy(a66 * h, k6, &r.end_state);
    OdeState const k7 = rhs(r.end_state);

    r.err_state = OdeState{{0, 0, 0}, {0, 0, 0}};
    axpy(d71 * h, k1, &r.err_state);
    axpy(d73 * h, k3, &r.err_state);
    axpy(d74 * h, k4, &r.err_state);
    axpy(d75 * h, k5, &r.err_state);
    axpy(d76 * h, k6, &r.err_state);
    axpy(d77 * h, k7, &r.err_state);

    double const hh = h / 2;
    r.mid_state = beg;
    axpy(c71 * hh, k1, &r.mid_state);
    axpy(c73 * hh, k3, &r.mid_state);
    axpy(c74 * hh, k4, &r.mid_state);
    axpy(c75 * hh, k5, &r.mid_state);
    axpy(c76 * hh, k6, &r.mid_state);
    axpy(c77 * hh, k7, &r.mid_state);
    return r;
}

//---------------------------------------------------------------------------//
FieldPropagatorState make_field_state(GeoState const& g, double momentum)
{
    FieldPropagatorState ps;
    for (int i = 0; i < 3; ++i)
    {
        ps.state.pos[i] = g.pos[i];
        ps.state.mom[i] = momentum * g.dir[i];
    }
    ps.max_chord = std::numeric_limits<double>::infinity();
    return ps;
}

FieldPropagation field_propagate(Geometry const& geo,
                                 GeoState& g,
                                 UniformFieldStepper const& stepper,
                                 double momentum,
                                 FieldOptions const& o,
                                 double step)
{
    FieldPropagatorState ps = make_field_state(g, momentum);
    return field_propagate(geo, g, stepper, o, ps, step);
}

FieldPropagation field_propagate(Geometry const& geo,
                                 GeoState& g,
                                 UniformFieldStepper const& stepper,
                                 FieldOptions const& o,
                                 FieldPropagatorState& ps,
                                 double step)
{
    FieldPropagation result;
    result.boundary = g.surf != none;
    Driver driver(o, stepper, &ps.max_chord);
    OdeState& state = ps.state;
    double const bump = o.delta_intersection * 0.1;
    double remaining = step;
    int substeps = o.max_substeps;
    do
    {
        DriverResult const sub = driver.advance(remaining, state);
        double chord_dir[3];
        for (int i = 0; i < 3; ++i)
            chord_dir[i] = sub.state.pos[i] - state.pos[i];
        double const chord = std::sqrt(dot3(chord_dir, chord_dir));
        for (int i = 0; i < 3; ++i)
            chord_dir[i] /= chord;
        if (chord >= o.minimum_step)
            geo_set_dir(geo, g, chord_dir);
        bool lin_boundary;
        double const lin = geo_find_next_step(geo, g, chord + o.delta_intersection,
                                              &lin_boundary);
        double const update = sub.step * lin / chord;
        if (!lin_boundary)
        {
            state = sub.state;
            result.boundary = false;
            result.distance += sub.step;
            remaining = step - result.distance;
            geo_move_internal_to(g, state.pos);
            --substeps;
        }
        else if (result.boundary && lin < bump)
        {
            remaining = sub.step / 2;
        }
        else if (update <= o.minimum_step
                 || is_intercept_close(state.pos, chord_dir, lin, sub.state.pos,
                                       o.delta_intersection)
                 || chord == 0)
        {
            result.boundary = (lin <= chord || result.distance + update <= step
                               || chord == 0);
            if (!result.boundary)
            {
                for (int i = 0; i < 3; ++i)
                    state.pos[i] = sub.state.pos[i];
                geo_move_internal_to(g, sub.state.pos);
            }
            result.distance += std::min(update, sub.step);
            for (int i = 0; i < 3; ++i)
                state.mom[i] = sub.state.mom[i];
            remaining = 0;
        }
        else
        {
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
            geo_move_to_boundary(g);
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
    geo_set_dir(geo, g, dir);

    if (result.distance == 0)
    {
        result.distance = std::min(bump, step);
        result.boundary = false;
        for (int i = 0; i < 3; ++i)
            state.pos[i] = std::fma(result.distance, dir[i], state.pos[i]);
        geo_move_internal_to(g, state.pos);
    }
    return result;
}

//---------------------------------------------------------------------------//
}  // namespace orc
