// Mixed-precision per-track step bodies for the host step loop
// (host_loop.cc): the same executors the mixed kernels run (pre-step,
// charged along-step, discrete select, interactions), compiled for the CPU
// in their own inline namespace, behind plain functions.
#define CELER_MIXED_PRECISION 1
#include "global/AlongStepExecutor.hh"
#include "global/StepExecutors.hh"

#include "host_loop.hh"

namespace hl
{
using namespace celeritas_amd;

void mixed_pre_step(CoreParamsRef const& p, CoreStateRef const& s, size_type slot,
                    TrackStatus st)
{
    exec::pre_step_track(p, s, slot, st);
}

void mixed_along_charged(CoreParamsRef const& p, CoreStateRef const& s, size_type slot,
                         TrackStatus st)
{
    if (p.field.enabled)
        exec::along_step_charged_track<true>(p, s, slot, st);
    else
        exec::along_step_charged_track<false>(p, s, slot, st);
}

void mixed_discrete(CoreParamsRef const& p, CoreStateRef const& s, size_type slot)
{
    exec::discrete_select_track(p, s, slot);
}

void mixed_interaction(CoreParamsRef const& p, CoreStateRef const& s, size_type slot,
                       ModelClass m)
{
    switch (m)
    {
#define HL_MODEL(M)                                          \
    case ModelClass::M:                                      \
        exec::interaction_track<ModelClass::M>(p, s, slot); \
        break;
        HL_MODEL(klein_nishina)
        HL_MODEL(livermore_pe)
        HL_MODEL(bethe_heitler)
        HL_MODEL(moller_bhabha)
        HL_MODEL(seltzer_berger)
        HL_MODEL(relativistic_brem)
        HL_MODEL(eplusgg)
        HL_MODEL(combined_brem)
#undef HL_MODEL
        default:
            break;
    }
}
}  // namespace hl
