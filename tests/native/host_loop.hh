// Interface between the host step loop's driver (FP64 bodies) and its
// mixed-precision translation unit.
#pragma once

#include "global/CoreData.hh"

namespace hl
{
using celeritas_amd::CoreParamsRef;
using celeritas_amd::CoreStateRef;
using celeritas_amd::ModelClass;
using celeritas_amd::size_type;
using celeritas_amd::TrackStatus;

void mixed_pre_step(CoreParamsRef const&, CoreStateRef const&, size_type, TrackStatus);
void mixed_along_charged(CoreParamsRef const&, CoreStateRef const&, size_type, TrackStatus);
void mixed_discrete(CoreParamsRef const&, CoreStateRef const&, size_type);
void mixed_interaction(CoreParamsRef const&, CoreStateRef const&, size_type, ModelClass);
}  // namespace hl
