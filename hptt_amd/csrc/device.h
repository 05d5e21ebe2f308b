// device.h -- which device a plan is created for.
//
// The reference's front end picks a plugin built for the host's instruction
// set and returns nullptr when none fits (src/arch/arch.cc:74-89,
// src/hptt.cc:86-87).  This library carries code objects for gfx950 only, so
// plan, batch and server creation check the current HIP device the same way
// and fail with HPTTX_EUNSUPPORTED on any other architecture (instead of
// failing later, at the first launch, with "invalid device function").
#pragma once

#include <string>

namespace hpttx {

// The architecture the library's kernels are built for.
constexpr const char *kTargetArch = "gfx950";

// Architecture name of the current HIP device ("gfx950"; feature suffixes
// such as ":sramecc+:xnack-" removed), looked up once per device.  Returns
// false (and says why in err) when there is no current device.
bool device_arch(std::string &arch, std::string &err);

// kOk when the current device is a gfx950; kEUnsupported for any other
// architecture, kEHip when no device can be queried.  err says why.
int require_target_device(std::string &err);

}  // namespace hpttx
