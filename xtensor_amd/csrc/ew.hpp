// ew.hpp — host entry points of the elementwise kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "xg_internal.hpp"

namespace xg {

// Launch out = program(leaves) for a plan built with vector width V.
// `used` (optional) receives the name of the kernel that ran.
hipError_t launch_ew(const EwPlan& p, int V, int cus, hipStream_t stream, const char** used);

// Name of the compiled specialisation that matches the plan, or nullptr.
const char* ew_find_spec(const EwPlan& p, int V);

// Route every launch through the dynamic interpreter (testing aid).
void ew_set_force_dynamic(int on);

}  // namespace xg
