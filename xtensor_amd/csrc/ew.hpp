// ew.hpp — host entry points of the elementwise kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "xg_internal.hpp"

namespace xg {

#if !defined(__HIPCC_RTC__)  // host API (absent from run-time compiled sources)
// Launch out = program(leaves) for a plan built with vector width V.
// `used` (optional) receives the name of the kernel that ran.
hipError_t launch_ew(const EwPlan& p, int V, int cus, hipStream_t stream, const char** used);
#endif

// Name of the compiled specialisation that matches the plan, or nullptr.
const char* ew_find_spec(const EwPlan& p, int V);

// Route every launch through the dynamic interpreter (testing aid).
void ew_set_force_dynamic(int on);
// The same for plans made and launched on the calling thread only; returns
// the previous setting of that thread.
int ew_set_thread_force_dynamic(int on);
// True when a plan without an ahead-of-time program runs a run-time compiled
// one (compiling it on first use).
bool ew_jit_ok(const EwPlan& p, int V);
// Force one (1) or four (4) work units per thread in compiled kernels, or
// choose by size (0) (measurement aid).
void ew_set_units_per_thread(int upt);

// One expression of a fused launch: the compact (rank <= 4, 32-bit offset)
// part of an EwPlan that a compiled program needs at run time.
struct CompactPlan {
    uint32_t shape32[4];
    FastDiv div[4];
    FastDiv div_chunks;
    uint32_t nchunks;
    uint32_t inner;
    uint64_t units;
    int32_t cst[kMaxLeaves + 1][4];
    const char* leaf[kMaxLeaves];
    char* out_ptr;
    uint64_t consts[kMaxConsts];
};

// Index of the compiled program matching the plan (-1: none), and its name.
int ew_spec_index(const EwPlan& p, int V);
const char* ew_spec_name(int index);
CompactPlan make_compact(const EwPlan& p);

// Evaluate n plans of compiled program `spec_index` at normalised rank
// `rank` in one launch.  d_plans / d_prefix live in device memory;
// d_prefix[i] is the first work unit of plan i, d_prefix[n] the total.
#if !defined(__HIPCC_RTC__)
hipError_t launch_ew_multi(int spec_index, int rank, const CompactPlan* d_plans, const uint64_t* d_prefix, uint32_t n,
                           uint64_t total_units, hipStream_t stream);
#endif

}  // namespace xg
