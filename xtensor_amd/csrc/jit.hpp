// jit.hpp — run-time compiled specialisations.
//
// xtensor instantiates every expression's loop at C++ compile time.  Here the
// ahead-of-time table covers the programs known when the library is built;
// any other program of a compact plan gets the same kernel templates
// instantiated on its own constexpr program, compiled in-process with hipRTC
// for gfx950 on first use and cached (per process, and on disk under
// $XG_JIT_CACHE or ~/.cache/xtensor_amd/jit).  Smaller launches, and plans a
// compile fails for, stay on the generic interpreter.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "planner.hpp"
#include "reduce.hpp"

namespace xg {

// which kernel template of a reduction plan
enum JitRedKind : int { JR_ROWS_FAST1 = 0, JR_ROWS_FAST2, JR_COLS_FAST, JR_ROWS, JR_COLS };

struct JitKernel {
    hipFunction_t fn = nullptr;
    std::string name;  // "jit:<hash>" for diagnostics
};

// 0: off; 1 (default): launches of at least the element threshold
// (XG_JIT_MIN_ELEMENTS, default 2^20); 2: every compact plan.  The
// environment variable XG_JIT sets the initial mode.
void jit_set_mode(int mode);
int jit_get_mode();
// whether a launch over `elements` elements should use a compiled kernel
bool jit_wanted(uint64_t elements);
// On the calling thread only: 1 = hand out run-time compiled kernels only
// when this process has already loaded them (no compile, no disk read, no
// module load, so nothing is allocated on the device); 0 = normal.  Returns
// the previous setting.
int jit_set_thread_loaded_only(int on);

// compiled kernel for a compact elementwise plan at vector width V, or nullptr
const JitKernel* jit_ew(const EwPlan& p, int V);
// compiled reduction kernel of the given kind for plan p at width V, or nullptr
const JitKernel* jit_reduce(const RedPlan& p, int V, int kind);
// assign_rc_kernel (assign_reduce.hpp) on plan p's program: RAK / CAK = the
// row / column accumulator kinds (-1: none), UL work units per lane.
const JitKernel* jit_assign_rc(const EwPlan& p, int V, int rak, int rop, int cak, int cop, int ul);

// launch with the plan struct as the only kernel argument
hipError_t jit_launch(const JitKernel* k, uint32_t grid, uint32_t block, const void* arg, size_t arg_bytes,
                      hipStream_t stream);

struct JitStats {
    int64_t compiles, disk_hits, memory_hits, failures;
};
JitStats jit_stats();
const char* jit_last_error();

}  // namespace xg
