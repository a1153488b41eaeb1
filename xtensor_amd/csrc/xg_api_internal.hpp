// xg_api_internal.hpp — helpers shared by the C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <exception>
#include <new>
#include <string>

#include "planner.hpp"
#include "reduce.hpp"

namespace xg {

void set_error(const std::string& m);
int device_cus();
void note_kernel(const char* name);
void check_hip(hipError_t e, const char* what);
// per (device, stream) scratch buffer of at least `bytes`
void* workspace(size_t bytes, hipStream_t stream);
// device address of xg_reduce_noalloc's fixed per-device array, resolved
// once per device (the current one)
char* scratch_setup();

// Device memory one call needs beyond its inputs and outputs, handed out by
// a bump allocator over a single block.  A call runs twice: a sizing pass
// (`measure`) plans every launch against placeholder addresses, launches
// nothing and records the bytes used; the real pass then takes the same
// regions, in the same order, from a block of that size.
struct Arena {
    bool measure = true;
    char* base = nullptr;
    size_t cap = 0;
    size_t used = 0;
    uint32_t* counters = nullptr;  // column reductions' arrival counters, or none
    // `bytes` (256-byte aligned) from the block; nullptr for 0 bytes
    char* take(size_t bytes);
};

// Where the block of a call comes from: `get(bytes)` returns at least
// `bytes` of device memory usable in stream order on the call's stream.
using BlockSource = char* (*)(size_t bytes, hipStream_t stream, void* ctx);

// out = reduce(e) (every reduction kind of xg_reduce), with its temporaries
// and split partials in one block from `src`; `counters`: the column
// reductions may merge their splits themselves with the stream's arrival
// counters (allocated on first use), else a combine kernel does.
void reduce_call(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, hipStream_t stream,
                 BlockSource src, void* ctx, bool counters);

// the stream's arrival counters (zeroed, kReduceCounters words)
uint32_t* reduce_counters(hipStream_t stream);

// Plan out = e with the compiled-kernel vector width, falling back to the
// generic interpreter's width when no compiled program matches.
EwPlan plan_assign_best(const xg_expr* e, const xg_tensor* out, int* V);
// Plan out = reduce(e) the same way.
RedPlan plan_reduce_best(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, int* V);

// Enqueue many elementwise plans: plans of one compiled program and rank go
// out as one fused launch (tables uploaded through the stream's staging
// buffer), the rest one launch each.  Returns the number of launches.
int launch_ew_many(const EwPlan* const* plans, const int* V, size_t n, hipStream_t stream);

// Run f, turning thrown errors into a status code and the thread's message.
template <class F> xg_status guard(F&& f) {
    try {
        f();
        set_error("");
        return XG_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return xg_status(e.code);
    } catch (const std::bad_alloc&) {
        set_error("host out of memory");
        return XG_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_error(e.what());
        return XG_INVALID_ARGUMENT;
    }
}

}  // namespace xg
