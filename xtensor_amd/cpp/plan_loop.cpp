// plan_loop.cpp — a C caller of the prepared-plan entry points: enqueue a
// list of plans in order, one xg_plan_execute each, `reps` times over, and
// time the whole loop on the host clock up to the final synchronisation of
// the stream.  It is the per-request path as a C or C++ host would drive it
// (no interpreter between the calls), so the time per plan is the library's
// own launch cost plus the kernels.  Built into its own small library, linked
// against libxg like any other caller.
#include <hip/hip_runtime.h>

#include <chrono>

#include "xg.h"

extern "C" int32_t xgl_time_plans(const xg_plan_t* plans, int64_t n, int32_t reps, xg_stream_t stream,
                                  double* seconds) {
    if (!plans || n < 0 || reps < 1 || !seconds) return XG_INVALID_ARGUMENT;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (hipStreamSynchronize(st) != hipSuccess) return XG_DEVICE_ERROR;
    const auto t0 = std::chrono::steady_clock::now();
    for (int32_t r = 0; r < reps; ++r)
        for (int64_t i = 0; i < n; ++i) {
            const xg_status s = xg_plan_execute(plans[i], nullptr, stream);
            if (s != XG_OK) return s;
        }
    if (hipStreamSynchronize(st) != hipSuccess) return XG_DEVICE_ERROR;
    *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return XG_OK;
}
