// scan.hpp — plans and host entry points of the accumulator (cumsum /
// cumprod) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "planner.hpp"
#include "reduce.hpp"
#include "xg_internal.hpp"

namespace xg {

enum ScanLayout : int32_t {
    SL_ROWS = 0,  // scan axis is the output's contiguous dim: a workgroup per row
    SL_COLS = 1   // scan axis strided: a thread per V adjacent columns walks the axis
};

struct ScanPlan {
    int32_t layout;
    int32_t rank;       // SL_ROWS: [outer..., axis]; SL_COLS: [outer..., axis, inner]
    int32_t acc;        // AccKind (sums compensated for floats)
    int32_t op;         // RO_ADD or RO_MUL
    int32_t rdt;        // result type R
    int32_t out_dtype;
    int32_t index64;
    int32_t out_mode;   // LM_CONTIG or LM_STRIDED, | LM_NT for large outputs
    int64_t shape[kMaxRank];
    FastDiv div[kMaxRank];
    uint32_t rows;      // SL_ROWS: rows; SL_COLS: outer * ceil(inner / V)
    uint32_t len;       // length of the scan axis
    uint32_t inner;     // SL_COLS: length of the last dim
    uint32_t nchunks_in;// SL_COLS: ceil(inner / V)
    FastDiv div_nchunks_in;
    uint64_t identity;  // 0 or 1 in acc representation
    char* out_ptr;
    int64_t out_stride[kMaxRank];
    LeafView leaves[kMaxLeaves];
    Program prog;       // element value converted to the accumulation input type
};

// Result dtype / shape of a scan.
void scan_result(const xg_expr* e, const xg_scan_desc* s, int* dtype, Shape* shape);

// Plan out = scan(e) along an axis (not the flattened form).
ScanPlan plan_scan(const xg_expr* e, const xg_scan_desc* s, const xg_tensor* out, int V);

#if !defined(__HIPCC_RTC__)
hipError_t launch_scan(const ScanPlan& p, int V, hipStream_t stream, const char** used);
#endif

}  // namespace xg
