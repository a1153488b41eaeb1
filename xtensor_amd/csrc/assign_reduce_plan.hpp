// assign_reduce_plan.hpp — the kernel argument of a fused assignment with
// row and / or column reductions (assign_reduce.hpp); shared by the host code
// that plans it and the kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "xg_internal.hpp"

namespace xg {

struct FusedRC {
    // row reduction: one output per row, finished as finish() does for FZ_NONE
    char* row_out;
    int64_t row_stride;   // elements between the outputs of consecutive rows
    int32_t row_dtype, row_vdt, row_rdt, row_has_initial;
    uint64_t row_identity, row_initial;
    // column reduction: per band, one accumulator per column
    Store* col_partials;
    uint64_t col_identity;
    uint32_t rows;           // R
    uint32_t cols;           // C
    uint32_t upr;            // work units per row
    uint32_t rows_per_band;
};

struct AssignRC {
    EwPlan plan;
    FusedRC f;
};

}  // namespace xg
