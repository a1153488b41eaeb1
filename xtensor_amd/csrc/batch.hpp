// batch.hpp — batched evaluation of many uneven instances in one persistent
// launch (row-wise templates: softmax, sum, max over the last axis).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "xg_internal.hpp"

namespace xg {

enum BatchOp : int32_t { BO_SOFTMAX = XG_BATCH_SOFTMAX, BO_SUM = XG_BATCH_SUM, BO_MAX = XG_BATCH_AMAX };

// one instance as seen by the kernel
struct BItem {
    const char* in;
    char* out;
    int64_t in_rs;   // row stride of the input, elements
    int64_t out_rs;  // row stride of the output, elements (sum/max: output element stride)
    int32_t rows;
    int32_t cols;
    int32_t lpr;     // lanes per row (4..64), a function of cols only
    int32_t pad_;
};

// one unit of work: `nrows` consecutive rows of one instance, processed by one wave
struct BTask {
    int32_t item;
    int32_t row0;
    int32_t nrows;
    int32_t pad_;
};

// Schedule: tasks[wbeg[w], wbeg[w + 1]) belong to wave w of the grid
// (assigned on the host, longest first onto the least loaded wave); the
// tasks from dyn_begin on are handed out by a work queue to waves that have
// finished their own list.
struct BatchArgs {
    const BItem* items;
    const BTask* tasks;
    const uint32_t* wbeg;  // grid waves + 1 entries
    uint32_t dyn_begin;
    uint32_t ntasks;
    uint32_t* counter;     // work-queue head, reset to 0 before each launch
};

// lanes per row for a row of `cols` elements; this fixes the per-row
// reduction tree, so it must depend on nothing but `cols`
inline int lanes_per_row(int64_t cols) {
    if (cols <= 4) return 4;
    if (cols <= 8) return 8;
    if (cols <= 16) return 16;
    if (cols <= 32) return 32;
    return 64;
}

// Elements per lane a row of `cols` is held in (a power of two), or 0 when
// the row is longer than the register budget (32 four-byte or 16 eight-byte
// values per lane) and is streamed from memory instead.
XG_HD constexpr int batch_ne_class(int cols, int lpr, int esz) {
    const int need = (cols + lpr - 1) / lpr;
    const int cap = esz == 8 ? 16 : 32;
    if (need > cap) return 0;
    int ne = 1;
    while (ne < need) ne <<= 1;
    return ne;
}

// Row groups (of 64 / lpr rows each) a wave holds at once for elements-per-lane
// class `ne`: enough to keep about `cap` values per lane in flight, at most 8.
XG_HD constexpr int batch_rows_at_once(int ne, int esz) {
    const int cap = esz == 8 ? 8 : 32;
    if (ne == 0) return 1;
    const int rb = cap / ne;
    return rb < 1 ? 1 : (rb > 8 ? 8 : rb);
}

// Kernel view of one instance (strides normalised, lanes per row chosen).
BItem batch_item(const void* in, void* out, int64_t rows, int64_t cols, int64_t in_rs, int64_t out_rs, int op,
                 int dtype);

// Split instances into tasks of about task_elems elements (whole row-group
// iterations of one wave), in instance order; and the estimated cost of one.
void batch_split_tasks(const BItem* items, size_t n, int op, int dtype, int task_elems, std::vector<BTask>& tasks);
int64_t batch_task_cost(const BItem& it, const BTask& t, int dtype);

// Enqueue the queue reset and the row kernel of (op, dtype): blocks x block
// threads; a.wbeg must hold blocks * block / 64 + 1 entries.
hipError_t launch_batch_rows(int op, int dtype, const BatchArgs& a, int blocks, int block, hipStream_t stream);

}  // namespace xg
