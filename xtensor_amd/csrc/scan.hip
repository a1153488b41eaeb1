// scan.hip — cumsum / cumprod along an axis on gfx950.
//
// Replaces accumulator_impl's copy-then-sequential recurrence
// (reducers/xaccumulator.hpp:215-297).  Rows layout (scan axis contiguous
// in the output): a workgroup owns a row and walks it in tiles of
// 256 x V elements; each thread scans its V elements, a wave scan (shuffles)
// and an LDS pass over the four wave totals give every thread its exclusive
// prefix, and the tile total carries into the next tile.  Cols layout (scan
// axis strided): a thread owns V adjacent columns and walks the axis.
// Float sums use the compensated accumulator of the reductions.
#include "scan_kernels.hpp"

namespace xg {

hipError_t launch_scan(const ScanPlan& p, int V, hipStream_t stream, const char** used) {
    if (p.rows == 0) return hipSuccess;
    if (V != 4) return hipErrorInvalidValue;
    ScanKernels k{};
    if (!scan_kernels_float(p.acc, p.op, &k) && !scan_kernels_int(p.acc, p.op, &k)) return hipErrorInvalidValue;
    dim3 block(256);
    if (p.layout == SL_ROWS) {
        uint32_t blocks = std::min<uint32_t>(p.rows, 1u << 20);
        if (used) *used = "scan_rows";
        hipLaunchKernelGGL(p.index64 ? k.rows64 : k.rows32, dim3(blocks), block, 0, stream, p);
    } else {
        uint32_t blocks = std::min<uint32_t>((p.rows + 255) / 256, 1u << 20);
        if (used) *used = "scan_cols";
        hipLaunchKernelGGL(p.index64 ? k.cols64 : k.cols32, dim3(blocks), block, 0, stream, p);
    }
    return hipGetLastError();
}

}  // namespace xg
