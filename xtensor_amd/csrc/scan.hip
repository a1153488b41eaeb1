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

// The direct rows kernel applies: rank <= 2 rows layout with 32-bit offsets,
// a program that is one load of a leaf of the accumulation input type, the
// leaf and the output contiguous along the axis, rows starting 16-byte
// aligned, output elements of the same size, not bf16.
static const int kScanInDtype[AK_NUM] = {XG_FLOAT32, XG_FLOAT64, XG_FLOAT32, XG_FLOAT64,
                                         XG_INT32,   XG_UINT32,  XG_INT64,   XG_UINT64};

static bool scan_direct_ok(const ScanPlan& p) {
    const int* in_dt = kScanInDtype;
    if (p.layout != SL_ROWS || p.index64 || p.rank < 1 || p.rank > 2 || p.acc < 0 || p.acc >= AK_NUM) return false;
    if (p.rank == 1 && p.rows != 1) return false;
    if (p.prog.nops != 1 || p.prog.ops[0].op != MO_LOAD || p.prog.nleaves != 1) return false;
    const LeafView& l = p.leaves[0];
    const int esz = dtype_size(l.dtype);
    if (l.dtype != in_dt[p.acc] || lm_base(l.mode) != LM_CONTIG || l.stride[p.rank - 1] != 1) return false;
    if (p.out_dtype == XG_BFLOAT16 || dtype_size(p.out_dtype) != esz || p.out_stride[p.rank - 1] != 1) return false;
    if (reinterpret_cast<uintptr_t>(l.ptr) % 16 || reinterpret_cast<uintptr_t>(p.out_ptr) % 16) return false;
    if (p.rank == 2 && ((l.stride[0] * esz) % 16 || (p.out_stride[0] * esz) % 16)) return false;
    return true;
}

hipError_t launch_scan(const ScanPlan& p, int V, hipStream_t stream, const char** used) {
    if (p.rows == 0) return hipSuccess;
    if (V != 4) return hipErrorInvalidValue;
    ScanKernels k{};
    if (!scan_kernels_float(p.acc, p.op, &k) && !scan_kernels_int(p.acc, p.op, &k)) return hipErrorInvalidValue;
    dim3 block(256);
    if (k.direct[0][0] && scan_direct_ok(p)) {
        const bool native = p.out_dtype == kScanInDtype[p.acc] && p.rdt == kScanInDtype[p.acc];
        // rows longer than one wave's tile take a block of four waves each
        // (measured on MI355X for f32 [8192, 4096]: one wave per row walking
        // the row in sequence reaches 5.4 TB/s as a plain copy, four waves
        // sharing each row 5.9 TB/s, the same as a flat copy)
        const bool block_rows = p.len > k.wave_tile && p.rows <= 0x7fffffffu;
        const uint64_t blocks = block_rows ? uint64_t(p.rows) : (uint64_t(p.rows) + 3) / 4;
        if (used) *used = block_rows ? "scan_rows_direct_block" : "scan_rows_direct";
        hipLaunchKernelGGL(k.direct[native][block_rows], dim3(uint32_t(blocks)), block, 0, stream, p);
        return hipGetLastError();
    }
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
