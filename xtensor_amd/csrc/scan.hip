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

// clears the one-pass tiled scan's ticket counters, timeout count and
// publications (a kernel, not a memset, so that the pair can be captured in a
// HIP graph); `n2` = publications / 2
__global__ __launch_bounds__(256) void scan_tiles_1p_init_kernel(uint32_t* ctrl, uint4* words, uint64_t n2) {
    const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
    if (i <= kTileMaxClasses) ctrl[i * kTileCtrlStride] = 0;  // the counters, then the timeouts
    if (i < n2) words[i] = make_uint4(0, 0, 0, 0);
}

static const int kScanInDtype[AK_NUM] = {XG_FLOAT32, XG_FLOAT64, XG_FLOAT32, XG_FLOAT64,
                                         XG_INT32,   XG_UINT32,  XG_INT64,   XG_UINT64};

hipError_t launch_scan(const ScanPlan& p, int V, void* ws, hipStream_t stream, const char** used) {
    if (p.rows == 0) return hipSuccess;
    if (V != 4) return hipErrorInvalidValue;
    ScanKernels k{};
    if (!scan_kernels_float(p.acc, p.op, &k) && !scan_kernels_int(p.acc, p.op, &k)) return hipErrorInvalidValue;
    dim3 block(256);
    if (p.tiled != ST_NONE) {
        const bool trans = p.tiled == ST_TRANS;
        TFn scan = k.tiles[p.tile_native ? 1 : 0][trans ? 1 : 0];
        if (!scan) return hipErrorInvalidValue;
        TileScan t = p.tile;
        t.part = static_cast<Store*>(ws);
        if (t.nchunks > 1 && !ws) return hipErrorInvalidValue;
        if (t.single) {  // one pass: chunks of kTileLbRows rows and a look-back per column
            TFn one = k.tiles1p[p.tile_native ? 1 : 0][trans ? 1 : 0];
            if (!one || t.chunk_rows != kTileLbRows) return hipErrorInvalidValue;
            const uint64_t total = uint64_t(t.outer) * t.nchunks * t.ntiles;
            if (t.classes < 1 || t.classes > kTileMaxClasses || (uint64_t(t.outer) * t.ntiles) % t.classes)
                return hipErrorInvalidValue;
            auto a256 = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
            char* w = static_cast<char*>(ws);
            t.ctrl = reinterpret_cast<uint32_t*>(w);
            w += kTileCtrlBytes;
            const uint64_t n = uint64_t(t.outer) * t.nchunks * t.ncols;
            uint64_t n2;  // 16-byte words the init kernel clears: the packed words, or the flags
            if (t.packed) {
                t.words = reinterpret_cast<uint64_t*>(w);
                n2 = n / 2;  // ncols: a multiple of 128
            } else {
                const uint64_t nf = uint64_t(t.outer) * t.nchunks * t.ntiles;
                t.flags = reinterpret_cast<uint32_t*>(w);
                w += a256(nf * 4);
                t.agg = reinterpret_cast<Store*>(w);
                w += a256(n * sizeof(Store));
                t.pre = reinterpret_cast<Store*>(w);
                n2 = (nf + 3) / 4;  // within the flags' 256-byte aligned block
            }
            hipLaunchKernelGGL(scan_tiles_1p_init_kernel, dim3(uint32_t((std::max<uint64_t>(n2, 16) + 255) / 256)), block, 0, stream, t.ctrl,
                               reinterpret_cast<uint4*>(t.packed ? static_cast<void*>(t.words) : static_cast<void*>(t.flags)), n2);
            hipLaunchKernelGGL(one, dim3(uint32_t(total)), block, 0, stream, t);
            if (used) *used = trans ? "scan_tiles_trans_1p" : "scan_tiles_1p";
            return hipGetLastError();
        }
        const uint32_t blocks = t.outer * t.nchunks * t.ntiles;
        if (t.nchunks > 1) {
            hipLaunchKernelGGL(k.tile_totals, dim3(blocks), block, 0, stream, t);
            hipLaunchKernelGGL(k.tile_carry, dim3(t.outer * ((t.ncols + 63u) / 64u)), block, 0, stream, t);
        }
        hipLaunchKernelGGL(scan, dim3(blocks), block, 0, stream, t);
        if (used) *used = trans ? "scan_tiles_trans" : "scan_tiles";
        return hipGetLastError();
    }
    if (p.lookback && p.lb_three) {  // reduce-then-scan over the look-back's chunks
        LbScan l = p.lb;
        if (!ws) return hipErrorInvalidValue;
        l.aggregate = static_cast<Store*>(ws);
        const bool native = p.out_dtype == kScanInDtype[p.acc] && p.rdt == kScanInDtype[p.acc];
        hipLaunchKernelGGL(k.lookback[kLbTotals][native ? 1 : 0], dim3(l.total), block, 0, stream, l);
        const bool sum = l.cpr <= kLbSumCarries;  // each chunk merges the totals before it itself
        if (!sum) hipLaunchKernelGGL(k.lb_carry, dim3(l.rows), block, 0, stream, l);
        hipLaunchKernelGGL(k.lb_apply[native ? 1 : 0][sum ? 1 : 0], dim3(l.total), block, 0, stream, l);
        if (used) *used = "scan_chunks3";
        return hipGetLastError();
    }
    if (p.lookback) {
        LbScan l = p.lb;
        if (!ws) return hipErrorInvalidValue;
        char* w = static_cast<char*>(ws);
        const size_t n = l.total;
        auto a256 = [](size_t b) { return (b + 255) & ~size_t(255); };
        l.ctrl = reinterpret_cast<uint32_t*>(w);
        w += 256;
        if (p.lb_packed) {
            l.words = reinterpret_cast<uint64_t*>(w);
        } else {
            l.flags = reinterpret_cast<uint32_t*>(w);
            w += a256(n * 4);
            l.aggregate = reinterpret_cast<Store*>(w);
            w += a256(n * sizeof(Store));
            l.prefix = reinterpret_cast<Store*>(w);
        }
        const bool native = p.out_dtype == kScanInDtype[p.acc] && p.rdt == kScanInDtype[p.acc];
        const uint32_t init_blocks = uint32_t((n + 255) / 256);
        if (p.lb_packed) hipLaunchKernelGGL(scan_lookback_init_kernel<true>, dim3(init_blocks), block, 0, stream, l);
        else hipLaunchKernelGGL(scan_lookback_init_kernel<false>, dim3(init_blocks), block, 0, stream, l);
        hipLaunchKernelGGL(k.lookback[kLbOnePass][native ? 1 : 0], dim3(l.total), block, 0, stream, l);
        if (used) *used = "scan_lookback";
        return hipGetLastError();
    }
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
