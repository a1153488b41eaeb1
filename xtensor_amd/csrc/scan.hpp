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

// Tiled scans (scan_tiled.hpp): the input viewed as [outer, len, inner]
// with `inner` contiguous, the scan along `len`.
enum ScanTiled : int32_t {
    ST_NONE = 0,
    ST_COLS = 1,   // output contiguous along inner as well (cumsum(X, 0))
    ST_TRANS = 2   // output contiguous along len (cumsum(transpose(X), 1))
};

// Kernel argument of the tiled scans (pointers filled in at launch).
struct TileScan {
    const char* in;
    char* out;
    Store* part;          // [outer][nchunks][ncols] chunk totals, then carries
    uint32_t outer, len, inner;
    uint32_t nchunks, ntiles, chunk_rows, ncols;  // ncols = ntiles * 64 * V
    int64_t in_s_outer, in_s_len;                 // elements; inner stride 1
    int64_t out_s_outer, out_s_len, out_s_inner;  // elements
    uint64_t identity;
    int32_t out_dtype, rdt;
    int32_t nt_in, nt_out;
    // one pass (scan_tiled_1p.hpp): chunks of kTileLbRows rows, look-back
    // across the chunks; the workspace holds the ticket counter and the
    // publications
    int32_t single;
    uint32_t classes;     // ticket counters: chain k (= o * ntiles + tile) draws from counter k % classes
    uint32_t* ctrl;       // counter k at ctrl[k * kTileCtrlStride]; timeouts at ctrl[kTileTimeouts]
    uint64_t* words;      // 32-bit accumulators: [outer][nchunks][ncols] packed value + flag
    uint32_t* flags;      // 64-bit accumulators: [outer][nchunks][ntiles] flag ...
    Store* agg;           // ... [outer][nchunks][ncols] column totals
    Store* pre;           // ... and inclusive prefixes
    int32_t packed;       // 1: `words`; 0: `flags`, `agg`, `pre`
};

// Rows of one chunk of the one-pass tiled scan (held in registers whole).
constexpr uint32_t kTileLbRows = 64;
// its ticket counters (at most kTileMaxClasses, 256 bytes apart) and timeout
// count at the head of the workspace
constexpr uint32_t kTileMaxClasses = 16;
constexpr uint32_t kTileCtrlStride = 64;  // words
constexpr uint32_t kTileTimeouts = kTileMaxClasses * kTileCtrlStride;
constexpr size_t kTileCtrlBytes = 8192;
static_assert((kTileTimeouts + 1) * 4 <= kTileCtrlBytes, "the counters fit the workspace head");

// Segments of 64 lanes x 16 bytes per wave in one chunk of the look-back scan
// (a chunk is 4 waves x kLbSegments segments).
constexpr int kLbSegments = 16;

// Kernel argument of the look-back scan of long contiguous rows
// (scan_lookback.hpp); pointers into the workspace filled in at launch.
struct LbScan {
    const char* in;
    char* out;
    uint32_t* ctrl;       // [0] ticket counter, [1] look-back timeouts
    uint64_t* words;      // packed value + flag per chunk (32-bit accumulators)
    uint32_t* flags;      // flag per chunk (other accumulators) ...
    Store* aggregate;     // ... with the chunk totals
    Store* prefix;        // ... and the inclusive prefixes
    uint32_t rows, len, cpr, total, tile;  // cpr = chunks per row, total = rows * cpr
    int64_t in_row_stride, out_row_stride; // elements
    uint64_t identity;
    int32_t out_dtype, rdt;
    int32_t nt_in, nt_out;
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
    int32_t tiled;      // ScanTiled: the tiled kernels apply (then `tile` holds their shape)
    int32_t tile_native;// result and output of the input's type
    TileScan tile;
    int32_t lookback;   // 1: few long rows, the look-back kernel (then `lb` holds its shape)
    int32_t lb_packed;  // its publications are packed 64-bit words
    int32_t lb_three;   // the same chunks in three launches (totals, carries, results) instead
    LbScan lb;
};

// Measurement aid: workgroups per compute unit the tiled scans' chunking
// aims at, for plans made after the call (0: the default, 8).
void set_scan_blocks_per_cu(int n);
// Measurement aid: whether the tiled scans may take the one-pass form
// (default 1), for plans made after the call.
void set_scan_single_pass(int on);
// Measurement aid: whether few long rows run the one-pass look-back (1) or
// the three-launch reduce-then-scan over the same chunks (0, the default),
// for plans made after the call.
void set_scan_flat_single_pass(int on);

// The direct rows kernels apply: rank <= 2 rows layout with 32-bit offsets,
// a program that is one load of a leaf of the accumulation input type, the
// leaf and the output contiguous along the axis, rows starting 16-byte
// aligned, output elements of the same size, not bf16.
bool scan_direct_ok(const ScanPlan& p);

// Result dtype / shape of a scan.
void scan_result(const xg_expr* e, const xg_scan_desc* s, int* dtype, Shape* shape);

// Plan out = scan(e) along an axis (not the flattened form); `cus` = compute
// units of the device (sizes the tiled scans' chunks).
ScanPlan plan_scan(const xg_expr* e, const xg_scan_desc* s, const xg_tensor* out, int V, int cus);

// Bytes of workspace launch_scan needs for the plan (the tiled scans' chunk
// totals; 0 for the other kernels).
size_t scan_workspace_bytes(const ScanPlan& p);

#if !defined(__HIPCC_RTC__)
// `ws`: scan_workspace_bytes(p) bytes of device memory, usable in stream order.
hipError_t launch_scan(const ScanPlan& p, int V, void* ws, hipStream_t stream, const char** used);
#endif

}  // namespace xg
