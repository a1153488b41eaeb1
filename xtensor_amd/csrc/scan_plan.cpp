// scan_plan.cpp — host planning of cumsum / cumprod along an axis.
//
// xtensor's accumulator (reducers/xaccumulator.hpp:215-297) copies the
// expression into the result type R = decltype(f(init_type, value_type))
// and runs res[k] = f(res[k-1], res[k]) along the axis.  The plan keeps the
// same result type and element conversion; the order of the float additions
// becomes a tree (with compensation for sums).
#include <algorithm>
#include <cstring>

#include "scan.hpp"

namespace xg {

namespace {

int scan_rdt(int op, int acc_dtype, int edt) {
    int it = acc_dtype >= 0 ? acc_dtype : edt;
    if (it < 0 || it >= XG_NUM_DTYPES || it == XG_BFLOAT16) raise(XG_INVALID_ARGUMENT, "bad accumulator dtype");
    int at[2] = {it, edt};
    return result_dtype(op == XG_SCAN_CUMSUM ? XG_OP_PLUS : XG_OP_MULTIPLIES, 2, at, 0);
}

struct SDim {
    int64_t n;
    int64_t out;
    int64_t leaf[kMaxLeaves];
};

std::vector<SDim> merge(const std::vector<SDim>& d, int nl) {
    std::vector<SDim> m;
    for (const auto& x : d) {
        if (!m.empty()) {
            SDim& o = m.back();
            bool ok = o.out == x.out * x.n;
            for (int l = 0; l < nl && ok; ++l) ok = o.leaf[l] == x.leaf[l] * x.n;
            if (ok) {
                o.n *= x.n;
                o.out = x.out;
                for (int l = 0; l < nl; ++l) o.leaf[l] = x.leaf[l];
                continue;
            }
        }
        m.push_back(x);
    }
    return m;
}

}  // namespace

void scan_result(const xg_expr* e, const xg_scan_desc* s, int* dtype, Shape* shape) {
    if (!s) raise(XG_INVALID_ARGUMENT, "null scan descriptor");
    if (s->op != XG_SCAN_CUMSUM && s->op != XG_SCAN_CUMPROD) raise(XG_INVALID_ARGUMENT, "unknown scan op");
    ExprInfo info = analyse(e);
    *dtype = scan_rdt(s->op, s->acc_dtype, info.dtype);
    int nd = int(info.shape.size());
    if (s->axis == XG_AXIS_NONE) {
        *shape = Shape{numel(info.shape)};
        return;
    }
    int a = s->axis < 0 ? s->axis + nd : s->axis;
    if (a < 0 || a >= nd) raise(XG_BAD_AXES, "Axis larger than expression dimension in accumulator.");
    *shape = info.shape;
}

namespace {

int g_tile_blocks_per_cu = 8;  // workgroups per compute unit the tiled scans aim at
int g_tile_single_pass = 1;    // the one-pass tiled form may apply
int g_flat_single_pass = 0;    // few long rows: one-pass look-back (1) or three launches (0)

// element type the direct and tiled kernels read, per accumulator kind
int scan_in_dtype(int acc) {
    static const int t[AK_NUM] = {XG_FLOAT32, XG_FLOAT64, XG_FLOAT32, XG_FLOAT64,
                                  XG_INT32,   XG_UINT32,  XG_INT64,   XG_UINT64};
    return acc >= 0 && acc < AK_NUM ? t[acc] : -1;
}

// The tiled scans (scan_tiled.hpp) apply when the program is one load of a
// leaf of the accumulation input type, the leaf is contiguous along a dim
// other than the scan axis with 16-byte aligned rows, and the output is
// contiguous along that dim (ST_COLS) or along the scan axis (ST_TRANS);
// outputs of another element size or bf16 keep the other kernels.
void plan_tiled(ScanPlan* pp, int cus) {
    ScanPlan& p = *pp;
    p.tiled = ST_NONE;
    const int in_dt = scan_in_dtype(p.acc);
    if (in_dt < 0 || p.prog.nops != 1 || p.prog.ops[0].op != MO_LOAD || p.prog.nleaves != 1 || p.prog.leaf_gather[0])
        return;
    const LeafView& l = p.leaves[0];
    const int esz = dtype_size(in_dt);
    const int64_t V = 16 / esz;
    if (l.dtype != in_dt || p.out_dtype == XG_BFLOAT16 || dtype_size(p.out_dtype) != esz) return;
    if (reinterpret_cast<uintptr_t>(l.ptr) % 16 || reinterpret_cast<uintptr_t>(p.out_ptr) % 16) return;
    const bool native = p.out_dtype == in_dt && p.rdt == in_dt;
    TileScan t{};
    int kind = ST_NONE;
    const int R = p.rank;
    if (p.layout == SL_COLS && (R == 2 || R == 3)) {
        if (l.stride[R - 1] != 1 || p.out_stride[R - 1] != 1) return;
        t.outer = R == 3 ? uint32_t(p.shape[0]) : 1;
        t.len = uint32_t(p.shape[R - 2]);
        t.inner = uint32_t(p.shape[R - 1]);
        t.in_s_outer = R == 3 ? l.stride[0] : 0;
        t.in_s_len = l.stride[R - 2];
        t.out_s_outer = R == 3 ? p.out_stride[0] : 0;
        t.out_s_len = p.out_stride[R - 2];
        t.out_s_inner = 1;
        if (t.out_s_outer % V || t.out_s_len % V) return;
        kind = ST_COLS;
    } else if (p.layout == SL_ROWS && (R == 2 || R == 3) && native) {
        if (l.stride[R - 2] != 1 || l.stride[R - 1] == 1 || p.out_stride[R - 1] != 1) return;
        t.outer = R == 3 ? uint32_t(p.shape[0]) : 1;
        t.inner = uint32_t(p.shape[R - 2]);
        t.len = uint32_t(p.shape[R - 1]);
        t.in_s_outer = R == 3 ? l.stride[0] : 0;
        t.in_s_len = l.stride[R - 1];
        t.out_s_outer = R == 3 ? p.out_stride[0] : 0;
        t.out_s_len = 1;
        t.out_s_inner = p.out_stride[R - 2];
        if (t.out_s_outer % V || t.out_s_inner % V) return;
        kind = ST_TRANS;
    } else {
        return;
    }
    if (t.in_s_outer % V || t.in_s_len % V || t.in_s_len < 0 || t.in_s_outer < 0) return;
    const uint32_t sub = 64u;  // rows of one sub-tile (scan_tiles_kernel)
    const uint64_t cols = uint64_t(64 * V);
    t.ntiles = uint32_t((t.inner + cols - 1) / cols);
    t.ncols = uint32_t(t.ntiles * cols);
    // about four workgroups per compute unit, chunks of whole sub-tiles
    const uint64_t per_chunk = uint64_t(t.outer) * t.ntiles;
    const uint64_t target = uint64_t(std::max(cus, 1)) * uint64_t(g_tile_blocks_per_cu);
    uint64_t want = (target + per_chunk - 1) / per_chunk;
    const uint64_t max_chunks = std::min<uint64_t>(256, (uint64_t(t.len) + sub - 1) / sub);
    want = std::max<uint64_t>(1, std::min(want, max_chunks));
    const uint64_t rows = (uint64_t(t.len) + want - 1) / want;
    t.chunk_rows = uint32_t((rows + sub - 1) / sub * sub);
    t.nchunks = uint32_t((uint64_t(t.len) + t.chunk_rows - 1) / t.chunk_rows);
    if (per_chunk * t.nchunks >= (uint64_t(1) << 31)) return;
    // One pass with a look-back per column when there are at least two
    // chunks.  Measured on MI355X faster than the three launches from one
    // column tile up (float32 cumsum along axis 0: [131072, 256] 99 against
    // 124 us, [32768, 1024] 71 against 88 us, [8192, 4096] 67 against 81 us).
    // 32-bit accumulators publish packed with their flag, 64-bit ones as
    // value, wait, flag per (chunk, tile).
    const bool packed = p.acc == AK_F32C || p.acc == AK_F32 || p.acc == AK_I32 || p.acc == AK_U32;
    const uint64_t chunks1 = (uint64_t(t.len) + kTileLbRows - 1) / kTileLbRows;
    if (g_tile_single_pass && chunks1 >= 2 && per_chunk * chunks1 < (uint64_t(1) << 31)) {
        t.packed = packed ? 1 : 0;
        t.single = 1;
        t.chunk_rows = kTileLbRows;
        t.nchunks = uint32_t(chunks1);
        t.classes = 1;  // the largest power of two up to kTileMaxClasses dividing the chains
        while (t.classes < kTileMaxClasses && per_chunk % (2 * t.classes) == 0) t.classes *= 2;
    }
    t.in = l.ptr;
    t.out = p.out_ptr;
    t.identity = p.identity;
    t.out_dtype = p.out_dtype;
    t.rdt = p.rdt;
    t.nt_in = lm_nt(l.mode) ? 1 : 0;
    t.nt_out = lm_nt(p.out_mode) ? 1 : 0;
    p.tile = t;
    p.tile_native = native ? 1 : 0;
    p.tiled = kind;
}

// Rows too few to fill the GPU with a workgroup each, and longer than one
// chunk, run the look-back kernel: chunks of 4 waves x kLbSegments segments
// x 64 lanes x 16 bytes (16384 float32 elements).
void plan_lookback(ScanPlan* pp, int cus) {
    ScanPlan& p = *pp;
    p.lookback = 0;
    if (!scan_direct_ok(p)) return;
    const int in_dt = scan_in_dtype(p.acc);
    const uint32_t tile = 4u * uint32_t(kLbSegments) * 64u * uint32_t(16 / dtype_size(in_dt));
    if (p.len <= tile || uint64_t(p.rows) * 2 >= uint64_t(std::max(cus, 1)) * 4) return;
    const uint64_t cpr = (uint64_t(p.len) + tile - 1) / tile;
    if (cpr * p.rows >= (uint64_t(1) << 31)) return;
    LbScan& l = p.lb;
    l = LbScan{};
    l.in = p.leaves[0].ptr;
    l.out = p.out_ptr;
    l.rows = p.rows;
    l.len = p.len;
    l.cpr = uint32_t(cpr);
    l.total = uint32_t(cpr * p.rows);
    l.tile = tile;
    l.in_row_stride = p.rank == 2 ? p.leaves[0].stride[0] : 0;
    l.out_row_stride = p.rank == 2 ? p.out_stride[0] : 0;
    l.identity = p.identity;
    l.out_dtype = p.out_dtype;
    l.rdt = p.rdt;
    l.nt_in = lm_nt(p.leaves[0].mode) ? 1 : 0;
    l.nt_out = lm_nt(p.out_mode) ? 1 : 0;
    p.lb_packed = (p.acc == AK_F32C || p.acc == AK_F32 || p.acc == AK_I32 || p.acc == AK_U32) ? 1 : 0;
    p.lb_three = g_flat_single_pass ? 0 : 1;
    p.lookback = 1;
}

size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

}  // namespace

void set_scan_blocks_per_cu(int n) { g_tile_blocks_per_cu = n > 0 ? n : 8; }
void set_scan_single_pass(int on) { g_tile_single_pass = on ? 1 : 0; }
void set_scan_flat_single_pass(int on) { g_flat_single_pass = on ? 1 : 0; }

bool scan_direct_ok(const ScanPlan& p) {
    // rows are addressed with 64-bit offsets and elements within a row with
    // 32-bit ones (a row is shorter than 2^31), so 64-bit plans qualify too
    if (p.layout != SL_ROWS || p.rank < 1 || p.rank > 2 || p.acc < 0 || p.acc >= AK_NUM) return false;
    if (p.rank == 1 && p.rows != 1) return false;
    if (p.prog.nops != 1 || p.prog.ops[0].op != MO_LOAD || p.prog.nleaves != 1 || p.prog.leaf_gather[0]) return false;
    const LeafView& l = p.leaves[0];
    const int esz = dtype_size(l.dtype);
    if (l.dtype != scan_in_dtype(p.acc) || lm_base(l.mode) != LM_CONTIG || l.stride[p.rank - 1] != 1) return false;
    if (p.out_dtype == XG_BFLOAT16 || dtype_size(p.out_dtype) != esz || p.out_stride[p.rank - 1] != 1) return false;
    if (reinterpret_cast<uintptr_t>(l.ptr) % 16 || reinterpret_cast<uintptr_t>(p.out_ptr) % 16) return false;
    if (p.rank == 2 && ((l.stride[0] * esz) % 16 || (p.out_stride[0] * esz) % 16)) return false;
    return true;
}

size_t scan_workspace_bytes(const ScanPlan& p) {
    if (p.rows == 0) return 0;
    if (p.lookback) {
        const size_t n = p.lb.total;
        if (p.lb_three) return align256(n * sizeof(Store));
        if (p.lb_packed) return 256 + align256(n * 8);
        return 256 + align256(n * 4) + 2 * align256(n * sizeof(Store));
    }
    if (p.tiled == ST_NONE || p.tile.nchunks <= 1) return 0;
    if (p.tile.single) {
        const size_t n = size_t(p.tile.outer) * p.tile.nchunks * p.tile.ncols;
        if (p.tile.packed) return kTileCtrlBytes + align256(n * 8);
        const size_t nf = size_t(p.tile.outer) * p.tile.nchunks * p.tile.ntiles;
        return kTileCtrlBytes + align256(nf * 4) + 2 * align256(n * sizeof(Store));
    }
    return size_t(p.tile.outer) * p.tile.nchunks * p.tile.ncols * sizeof(Store);
}

ScanPlan plan_scan(const xg_expr* e, const xg_scan_desc* s, const xg_tensor* out, int V, int cus) {
    if (!out || out->ndim < 0 || out->ndim > XG_MAX_RANK || out->dtype < 0 || out->dtype >= XG_NUM_DTYPES)
        raise(XG_INVALID_ARGUMENT, "malformed output tensor");
    ExprInfo info = analyse(e);
    int nd = int(info.shape.size());
    int axis = s->axis < 0 ? s->axis + nd : s->axis;
    if (axis < 0 || axis >= nd) raise(XG_BAD_AXES, "Axis larger than expression dimension in accumulator.");
    Shape os(out->shape, out->shape + out->ndim);
    if (os != info.shape) raise(XG_BROADCAST_ERROR, "scan output has shape " + shape_str(os) + ", expected " + shape_str(info.shape));
    ScanPlan p{};
    p.rdt = scan_rdt(s->op, s->acc_dtype, info.dtype);
    p.op = s->op == XG_SCAN_CUMSUM ? RO_ADD : RO_MUL;
    p.out_dtype = out->dtype;
    int r = p.rdt;
    if (r == XG_FLOAT32) p.acc = p.op == RO_ADD ? AK_F32C : AK_F32;
    else if (r == XG_FLOAT64) p.acc = p.op == RO_ADD ? AK_F64C : AK_F64;
    else if (r == XG_INT64) p.acc = AK_I64;
    else if (r == XG_UINT64) p.acc = AK_U64;
    else if (r == XG_UINT32) p.acc = AK_U32;
    else p.acc = AK_I32;
    p.prog = lower(e, info, r);
    uint64_t one = 0;
    if (p.op == RO_MUL) {
        if (r == XG_FLOAT32) { float f = 1.f; uint32_t b; std::memcpy(&b, &f, 4); one = b; }
        else if (r == XG_FLOAT64) { double d = 1.0; std::memcpy(&one, &d, 8); }
        else one = 1;
    }
    p.identity = one;
    if (numel(info.shape) == 0) { p.rows = 0; return p; }

    auto ls = leaf_strides(e, info.shape);
    int nl = p.prog.nleaves;
    SDim ad{};
    std::vector<SDim> before, after;
    for (int k = 0; k < nd; ++k) {
        SDim d{};
        d.n = info.shape[size_t(k)];
        d.out = out->shape[k] == 1 ? 0 : out->strides[k];
        for (int pl = 0; pl < nl; ++pl) d.leaf[pl] = p.prog.leaf_gather[pl] ? 0 : ls[size_t(p.prog.leaf_src[pl])][size_t(k)];
        if (k == axis) ad = d;
        else if (d.n != 1) (k < axis ? before : after).push_back(d);
    }
    if (ad.out < 0) raise(XG_UNSUPPORTED, "scan into a reversed output view");
    bool rows = ad.out == 1 || ad.n == 1 || after.empty();
    if (!rows) {
        after = merge(after, nl);
        if (after.back().out != 1) rows = true;  // no contiguous column dim: scan row by row
    }
    std::vector<SDim> order;
    if (rows) {
        p.layout = SL_ROWS;
        std::vector<SDim> outer = before;
        outer.insert(outer.end(), after.begin(), after.end());
        order = merge(outer, nl);
        order.push_back(ad);
    } else {
        p.layout = SL_COLS;
        std::vector<SDim> outer = merge(before, nl);
        for (size_t k = 0; k + 1 < after.size(); ++k) outer.push_back(after[k]);
        order = outer;
        order.push_back(ad);
        order.push_back(after.back());
    }
    p.rank = int(order.size());
    if (p.rank > kMaxRank) raise(XG_UNSUPPORTED, "scan rank too large");
    int64_t max_off = 0;
    for (int k = 0; k < p.rank; ++k) {
        if (order[size_t(k)].n >= (int64_t(1) << 31)) raise(XG_UNSUPPORTED, "dimension too large for 32-bit indexing");
        p.shape[k] = order[size_t(k)].n;
        p.div[k] = make_fastdiv(uint32_t(order[size_t(k)].n));
        p.out_stride[k] = order[size_t(k)].out;
        max_off += (order[size_t(k)].n - 1) * std::abs(order[size_t(k)].out);
    }
    int osz = dtype_size(out->dtype);
    p.out_ptr = static_cast<char*>(out->data) + out->offset * osz;
    uint64_t outer = 1;
    if (p.layout == SL_ROWS) {
        for (int k = 0; k + 1 < p.rank; ++k) outer *= uint64_t(p.shape[k]);
        p.len = uint32_t(p.shape[p.rank - 1]);
        p.rows = uint32_t(outer);
    } else {
        for (int k = 0; k + 2 < p.rank; ++k) outer *= uint64_t(p.shape[k]);
        p.len = uint32_t(p.shape[p.rank - 2]);
        p.inner = uint32_t(p.shape[p.rank - 1]);
        p.nchunks_in = (p.inner + uint32_t(V) - 1) / uint32_t(V);
        p.div_nchunks_in = make_fastdiv(p.nchunks_in);
        p.rows = uint32_t(outer * p.nchunks_in);
    }
    if (outer >= (uint64_t(1) << 31)) raise(XG_UNSUPPORTED, "scan too large for 32-bit indexing");
    // out vector stores along the last dim when contiguous and aligned
    bool oal = p.out_stride[p.rank - 1] == 1 && reinterpret_cast<uintptr_t>(p.out_ptr) % uintptr_t(osz * V) == 0;
    for (int k = 0; k + 1 < p.rank; ++k)
        if (p.out_stride[k] % V != 0) oal = false;
    p.out_mode = oal ? LM_CONTIG : LM_STRIDED;
    {
        int64_t total = 1;
        for (int k = 0; k < p.rank; ++k) total *= std::max<int64_t>(p.shape[k], 0);
        if (total * osz >= kNtMinBytes) p.out_mode |= LM_NT;
    }
    for (int pl = 0; pl < nl; ++pl) {
        LeafView& lv = p.leaves[pl];
        const xg_tensor& t = e->leaves[p.prog.leaf_src[pl]];
        if (p.prog.leaf_gather[pl]) {
            set_gather_leaf(t, &lv);
            continue;
        }
        int esz = dtype_size(t.dtype);
        lv.ptr = static_cast<const char*>(t.data) + t.offset * esz;
        lv.dtype = t.dtype;
        bool al = reinterpret_cast<uintptr_t>(lv.ptr) % uintptr_t(esz * V) == 0;
        int64_t lmax = 0;
        for (int k = 0; k < p.rank; ++k) {
            lv.stride[k] = order[size_t(k)].leaf[pl];
            lmax += (order[size_t(k)].n - 1) * std::abs(lv.stride[k]);
            if (k + 1 < p.rank && lv.stride[k] % V != 0) al = false;
        }
        max_off = std::max(max_off, lmax);
        int64_t is = lv.stride[p.rank - 1];
        lv.mode = is == 0 ? LM_BCAST : (is == 1 && al) ? LM_CONTIG : LM_STRIDED;
        if (lv.mode != LM_BCAST && streams(lv.stride, p.shape, p.rank, esz)) lv.mode |= LM_NT;
    }
    p.index64 = max_off >= (int64_t(1) << 31) ? 1 : 0;
    plan_tiled(&p, cus);
    if (p.tiled == ST_NONE) plan_lookback(&p, cus);
    return p;
}

}  // namespace xg
