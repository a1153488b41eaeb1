// scan_kernels.hpp — scan kernel templates (see scan.hip for the design).
//
// Replaces accumulator_impl's copy-then-sequential recurrence
// (reducers/xaccumulator.hpp:215-297).  Rows layout (scan axis contiguous
// in the output): a workgroup owns a row and walks it in tiles of
// 256 x V elements; each thread scans its V elements, a wave scan (shuffles)
// and an LDS pass over the four wave totals give every thread its exclusive
// prefix, and the tile total carries into the next tile.  Cols layout (scan
// axis strided): a thread owns V adjacent columns and walks the axis.
// Float sums use the compensated accumulator of the reductions.
#pragma once

#include <hip/hip_runtime.h>

#include "acc.hpp"
#include "chunk.hpp"
#include "scan.hpp"

namespace xg {

template <int V, bool I64>
__device__ __forceinline__ void scan_store(const ScanPlan& p, const uint32_t* idx, uint32_t j0, int nvalid,
                                           const Reg (&val)[V]) {
    const int odt = p.out_dtype;
    const int esz = dtype_size(odt);
    auto off = unit_offset<I64>(p.out_stride, idx, p.rank, j0);
    char* ptr = p.out_ptr + int64_t(off) * esz;
    Reg o[V];
#pragma unroll
    for (int v = 0; v < V; ++v) o[v] = odt == XG_BFLOAT16 ? cvt(val[v], p.rdt, XG_FLOAT32) : cvt(val[v], p.rdt, odt);
    if (lm_base(p.out_mode) == LM_CONTIG && nvalid == V) {
        store_vec<V>(ptr, esz, odt, o);
    } else {
        const int64_t os = p.out_stride[p.rank - 1];
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (v < nvalid) store_raw(ptr + int64_t(v) * os * esz, esz, to_storage(o[v], odt));
    }
}

template <int V, int NR, class S, bool I64, int AK, int OP>
__global__ __launch_bounds__(256) void scan_rows_kernel(const ScanPlan p_arg) {
    const ScanPlan& p = kernel_plan<S>(p_arg);
    using A = Acc<AK, OP>;
    __shared__ Store wtot[4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    for (uint32_t row = blockIdx.x; row < p.rows; row += gridDim.x) {
        uint32_t idx[kMaxRank];
#pragma unroll
        for (int k = 0; k < kMaxRank; ++k) idx[k] = 0;
        uint32_t q = row;
        for (int k = p.rank - 2; k >= 0; --k) {
            uint32_t d = fastdiv(q, p.div[k]);
            idx[k] = q - d * uint32_t(p.shape[k]);
            q = d;
        }
        A carry;
        carry.init(p.identity);
        for (uint32_t base = 0; base < p.len; base += 256u * V) {
            const uint32_t j0 = base + threadIdx.x * V;
            const int nvalid = j0 < p.len ? min(V, int(p.len - j0)) : 0;
            Reg vals[V];
            if (nvalid > 0) eval_chunk<V, NR, S, I64>(p.leaves, p.prog, p.rank, idx, j0, nvalid, vals);
            // thread-local inclusive scan
            A run[V];
            A t;
            t.init(p.identity);
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if (v < nvalid) t.add(vals[v]);
                run[v] = t;
            }
            // wave inclusive scan of thread totals (earlier values merge first)
            A x = t;
            for (int d = 1; d < 64; d <<= 1) {
                A y = x.shfl_up(d);
                if (lane >= d) {
                    y.merge(x);
                    x = y;
                }
            }
            if (lane == 63) wtot[wave] = x.store();
            A ex = x.shfl_up(1);
            if (lane == 0) ex.init(p.identity);
            __syncthreads();
            A pre = carry;
            for (int w = 0; w < wave; ++w) {
                A o;
                o.load(wtot[w]);
                pre.merge(o);
            }
            pre.merge(ex);
            Reg outv[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                A o = pre;
                o.merge(run[v]);
                outv[v] = o.value(p.rdt);
            }
            if (nvalid > 0) scan_store<V, I64>(p, idx, j0, nvalid, outv);
            for (int w = 0; w < 4; ++w) {
                A o;
                o.load(wtot[w]);
                carry.merge(o);
            }
            __syncthreads();
        }
    }
}

template <int V, int NR, class S, bool I64, int AK, int OP>
__global__ __launch_bounds__(256) void scan_cols_kernel(const ScanPlan p_arg) {
    const ScanPlan& p = kernel_plan<S>(p_arg);
    using A = Acc<AK, OP>;
    for (uint32_t item = blockIdx.x * 256 + threadIdx.x; item < p.rows; item += gridDim.x * 256) {
        uint32_t oo = fastdiv(item, p.div_nchunks_in);
        const uint32_t j0 = (item - oo * p.nchunks_in) * V;
        const int nvalid = min(V, int(p.inner - j0));
        uint32_t idx[kMaxRank];
#pragma unroll
        for (int k = 0; k < kMaxRank; ++k) idx[k] = 0;
        for (int k = p.rank - 3; k >= 0; --k) {
            uint32_t d = fastdiv(oo, p.div[k]);
            idx[k] = oo - d * uint32_t(p.shape[k]);
            oo = d;
        }
        A acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v].init(p.identity);
        for (uint32_t i = 0; i < p.len; ++i) {
            idx[p.rank - 2] = i;
            Reg vals[V];
            eval_chunk<V, NR, S, I64>(p.leaves, p.prog, p.rank, idx, j0, nvalid, vals);
            Reg outv[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                acc[v].add(vals[v]);
                outv[v] = acc[v].value(p.rdt);
            }
            scan_store<V, I64>(p, idx, j0, nvalid, outv);
        }
    }
}

template <class T> struct ScanInDtype;
template <> struct ScanInDtype<float> { static constexpr int value = XG_FLOAT32; };
template <> struct ScanInDtype<double> { static constexpr int value = XG_FLOAT64; };
template <> struct ScanInDtype<int32_t> { static constexpr int value = XG_INT32; };
template <> struct ScanInDtype<uint32_t> { static constexpr int value = XG_UINT32; };
template <> struct ScanInDtype<int64_t> { static constexpr int value = XG_INT64; };
template <> struct ScanInDtype<uint64_t> { static constexpr int value = XG_UINT64; };

// element type the direct kernel reads for an accumulator kind
template <int AK> struct ScanIn;
template <> struct ScanIn<AK_F32C> { using T = float; };
template <> struct ScanIn<AK_F64C> { using T = double; };
template <> struct ScanIn<AK_F32> { using T = float; };
template <> struct ScanIn<AK_F64> { using T = double; };
template <> struct ScanIn<AK_I32> { using T = int32_t; };
template <> struct ScanIn<AK_U32> { using T = uint32_t; };
template <> struct ScanIn<AK_I64> { using T = int64_t; };
template <> struct ScanIn<AK_U64> { using T = uint64_t; };

// Cross-lane moves of an accumulator through DPP (the data moves between
// lanes inside a VALU instruction, no LDS permute), one 32-bit word at a
// time.  Lanes the pattern gives no source, or that ROW_MASK leaves out,
// take `old`.  dpp_ctrl codes of GFX9 DPP: row_shr:n = 0x110 + n,
// wave_shr:1 = 0x138, row_bcast:15 = 0x142, row_bcast:31 = 0x143.
template <int CTRL, int ROW_MASK, class A>
__device__ __forceinline__ A dpp_move(const A& src, const A& old) {
    static_assert(sizeof(A) % 4 == 0, "accumulators are whole 32-bit words");
    constexpr int N = int(sizeof(A) / 4);
    int s[N], o[N], r[N];
    __builtin_memcpy(s, &src, sizeof(A));
    __builtin_memcpy(o, &old, sizeof(A));
#pragma unroll
    for (int i = 0; i < N; ++i) r[i] = __builtin_amdgcn_update_dpp(o[i], s[i], CTRL, ROW_MASK, 0xf, false);
    A out;
    __builtin_memcpy(&out, r, sizeof(A));
    return out;
}

// the accumulator of lane `L`, in every lane
template <int L, class A>
__device__ __forceinline__ A lane_value(const A& src) {
    constexpr int N = int(sizeof(A) / 4);
    int s[N];
    __builtin_memcpy(s, &src, sizeof(A));
#pragma unroll
    for (int i = 0; i < N; ++i) s[i] = __builtin_amdgcn_readlane(s[i], L);
    A out;
    __builtin_memcpy(&out, s, sizeof(A));
    return out;
}

// Inclusive scan of one accumulator per lane over the wave, earlier lanes
// merged first: shifts of 1, 2, 4, 8 inside each row of 16 lanes, then row
// 0's total into row 1 and row 2's into row 3, then the total of rows 0-1
// into rows 2 and 3.  `id` is the identity.
template <class A>
__device__ __forceinline__ A wave_inclusive_scan(A x, const A& id) {
    { A y = dpp_move<0x111, 0xf>(x, id); y.merge(x); x = y; }
    { A y = dpp_move<0x112, 0xf>(x, id); y.merge(x); x = y; }
    { A y = dpp_move<0x114, 0xf>(x, id); y.merge(x); x = y; }
    { A y = dpp_move<0x118, 0xf>(x, id); y.merge(x); x = y; }
    { A y = dpp_move<0x142, 0xa>(x, id); y.merge(x); x = y; }
    { A y = dpp_move<0x143, 0xc>(x, id); y.merge(x); x = y; }
    return x;
}

// Direct rows scan for a program that only loads its one leaf (already of
// the accumulation input type), contiguous and 16-byte aligned along the
// axis, with a contiguous aligned output (rank 2, [rows | axis]).  W waves
// share a row: W = 1 gives one row per wave (four rows per block), for rows
// that fit one wave's tile; W = 4 gives one row per block.  A tile of the
// row is W * U segments of 64 * V elements (16 bytes per lane, so a wave
// instruction covers 1 KiB); wave w owns segments [w*U, (w+1)*U) and loads
// them at once.  Per segment each lane sums its V elements, a wave scan
// (DPP moves, earlier lanes first) gives the exclusive prefix and the
// segment total carries on inside the wave; the W wave totals meet in LDS,
// so every wave knows the tile's carry plus the waves before it, and each
// lane then runs its V outputs up from that prefix.  The merge order
// depends on the row length only.  NATIVE: the result and the output are
// of the input's type (the usual cumsum / cumprod), fixed at compile time
// so that no per-element type conversion is dispatched at run time.
// One row of `len` elements scanned from `carry` in tiles of W waves x U
// segments (the body of the direct rows kernel; the chunked reduce-then-scan
// runs it on each chunk from the chunk's carry).  `wr` = this wave's place in
// a tile, `wtot` = W slots of LDS.
template <int AK, int OP, bool NATIVE, int W>
__device__ __forceinline__ void scan_row_tiles(const typename ScanIn<AK>::T* in, char* out, uint32_t len,
                                               Acc<AK, OP> carry, const Acc<AK, OP>& id, bool nt_in, bool nt_out,
                                               int odt, int rdt, int wr, Store* wtot) {
    using A = Acc<AK, OP>;
    using T = typename ScanIn<AK>::T;
    constexpr int V = int(16 / sizeof(T));
    constexpr int U = 4;
    constexpr uint32_t SEG = 64u * V;
    constexpr uint32_t TILE = SEG * U * W;
    const int lane = threadIdx.x & 63;
    const int wave = wr;
    const int oesz = dtype_size(odt);
    for (uint32_t base = 0; base < len; base += TILE) {
        const uint32_t wbase = base + uint32_t(wr) * SEG * U;
        T x[U][V];
        int nv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = wbase + uint32_t(u) * SEG + uint32_t(lane * V);
            nv[u] = j < len ? min(V, int(len - j)) : 0;
            if (nv[u] == V) {
                Reg r[V];
                if (nt_in) load_vec<V, true>(reinterpret_cast<const char*>(in + j), int(sizeof(T)), ScanInDtype<T>::value, r);
                else load_vec<V>(reinterpret_cast<const char*>(in + j), int(sizeof(T)), ScanInDtype<T>::value, r);
#pragma unroll
                for (int v = 0; v < V; ++v) x[u][v] = reg_get<T>(r[v]);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) x[u][v] = v < nv[u] ? in[j + uint32_t(v)] : T(0);
            }
        }
        // wave-local exclusive prefixes of the segments
        A local = id;
        A pre[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            pre[u] = local;
            if (wbase + uint32_t(u) * SEG >= len) continue;  // uniform in the wave
            A t = id;
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (v < nv[u]) t.add(reg_put<T>(x[u][v]));
            const A incl = wave_inclusive_scan(t, id);
            pre[u].merge(dpp_move<0x138, 0xf>(incl, id));  // wave_shr:1, lane 0 takes the identity
            local.merge(lane_value<63>(incl));
        }
        A before = carry;  // the tile's carry and the waves before this one
        if constexpr (W > 1) {
            if (lane == 0) wtot[wave] = local.store();
            __syncthreads();
#pragma unroll
            for (int w = 0; w < W; ++w) {
                if (w == wave) before = carry;
                A o;
                o.load(wtot[w]);
                carry.merge(o);
            }
            __syncthreads();  // wtot is written again for the next tile
        } else {
            carry.merge(local);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (wbase + uint32_t(u) * SEG >= len) continue;  // uniform in the wave
            A t = before;
            t.merge(pre[u]);
            Reg o[V];
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if (v < nv[u]) t.add(reg_put<T>(x[u][v]));
                const Reg val = t.value(rdt);
                if constexpr (NATIVE) o[v] = val;
                else o[v] = odt == XG_BFLOAT16 ? cvt(val, rdt, XG_FLOAT32) : cvt(val, rdt, odt);
            }
            const uint32_t j = wbase + uint32_t(u) * SEG + uint32_t(lane * V);
            char* ptr = out + int64_t(j) * oesz;
            if (nv[u] == V) {
                store_vec_rt(nt_out, ptr, oesz, odt, o);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v)
                    if (v < nv[u]) store_raw(ptr + int64_t(v) * oesz, oesz, to_storage(o[v], odt));
            }
        }
    }
}

template <int AK, int OP, bool NATIVE, int W>
__global__ __launch_bounds__(256) void scan_rows_direct_kernel(const ScanPlan p) {
    using A = Acc<AK, OP>;
    using T = typename ScanIn<AK>::T;
    __shared__ Store wtot[4];
    const int wave = threadIdx.x >> 6;
    const uint32_t row = W == 1 ? blockIdx.x * 4u + uint32_t(wave) : blockIdx.x;
    const int wr = W == 1 ? 0 : wave;  // this wave's place in a tile
    if (row >= p.rows) return;  // uniform in a wave (W = 1) or in the block (W = 4)
    const T* in = reinterpret_cast<const T*>(p.leaves[0].ptr) + int64_t(row) * p.leaves[0].stride[0];
    const bool nt_in = lm_nt(p.leaves[0].mode), nt_out = lm_nt(p.out_mode);
    const int odt = NATIVE ? int(ScanInDtype<T>::value) : p.out_dtype;
    const int rdt = NATIVE ? int(ScanInDtype<T>::value) : p.rdt;
    const int oesz = dtype_size(odt);
    char* out = p.out_ptr + int64_t(row) * p.out_stride[0] * oesz;
    A id;
    id.init(p.identity);
    scan_row_tiles<AK, OP, NATIVE, W>(in, out, p.len, id, id, nt_in, nt_out, odt, rdt, wr, wtot);
}

}  // namespace xg

#include "scan_tiled.hpp"
#include "scan_lookback.hpp"
#include "scan_tiled_1p.hpp"

namespace xg {

// elements of one wave's tile in the direct rows scan (rows up to this
// length take one wave each)
template <class T> constexpr uint32_t scan_wave_tile() { return 64u * uint32_t(16 / sizeof(T)) * 4u; }

using SFn = void (*)(const ScanPlan);
using TFn = void (*)(const TileScan);
using LFn = void (*)(const LbScan);

// the one-pass tiled kernel (the transposed form stores the input's type only)
template <int AK, int OP, bool NATIVE, bool TRANS> constexpr TFn tile_1p_fn() {
    if constexpr (NATIVE || !TRANS) return scan_tiles_1p_kernel<AK, OP, NATIVE, TRANS>;
    else return nullptr;
}

struct ScanKernels {
    SFn rows32, rows64, cols32, cols64;
    SFn direct[2][2];   // [native][waves per row: 1, 4]
    uint32_t wave_tile;  // elements of one wave's tile in the direct kernels
    TFn tile_totals, tile_carry;
    TFn tiles[2][2];     // [native][transposed output]; the transposed one is native only
    TFn tiles1p[2][2];   // the one-pass form, likewise
    LFn lookback[2][2];  // [kLbOnePass / kLbTotals][native]
    LFn lb_carry;
    LFn lb_apply[2][2];  // [native][carries summed in the kernel]
};

template <int AK, int OP> ScanKernels scan_kernels() {
    return ScanKernels{scan_rows_kernel<4, kMaxRegs, DynSpec, false, AK, OP>,
                       scan_rows_kernel<4, kMaxRegs, DynSpec, true, AK, OP>,
                       scan_cols_kernel<4, kMaxRegs, DynSpec, false, AK, OP>,
                       scan_cols_kernel<4, kMaxRegs, DynSpec, true, AK, OP>,
                       {{scan_rows_direct_kernel<AK, OP, false, 1>, scan_rows_direct_kernel<AK, OP, false, 4>},
                        {scan_rows_direct_kernel<AK, OP, true, 1>, scan_rows_direct_kernel<AK, OP, true, 4>}},
                       scan_wave_tile<typename ScanIn<AK>::T>(),
                       scan_tiles_totals_kernel<AK, OP>,
                       scan_tiles_carry_kernel<AK, OP>,
                       {{scan_tiles_kernel<AK, OP, false, false>, nullptr},
                        {scan_tiles_kernel<AK, OP, true, false>, scan_tiles_kernel<AK, OP, true, true>}},
                       {{tile_1p_fn<AK, OP, false, false>(), nullptr},
                        {tile_1p_fn<AK, OP, true, false>(), tile_1p_fn<AK, OP, true, true>()}},
                       {{scan_lookback_kernel<AK, OP, false, kLbOnePass>, scan_lookback_kernel<AK, OP, true, kLbOnePass>},
                        {scan_lookback_kernel<AK, OP, false, kLbTotals>, scan_lookback_kernel<AK, OP, true, kLbTotals>}},
                       scan_lookback_carry_kernel<AK, OP>,
                       {{scan_chunks_apply_kernel<AK, OP, false, false>, scan_chunks_apply_kernel<AK, OP, false, true>},
                        {scan_chunks_apply_kernel<AK, OP, true, false>, scan_chunks_apply_kernel<AK, OP, true, true>}}};
}

bool scan_kernels_float(int ak, int op, ScanKernels* k);
bool scan_kernels_int(int ak, int op, ScanKernels* k);

}  // namespace xg
