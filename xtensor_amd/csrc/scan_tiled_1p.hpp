// scan_tiled_1p.hpp — scans along a strided axis in one pass, with a
// look-back across the chunks of the axis for every column.
//
// accumulator_impl (reducers/xaccumulator.hpp:240-297) runs the recurrence
// res[k] = res[k-1] + res[k] down the axis for every column.  The three
// launches of scan_tiled.hpp read the input twice; here a workgroup holds its
// whole chunk (64 rows of one tile of 64 lanes x 16 bytes: 256 bytes a lane)
// in registers, so the input is read once:
//   * a ticket from an atomic counter orders the workgroups (chunk-major), so
//     the chunk before this one in the same column tile was taken earlier and
//     is resident or finished when this one waits for it.  The column tiles
//     (chains) are split into `classes` groups with a counter each: the
//     workgroup with blockIdx b draws from counter b % classes and takes a
//     chunk of that group's chains, so the ordering holds inside each group
//     and one counter serves 1/classes of the workgroups (a single counter
//     for 2048 workgroups measured ~11 us of 67 on MI355X);
//   * the chunk's column totals meet in LDS and every thread publishes its
//     column's total as an aggregate;
//   * one wave looks back along column 0 over the 64 preceding chunks at
//     once and finds how far back the nearest published inclusive prefix
//     lies (chunk 0 always publishes one); then every thread merges its own
//     column's aggregates back to that prefix and publishes its column's
//     inclusive prefix;
//   * the chunk is scanned from those prefixes and stored, as in the scan
//     phase of the three-launch form.
// Publications of 32-bit accumulators (float32 sums with compensation,
// float32 products, 32-bit integers) are one 64-bit word per (chunk, column)
// holding the value and its flag (the packed form of scan_lookback.hpp),
// written and read with single device-scope atomic accesses: no fence.
// 64-bit accumulators publish values, wait for them, then one flag per
// (chunk, tile) (below).  Every wait is bounded (kSpinLimit polls; then the chunk counts as a
// prefix of the identity and the workspace counts the timeout: a wrong result,
// never a hang).  The look-back merges in an order that depends on timing, so
// float sums can differ in the last bit between runs (integers are exact).
//
// Included by scan_kernels.hpp, after scan_tiled.hpp and scan_lookback.hpp.
#pragma once

#include <hip/hip_runtime.h>

#include "acc.hpp"
#include "scan.hpp"

namespace xg {

constexpr int kTileLbBatch = 4;  // publications one thread loads at once when it follows the leader

// The leader: one whole wave polls column 0 of the tile only, its lanes the
// 64 preceding chunks at once (further windows while all 64 hold
// aggregates), and returns in every lane the depth d such that chunk c-1-d
// holds column 0's nearest published inclusive prefix and the chunks between
// hold its aggregates.  The other columns are published by the same
// workgroups at the same time, so polling one column keeps the waiting
// traffic to one 64-bit word per chunk.
__device__ __forceinline__ uint32_t tile_lead_depth(uint64_t* col0, uint32_t stride, uint32_t c) {
    const int lane = threadIdx.x & 63;
    int64_t nearest = int64_t(c) - 1;
    uint32_t depth = 0;
    while (true) {
        const int64_t j = nearest - lane;
        uint32_t f = j >= 0 ? 0u : kLbPrefix;  // before chunk 0: the identity, as a prefix
        int spins = 0;
        int lim = 64;
        while (true) {
            if (f == 0) f = uint32_t(lb_load(col0 + j * int64_t(stride))) & 3u;
            const uint64_t pm = __ballot(f == kLbPrefix);
            lim = pm ? __builtin_ctzll(pm) : 64;
            const uint64_t zm = __ballot(f == 0 && lane < lim);
            if (zm == 0) break;
            if (++spins >= kSpinLimit) {  // the followers' own bounded waits settle it
                lim = __builtin_ctzll(zm);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (lim < 64) return depth + uint32_t(lim);
        depth += 64;
        nearest -= 64;
    }
}

// A follower: the exclusive prefix of chunk c (> 0) of one column, given the
// leader's depth d.  It merges the column's aggregates of chunks c-1 ...
// back to the nearest inclusive prefix, which lies at depth d or nearer (the
// column's publications follow column 0's closely: a slot not yet published,
// or slot d still holding the aggregate, is polled again).  Bounded waits:
// after kSpinLimit polls the slot counts as a prefix of the identity and
// `timeouts` counts it.
template <int AK, int OP>
__device__ __forceinline__ Acc<AK, OP> tile_follow(uint64_t* col, uint32_t stride, uint32_t c, uint32_t d,
                                                   const Acc<AK, OP>& id, uint32_t* timeouts) {
    using A = Acc<AK, OP>;
    constexpr int B = kTileLbBatch;
    A excl = id;
    for (uint32_t m0 = 0; m0 <= d; m0 += B) {
        uint64_t w[B];
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int64_t j = int64_t(c) - 1 - int64_t(m0 + uint32_t(b));
            w[b] = (m0 + uint32_t(b) <= d && j >= 0) ? lb_load(col + j * int64_t(stride)) : 0;
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const uint32_t m = m0 + uint32_t(b);
            const int64_t j = int64_t(c) - 1 - int64_t(m);
            if (m > d || j < 0) return excl;  // before chunk 0: the identity
            uint32_t f = uint32_t(w[b]) & 3u;
            int spins = 0;
            while (f == 0 || (m == d && f != kLbPrefix)) {
                if (++spins >= kSpinLimit) {
                    atomicAdd(timeouts, 1u);
                    return excl;
                }
                __builtin_amdgcn_s_sleep(1);
                w[b] = lb_load(col + j * int64_t(stride));
                f = uint32_t(w[b]) & 3u;
            }
            A v = lb_unpack<AK, OP>(w[b]);
            v.merge(excl);  // this chunk lies before what was merged so far
            excl = v;
            if (f == kLbPrefix) return excl;
        }
    }
    return excl;
}

// 64-bit accumulators do not fit beside a flag in one word: each (chunk,
// tile) has one flag word, and the columns' totals and inclusive prefixes go
// to two arrays of their own.  The workgroup stores its columns' values with
// device-scope atomic stores, waits for them to complete, synchronises, and
// only then one thread sets the flag (value, wait, flag); a value array is
// written once per launch, so a reader that saw a flag reads complete values.
// The leader polls the flags, 64 preceding chunks at once, for the depth d of
// the nearest published prefix (as tile_lead_depth does on column 0's words);
// `*timed_out` tells the followers that slot d never published (it then
// counts as a prefix of the identity).
__device__ __forceinline__ uint32_t tile_lead_depth_flags(uint32_t* flags, uint32_t stride, uint32_t c,
                                                          bool* timed_out) {
    const int lane = threadIdx.x & 63;
    int64_t nearest = int64_t(c) - 1;
    uint32_t depth = 0;
    *timed_out = false;
    while (true) {
        const int64_t j = nearest - lane;
        uint32_t f = j >= 0 ? 0u : kLbPrefix;  // before chunk 0: the identity, as a prefix
        int spins = 0;
        int lim = 64;
        while (true) {
            if (f == 0) f = lb_load(flags + j * int64_t(stride));
            const uint64_t pm = __ballot(f == kLbPrefix);
            lim = pm ? __builtin_ctzll(pm) : 64;
            const uint64_t zm = __ballot(f == 0 && lane < lim);
            if (zm == 0) break;
            if (++spins >= kSpinLimit) {
                lim = __builtin_ctzll(zm);
                *timed_out = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (lim < 64) return depth + uint32_t(lim);
        depth += 64;
        nearest -= 64;
    }
}

// A follower of the flagged form: the merge of the aggregates of chunks
// c-1 ... c-d and the inclusive prefix of chunk c-1-d (the identity where
// that lies before chunk 0 or never published), all from complete values.
template <int AK, int OP>
__device__ __forceinline__ Acc<AK, OP> tile_follow_values(const Store* agg, const Store* pre, uint32_t stride,
                                                          uint32_t c, uint32_t d, bool timed_out,
                                                          const Acc<AK, OP>& id) {
    using A = Acc<AK, OP>;
    A excl = id;
    for (uint32_t m = 0; m <= d; ++m) {
        const int64_t j = int64_t(c) - 1 - int64_t(m);
        if (j < 0 || (m == d && timed_out)) break;  // the identity
        const Store* src = (m == d ? pre : agg) + j * int64_t(stride);
        Store st{lb_load(&src->a), lb_load(&src->b)};
        A v;
        v.load(st);
        v.merge(excl);  // this chunk lies before what was merged so far
        excl = v;
    }
    return excl;
}

// the column's values, then (after every thread's stores completed) the flag
template <int AK, int OP>
__device__ __forceinline__ void tile_publish_flagged(Store* dst, bool has_col, const Acc<AK, OP>& a, uint32_t* flag,
                                                     uint32_t f) {
    if (has_col) {
        const Store st = a.store();
        lb_store(&dst->a, st.a);
        lb_store(&dst->b, st.b);
    }
    __builtin_amdgcn_s_waitcnt(0);  // this thread's stores have completed
    __syncthreads();
    if (threadIdx.x == 0) lb_store(flag, f);
}

template <int AK, int OP, bool NATIVE, bool TRANS>
__global__ __launch_bounds__(256) void scan_tiles_1p_kernel(const TileScan p) {
    using A = Acc<AK, OP>;
    using T = typename ScanIn<AK>::T;
    constexpr int V = int(16 / sizeof(T));
    constexpr int COLS = 64 * V;
    constexpr bool PACKED = LbPacked<AK>::value;
    static_assert(COLS == 256 || !PACKED, "one column of the tile per thread in the packed look-back");
    constexpr int CH = int(kTileLbRows);
    constexpr int RW = CH / 4;                  // rows per wave in a sub-tile
    constexpr uint32_t SUB = 4u * RW;           // rows of a sub-tile
    constexpr int NS = int(kTileLbRows / SUB);  // sub-tiles of a chunk, all held in registers
    static_assert(NS == 1 && int(SUB) == CH, "a chunk is one sub-tile");
    constexpr int HC = COLS / 2;                // TRANS: columns of one half of the tile in LDS
    constexpr int PITCH = CH + 1;
    static_assert(!TRANS || NATIVE, "the transposed scan stores the input's type");
    union Shared {
        Store wtot[NS][4][COLS];
        T tile[TRANS ? HC : 1][TRANS ? PITCH : 1];
    };
    __shared__ Shared sm;
    __shared__ Store s_excl[COLS];
    __shared__ uint32_t s_ticket, s_depth, s_timed_out;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint32_t cls = blockIdx.x % p.classes;
    if (threadIdx.x == 0) s_ticket = atomicAdd(p.ctrl + cls * kTileCtrlStride, 1u);
    __syncthreads();
    // the group's chains, chunk-major: ticket i -> chunk i / per, chain cls + classes * (i % per)
    const uint32_t per = p.outer * p.ntiles / p.classes;
    const uint32_t i = s_ticket;
    if (i >= per * p.nchunks) return;  // uniform (the grid holds exactly classes * per * nchunks)
    const uint32_t chunk = i / per;
    const uint32_t chain = cls + p.classes * (i - chunk * per);
    const uint32_t tile = chain % p.ntiles;
    const uint32_t o = chain / p.ntiles;
    const uint32_t col = tile * uint32_t(COLS) + uint32_t(lane * V);
    const int nv = col < p.inner ? min(V, int(p.inner - col)) : 0;
    const uint32_t r0 = chunk * kTileLbRows;
    const uint32_t r1 = min(p.len, r0 + kTileLbRows);
    const T* base = reinterpret_cast<const T*>(p.in) + int64_t(o) * p.in_s_outer + col;
    const int odt = NATIVE ? int(ScanInDtype<T>::value) : p.out_dtype;
    const int rdt = NATIVE ? int(ScanInDtype<T>::value) : p.rdt;
    const int oesz = dtype_size(odt);
    char* obase = p.out + int64_t(o) * p.out_s_outer * oesz;
    const bool nt_in = p.nt_in != 0, nt_out = p.nt_out != 0;
    A id;
    id.init(p.identity);
    // the whole chunk: sub-tile s, wave w holds rows r0 + s*SUB + w*RW + [0, RW)
    // (rows past the chunk clamped to its last row; their values are never used)
    const bool whole = (tile + 1u) * uint32_t(COLS) <= p.inner;  // uniform in the block
    T x[NS][RW][V];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const uint32_t rw0 = r0 + uint32_t(s) * SUB + uint32_t(wave * RW);
        if (whole && nt_in) {
#pragma unroll
            for (int i = 0; i < RW; ++i)
                load_row<V, true>(base + int64_t(min(rw0 + uint32_t(i), r1 - 1)) * p.in_s_len, x[s][i]);
        } else if (whole) {
#pragma unroll
            for (int i = 0; i < RW; ++i)
                load_row<V, false>(base + int64_t(min(rw0 + uint32_t(i), r1 - 1)) * p.in_s_len, x[s][i]);
        } else {
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                const uint32_t row = rw0 + uint32_t(i);
                load_row_partial<V>(base + int64_t(min(row, r1 - 1)) * p.in_s_len, row < r1, nv, x[s][i]);
            }
        }
    }
    // every (sub-tile, wave) total of every column
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const uint32_t rw0 = r0 + uint32_t(s) * SUB + uint32_t(wave * RW);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            A t = id;
#pragma unroll
            for (int i = 0; i < RW; ++i)
                if (rw0 + uint32_t(i) < r1) t.add(reg_put<T>(x[s][i][v]));
            sm.wtot[s][wave][lane * V + v] = t.store();
        }
    }
    __syncthreads();
    if constexpr (!PACKED) {  // thread c < COLS: column c; value, wait, flag per (chunk, tile)
        const int c = int(threadIdx.x);
        const bool has_col = c < COLS;
        A agg = id;
        if (has_col) {
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    A t;
                    t.load(sm.wtot[s][w][c]);
                    agg.merge(t);
                }
        }
        const int64_t vbase = int64_t(o) * p.nchunks * p.ncols + tile * uint32_t(COLS) + (has_col ? c : 0);
        uint32_t* ftile = p.flags + int64_t(o) * p.nchunks * p.ntiles + tile;
        uint32_t* fmine = ftile + int64_t(chunk) * p.ntiles;
        const int64_t vmine = vbase + int64_t(chunk) * p.ncols;
        A excl = id;
        if (chunk == 0) {  // uniform in the block
            tile_publish_flagged<AK, OP>(p.pre + vmine, has_col, agg, fmine, kLbPrefix);
        } else {
            tile_publish_flagged<AK, OP>(p.agg + vmine, has_col, agg, fmine, kLbAggregate);
            if (wave == 0) {
                bool to = false;
                const uint32_t d = tile_lead_depth_flags(ftile, p.ntiles, chunk, &to);
                if (lane == 0) {
                    s_depth = d;
                    s_timed_out = to ? 1u : 0u;
                    if (to) atomicAdd(p.ctrl + kTileTimeouts, 1u);
                }
            }
            __syncthreads();
            if (has_col)
                excl = tile_follow_values<AK, OP>(p.agg + vbase, p.pre + vbase, p.ncols, chunk, s_depth,
                                                  s_timed_out != 0, id);
            A incl = excl;
            incl.merge(agg);
            tile_publish_flagged<AK, OP>(p.pre + vmine, has_col, incl, fmine, kLbPrefix);
        }
        if (has_col) s_excl[c] = excl.store();
    } else {  // thread c: column c of the tile, the chunk's total, look-back, publications
        const int c = int(threadIdx.x);
        A agg = id;
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                A t;
                t.load(sm.wtot[s][w][c]);
                agg.merge(t);
            }
        uint64_t* wtile = p.words + int64_t(o) * p.nchunks * p.ncols + tile * uint32_t(COLS);
        uint64_t* mine = wtile + int64_t(chunk) * p.ncols + c;
        A excl = id;
        if (chunk == 0) {  // uniform in the block
            lb_store(mine, lb_pack(agg, kLbPrefix));
        } else {
            lb_store(mine, lb_pack(agg, kLbAggregate));
            if (wave == 0) {
                const uint32_t d = tile_lead_depth(wtile, p.ncols, chunk);
                if (lane == 0) s_depth = d;
            }
            __syncthreads();
            excl = tile_follow<AK, OP>(wtile + c, p.ncols, chunk, s_depth, id, p.ctrl + kTileTimeouts);
            A incl = excl;
            incl.merge(agg);
            lb_store(mine, lb_pack(incl, kLbPrefix));
        }
        s_excl[c] = excl.store();
    }
    __syncthreads();
    // each lane's start for every sub-tile: the chunk's carry, then the
    // (sub-tile, wave) totals before it in row order
    A pre[NS][V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        A carry;
        carry.load(s_excl[lane * V + v]);
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                A t;
                t.load(sm.wtot[s][w][lane * V + v]);
                if (w == wave) pre[s][v] = carry;
                carry.merge(t);
            }
    }
    __syncthreads();  // wtot is read; the transposed form reuses the space
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const uint32_t sub = r0 + uint32_t(s) * SUB;
        const uint32_t rw0 = sub + uint32_t(wave * RW);
        A run[V];
#pragma unroll
        for (int v = 0; v < V; ++v) run[v] = pre[s][v];
        if constexpr (!TRANS) {
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                const uint32_t row = rw0 + uint32_t(i);
                if (row >= r1) continue;  // uniform in the wave
                Reg y[V];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    run[v].add(reg_put<T>(x[s][i][v]));
                    const Reg val = run[v].value(rdt);
                    if constexpr (NATIVE) y[v] = val;
                    else y[v] = cvt(val, rdt, odt);
                }
                char* ptr = obase + (int64_t(row) * p.out_s_len + int64_t(col) * p.out_s_inner) * oesz;
                if (nv == V) {
                    store_vec_rt(nt_out, ptr, oesz, odt, y);
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (v < nv) store_raw(ptr + int64_t(v) * p.out_s_inner * oesz, oesz, y[v]);
                }
            }
        } else {
            // the results in place of the values, then the chunk as [column][row] in
            // LDS in two halves of 128 columns (lanes 0-31, then 32-63), each stored
            // along the output's rows: a column's 64 rows are one 256-byte segment
#pragma unroll
            for (int i = 0; i < RW; ++i) {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    if (rw0 + uint32_t(i) < r1) run[v].add(reg_put<T>(x[s][i][v]));
                    x[s][i][v] = reg_get<T>(run[v].value(rdt));
                }
            }
            constexpr int TPC = CH / V;     // threads per column: 16 bytes each
            constexpr int CPP = 256 / TPC;  // columns per pass
            const int cc = threadIdx.x / TPC;
            const int rr = (threadIdx.x - cc * TPC) * V;
            const uint32_t row = sub + uint32_t(rr);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if ((lane >> 5) == h) {
#pragma unroll
                    for (int i = 0; i < RW; ++i)
#pragma unroll
                        for (int v = 0; v < V; ++v) sm.tile[(lane & 31) * V + v][wave * RW + i] = x[s][i][v];
                }
                __syncthreads();
#pragma unroll
                for (int c = cc; c < HC; c += CPP) {
                    const uint32_t gcol = tile * uint32_t(COLS) + uint32_t(h * HC + c);
                    if (gcol >= p.inner || row >= r1) continue;
                    char* ptr = obase + (int64_t(row) * p.out_s_len + int64_t(gcol) * p.out_s_inner) * oesz;
                    Reg q[V];
#pragma unroll
                    for (int v = 0; v < V; ++v) q[v] = reg_put<T>(sm.tile[c][rr + v]);
                    if (row + uint32_t(V) <= r1) {
                        store_vec_rt(nt_out, ptr, oesz, odt, q);
                    } else {
#pragma unroll
                        for (int v = 0; v < V; ++v)
                            if (row + uint32_t(v) < r1) store_raw(ptr + int64_t(v) * oesz, oesz, q[v]);
                    }
                }
                __syncthreads();  // the tile space is written again for the second half
            }
        }
    }
}

}  // namespace xg
