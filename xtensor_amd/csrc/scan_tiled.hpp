// scan_tiled.hpp — scans along a strided axis, in tiles of the plane the
// scan axis spans with the contiguous one.
//
// accumulator_impl (reducers/xaccumulator.hpp:240-297) runs the recurrence
// res[k] = res[k-1] + res[k] down the axis for every position of the other
// axes.  When the scan axis is strided and another axis is contiguous (the
// column scan cumsum(X, 0) of a row-major X, or cumsum(transpose(X), 1) whose
// input is contiguous across the scan lines), a thread per column walking
// the whole axis leaves the GPU nearly idle.  Here the axis is cut into
// chunks and every (chunk, column tile) is a workgroup, in three launches:
//   1. totals:  each workgroup sums its chunk of every column of its tile;
//   2. carries: per column, the exclusive scan of the chunk totals, in place;
//   3. scan:    each workgroup scans its chunk starting from its carry and
//               stores the results.
// Every merge happens in a fixed order (rows inside a wave, then waves in
// order, then chunks in order), so the results do not depend on timing.
// The input is read twice (phases 1 and 3); phase 1 uses the default cache
// policy so the second read can be served by the last-level cache.
//
// Tiles: a lane owns V adjacent columns (16 bytes), a wave 64 * V columns,
// a workgroup of four waves the same columns over SUB rows at a time
// (RW consecutive rows per wave).  TRANS: the output is contiguous along the
// scan axis instead (the transposed case); a tile passes through LDS so that
// stores run along the output's rows.
//
// Included by scan_kernels.hpp, after ScanIn / ScanInDtype.
#pragma once

#include <hip/hip_runtime.h>

#include "acc.hpp"
#include "scan.hpp"

namespace xg {

// V elements of type T (16 bytes, aligned) from ptr
template <int V, bool NT, class T> __device__ __forceinline__ void load_row(const T* ptr, T (&x)[V]) {
    Reg q[V];
    load_vec<V, NT>(reinterpret_cast<const char*>(ptr), int(sizeof(T)), ScanInDtype<T>::value, q);
#pragma unroll
    for (int v = 0; v < V; ++v) x[v] = reg_get<T>(q[v]);
}

// the same V elements, each only if v < nv and `ok` (else 0): the ragged edges
template <int V, class T> __device__ __forceinline__ void load_row_partial(const T* ptr, bool ok, int nv, T (&x)[V]) {
#pragma unroll
    for (int v = 0; v < V; ++v) x[v] = (ok && v < nv) ? ptr[v] : T(0);
}

template <int AK, int OP>
__global__ __launch_bounds__(256) void scan_tiles_totals_kernel(const TileScan p) {
    using A = Acc<AK, OP>;
    using T = typename ScanIn<AK>::T;
    constexpr int V = int(16 / sizeof(T));
    constexpr int U = 8;  // rows in flight per wave
    constexpr int COLS = 64 * V;
    __shared__ Store part[4][COLS];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t b = blockIdx.x;
    const uint32_t tile = b % p.ntiles;
    b /= p.ntiles;
    const uint32_t chunk = b % p.nchunks;
    const uint32_t o = b / p.nchunks;
    const uint32_t col = tile * uint32_t(COLS) + uint32_t(lane * V);
    const int nv = col < p.inner ? min(V, int(p.inner - col)) : 0;
    const uint32_t r0 = chunk * p.chunk_rows;
    const uint32_t r1 = min(p.len, r0 + p.chunk_rows);
    const T* base = reinterpret_cast<const T*>(p.in) + int64_t(o) * p.in_s_outer + col;
    A id;
    id.init(p.identity);
    A acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = id;
    // whole tiles (every lane V columns): the U rows' loads issue together,
    // rows past the chunk clamped to its last row and their values skipped
    const bool whole = (tile + 1u) * uint32_t(COLS) <= p.inner;  // uniform in the block
    for (uint32_t r = r0 + uint32_t(wave); r < r1; r += 4u * U) {
        T x[U][V];
        if (whole) {
#pragma unroll
            for (int u = 0; u < U; ++u)
                load_row<V, false>(base + int64_t(min(r + 4u * uint32_t(u), r1 - 1)) * p.in_s_len, x[u]);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t row = r + 4u * uint32_t(u);
                load_row_partial<V>(base + int64_t(min(row, r1 - 1)) * p.in_s_len, row < r1, nv, x[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (r + 4u * uint32_t(u) >= r1) continue;  // uniform in the wave
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (v < nv) acc[v].add(reg_put<T>(x[u][v]));
        }
    }
#pragma unroll
    for (int v = 0; v < V; ++v) part[wave][lane * V + v] = acc[v].store();
    __syncthreads();
    // one thread per column of the tile merges the four waves in order
    for (int c = threadIdx.x; c < COLS; c += 256) {
        A t = id;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            A q;
            q.load(part[w][c]);
            t.merge(q);
        }
        p.part[(int64_t(o) * p.nchunks + chunk) * p.ncols + tile * uint32_t(COLS) + uint32_t(c)] = t.store();
    }
}

// In place: the chunk totals of every column become exclusive prefixes over
// the chunks.  A workgroup takes 64 columns; wave w the w-th quarter of the
// chunks.
template <int AK, int OP>
__global__ __launch_bounds__(256) void scan_tiles_carry_kernel(const TileScan p) {
    using A = Acc<AK, OP>;
    __shared__ Store wsum[4][64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint32_t groups = (p.ncols + 63u) / 64u;
    const uint32_t o = blockIdx.x / groups;
    const uint32_t col = (blockIdx.x - o * groups) * 64u + uint32_t(lane);
    const uint32_t q = (p.nchunks + 3u) / 4u;
    const uint32_t c0 = min(p.nchunks, uint32_t(wave) * q), c1 = min(p.nchunks, c0 + q);
    Store* base = p.part + int64_t(o) * p.nchunks * p.ncols + col;
    A id;
    id.init(p.identity);
    A loc = id;
    if (col < p.ncols) {
#pragma unroll 8
        for (uint32_t c = c0; c < c1; ++c) {
            A t;
            t.load(base[int64_t(c) * p.ncols]);
            loc.merge(t);
        }
    }
    wsum[wave][lane] = loc.store();
    __syncthreads();
    A pre = id;
    for (int w = 0; w < wave; ++w) {
        A t;
        t.load(wsum[w][lane]);
        pre.merge(t);
    }
    if (col < p.ncols) {
#pragma unroll 8
        for (uint32_t c = c0; c < c1; ++c) {
            A t;
            t.load(base[int64_t(c) * p.ncols]);
            base[int64_t(c) * p.ncols] = pre.store();
            pre.merge(t);
        }
    }
}

template <int AK, int OP, bool NATIVE, bool TRANS>
__global__ __launch_bounds__(256) void scan_tiles_kernel(const TileScan p) {
    using A = Acc<AK, OP>;
    using T = typename ScanIn<AK>::T;
    constexpr int V = int(16 / sizeof(T));
    constexpr int COLS = 64 * V;
    constexpr int RW = 16;  // rows per wave in a sub-tile
    constexpr uint32_t SUB = 4u * RW;
    constexpr int HC = COLS / 2;         // TRANS: columns of one half of the tile in LDS
    constexpr int PITCH = int(SUB) + 1;  // output tile row pitch (TRANS), in elements
    static_assert(!TRANS || NATIVE, "the transposed scan stores the input's type");
    union Shared {
        Store wtot[4][COLS];
        T tile[TRANS ? HC : 1][TRANS ? PITCH : 1];
    };
    __shared__ Shared sm;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t b = blockIdx.x;
    const uint32_t tile = b % p.ntiles;
    b /= p.ntiles;
    const uint32_t chunk = b % p.nchunks;
    const uint32_t o = b / p.nchunks;
    const uint32_t col = tile * uint32_t(COLS) + uint32_t(lane * V);
    const int nv = col < p.inner ? min(V, int(p.inner - col)) : 0;
    const uint32_t r0 = chunk * p.chunk_rows;
    const uint32_t r1 = min(p.len, r0 + p.chunk_rows);
    const T* base = reinterpret_cast<const T*>(p.in) + int64_t(o) * p.in_s_outer + col;
    const int odt = NATIVE ? int(ScanInDtype<T>::value) : p.out_dtype;
    const int rdt = NATIVE ? int(ScanInDtype<T>::value) : p.rdt;
    const int oesz = dtype_size(odt);
    char* obase = p.out + int64_t(o) * p.out_s_outer * oesz;
    const bool nt_in = p.nt_in != 0, nt_out = p.nt_out != 0;
    A id;
    id.init(p.identity);
    A carry[V];
#pragma unroll
    for (int v = 0; v < V; ++v) {
        carry[v] = id;
        if (p.nchunks > 1) carry[v].load(p.part[(int64_t(o) * p.nchunks + chunk) * p.ncols + col + uint32_t(v)]);
    }
    const bool whole = (tile + 1u) * uint32_t(COLS) <= p.inner;  // uniform in the block
    for (uint32_t sub = r0; sub < r1; sub += SUB) {
        const uint32_t rw0 = sub + uint32_t(wave * RW);
        T x[RW][V];
        // whole tiles: all RW rows' loads issue together (rows past the chunk
        // clamped to its last row; their values are never used)
        if (whole && nt_in) {
#pragma unroll
            for (int i = 0; i < RW; ++i)
                load_row<V, true>(base + int64_t(min(rw0 + uint32_t(i), r1 - 1)) * p.in_s_len, x[i]);
        } else if (whole) {
#pragma unroll
            for (int i = 0; i < RW; ++i)
                load_row<V, false>(base + int64_t(min(rw0 + uint32_t(i), r1 - 1)) * p.in_s_len, x[i]);
        } else {
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                const uint32_t row = rw0 + uint32_t(i);
                load_row_partial<V>(base + int64_t(min(row, r1 - 1)) * p.in_s_len, row < r1, nv, x[i]);
            }
        }
        // this wave's totals, then the prefix of the waves before it
        A pre[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            A t = id;
#pragma unroll
            for (int i = 0; i < RW; ++i)
                if (rw0 + uint32_t(i) < r1) t.add(reg_put<T>(x[i][v]));
            sm.wtot[wave][lane * V + v] = t.store();
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < V; ++v) {
            pre[v] = carry[v];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                A t;
                t.load(sm.wtot[w][lane * V + v]);
                if (w == wave) pre[v] = carry[v];
                carry[v].merge(t);
            }
        }
        __syncthreads();  // wtot is read; the space is reused below / next sub-tile
        // results, row by row (one row of V values live at a time)
        A run[V];
#pragma unroll
        for (int v = 0; v < V; ++v) run[v] = pre[v];
        if constexpr (!TRANS) {
#pragma unroll
            for (int i = 0; i < RW; ++i) {
                const uint32_t row = rw0 + uint32_t(i);
                if (row >= r1) continue;  // uniform in the wave
                Reg y[V];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    run[v].add(reg_put<T>(x[i][v]));
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
            // the results in place of the values, then the sub-tile as [column][row]
            // in LDS in two halves of the columns (lanes 0-31, then 32-63), each
            // stored along the output's rows: a column's 64 rows leave as one
            // segment of 64 elements (32-row sub-tiles, half that, measured slower
            // on the one-pass form)
#pragma unroll
            for (int i = 0; i < RW; ++i) {
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    if (rw0 + uint32_t(i) < r1) run[v].add(reg_put<T>(x[i][v]));
                    x[i][v] = reg_get<T>(run[v].value(rdt));
                }
            }
            constexpr int TPC = int(SUB) / V;  // threads per column: 16 bytes each
            constexpr int CPP = 256 / TPC;     // columns per pass
            const int cc = threadIdx.x / TPC;
            const int rr = (threadIdx.x - cc * TPC) * V;
            const uint32_t row = sub + uint32_t(rr);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if ((lane >> 5) == h) {
#pragma unroll
                    for (int i = 0; i < RW; ++i)
#pragma unroll
                        for (int v = 0; v < V; ++v) sm.tile[(lane & 31) * V + v][wave * RW + i] = x[i][v];
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
                __syncthreads();  // the tile space is written again (second half, next sub-tile)
            }
        }
    }
}

}  // namespace xg
