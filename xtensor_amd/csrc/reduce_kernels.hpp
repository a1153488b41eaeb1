// reduce_kernels.hpp — the reduction kernel templates, included by the
// translation units that instantiate them (split so they compile in
// parallel).  See reduce.hip for the design.
#pragma once

#include <hip/hip_runtime.h>

#include "acc.hpp"
#include "chunk.hpp"
#include "spec_eval.hpp"
#include "reduce.hpp"

namespace xg {

// ---------------------------------------------------------------- finishing
// output element o (row-major over the kept dims) -> output element offset
__device__ __forceinline__ int64_t out_offset(const RedPlan& p, uint32_t o) {
    int64_t off = 0;
    if (p.layout == RL_COLS) {
        uint32_t q = fastdiv(o, p.div[p.rank - 1]);
        off += int64_t(o - q * p.lf) * p.out_stride[p.rank - 1];
        o = q;
    }
    for (int k = p.n_kept - 1; k >= 0; --k) {
        uint32_t q = fastdiv(o, p.div[k]);
        off += int64_t(o - q * uint32_t(p.shape[k])) * p.out_stride[k];
        o = q;
    }
    return off;
}

// a / b in value type dt (int32 or wider; integer division truncates, and
// a zero divisor gives 0 where C++ leaves it undefined)
__device__ __forceinline__ Reg div_in(Reg a, Reg b, int dt) {
    switch (dt) {
    case XG_FLOAT32: return flt_binop<float>(MO_DIV, a, b);
    case XG_FLOAT64: return flt_binop<double>(MO_DIV, a, b);
    case XG_INT32: return int_binop<int32_t>(MO_DIV, a, b);
    case XG_UINT32: return int_binop<uint32_t>(MO_DIV, a, b);
    case XG_INT64: return int_binop<int64_t>(MO_DIV, a, b);
    default: return int_binop<uint64_t>(MO_DIV, a, b);
    }
}

template <class A> __device__ void finish(const RedPlan& p, A acc, uint32_t o) {
    if (p.has_initial) {  // xreducer merges `initial` in front of the folded value
        A ini;
        ini.init(p.initial);
        ini.merge(acc);
        acc = ini;
    }
    Reg v = acc.value(p.vdt);
    Reg r = v;
    int rt = p.vdt;
    if (p.fin == FZ_MEAN || p.fin == FZ_MEAN_SQRT) {  // s / div in their common type
        r = div_in(cvt(v, p.vdt, p.mdt), Reg(p.div_bits), p.mdt);
        rt = p.mdt;
    }
    if (p.fin == FZ_SQRT || p.fin == FZ_MEAN_SQRT) {
        if (p.rdt == XG_FLOAT32) r = from_f32(sqrtf(as_f32(cvt(r, rt, XG_FLOAT32))));
        else r = from_f64(sqrt(as_f64(cvt(r, rt, XG_FLOAT64))));
    }
    const int odt = p.out_dtype;
    Reg ov = odt == XG_BFLOAT16 ? cvt(r, p.rdt, XG_FLOAT32) : cvt(r, p.rdt, odt);
    const int esz = dtype_size(odt);
    store_raw(p.out_ptr + out_offset(p, o) * esz, esz, to_storage(ov, odt));
}

// group-wide combine of G (power of two) threads; all threads of the block call it
template <class A>
__device__ A group_combine(A a, int G, Store* lds) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wg = G < 64 ? G : 64;
    for (int m = 1; m < wg; m <<= 1) {
        A o = a.shfl_xor(m);
        if (lane & m) {  // combine in index order: lower lane's value first
            o.merge(a);
            a = o;
        } else {
            a.merge(o);
        }
    }
    if (G > 64) {
        if (lane == 0) lds[wave] = a.store();
        __syncthreads();
        const int per = G / 64;
        const int first = (wave / per) * per;
        A r;
        r.load(lds[first]);
        for (int w = 1; w < per; ++w) {
            A o;
            o.load(lds[first + w]);
            r.merge(o);
        }
        __syncthreads();
        a = r;
    }
    return a;
}


// ---------------------------------------------------------------- rows layout
template <int V, int NR, class S, bool I64, int AK, int OP, int kUnroll>
__global__ __launch_bounds__(256) void reduce_rows_kernel(const RedPlan p_arg) {
    const RedPlan& p = kernel_plan<S>(p_arg);
    using A = Acc<AK, OP>;
    __shared__ Store lds[4];
    const int G = p.group;
    const int per_block = 256 / G;
    const uint32_t grp = threadIdx.x / G;
    const uint32_t gl = threadIdx.x % G;
    const uint32_t items = p.n_out * uint32_t(p.splits);
    for (uint32_t base = blockIdx.x * per_block; base < items; base += gridDim.x * per_block) {
        const uint32_t wi = base + grp;
        const bool active = wi < items;
        A acc;
        acc.init(p.identity);
        uint32_t o = 0;
        if (active) {
            o = fastdiv(wi, p.div_splits);
            const uint32_t sp = wi - o * uint32_t(p.splits);
            uint32_t idx[kMaxRank];
#pragma unroll
            for (int k = 0; k < kMaxRank; ++k) idx[k] = 0;
            uint32_t q = o;
            for (int k = p.n_kept - 1; k >= 0; --k) {
                uint32_t d = fastdiv(q, p.div[k]);
                idx[k] = q - d * uint32_t(p.shape[k]);
                q = d;
            }
            const uint32_t c0 = sp * p.per_split;
            const uint32_t c1 = min(p.red_count, c0 + p.per_split);
            for (uint32_t c = c0 + gl; c < c1; c += G * kUnroll) {
                Reg vals[kUnroll][V];
                int nv[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    const uint32_t cc = c + u * G;
                    nv[u] = 0;
                    if (cc < c1) {
                        uint32_t ro = fastdiv(cc, p.div_nchunks_f);
                        const uint32_t j0 = (cc - ro * p.nchunks_f) * V;
                        uint32_t id2[kMaxRank];
#pragma unroll
                        for (int k = 0; k < kMaxRank; ++k) id2[k] = idx[k];
                        for (int k = p.rank - 2; k >= p.n_kept; --k) {
                            uint32_t d = fastdiv(ro, p.div[k]);
                            id2[k] = ro - d * uint32_t(p.shape[k]);
                            ro = d;
                        }
                        nv[u] = min(V, int(p.lf - j0));
                        eval_chunk<V, NR, S, I64>(p.leaves, p.prog, p.rank, id2, j0, nv[u], vals[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < kUnroll; ++u)
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (v < nv[u]) acc.add(vals[u][v]);
            }
        }
        acc = group_combine(acc, G, lds);
        if (active && gl == 0) {
            if (p.splits == 1) finish(p, acc, o);
            else reinterpret_cast<Store*>(p.partials)[wi] = acc.store();
        }
    }
}

// ---------------------------------------------------------------- cols layout
template <int V, int NR, class S, bool I64, int AK, int OP, int kUnroll>
__global__ __launch_bounds__(256) void reduce_cols_kernel(const RedPlan p_arg) {
    const RedPlan& p = kernel_plan<S>(p_arg);
    using A = Acc<AK, OP>;
    __shared__ typename A::Lds lds[4][64][V];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint32_t n_ko = p.n_out / p.lf;
    const uint32_t items = n_ko * p.col_tiles * uint32_t(p.splits);
    for (uint32_t wi = blockIdx.x; wi < items; wi += gridDim.x) {
        const uint32_t rest = fastdiv(wi, p.div_splits);
        const uint32_t sp = wi - rest * uint32_t(p.splits);
        const uint32_t ko = fastdiv(rest, p.div_col_tiles);
        const uint32_t tile = rest - ko * p.col_tiles;
        const uint32_t cc = tile * 64 + lane;
        const bool active = cc < p.nchunks_f;
        const uint32_t j0 = cc * V;
        const int nvalid = active ? min(V, int(p.lf - j0)) : 0;
        uint32_t idx[kMaxRank];
#pragma unroll
        for (int k = 0; k < kMaxRank; ++k) idx[k] = 0;
        uint32_t q = ko;
        for (int k = p.n_kept - 1; k >= 0; --k) {
            uint32_t d = fastdiv(q, p.div[k]);
            idx[k] = q - d * uint32_t(p.shape[k]);
            q = d;
        }
        A acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v].init(p.identity);
        if (active) {
            const uint32_t r0 = sp * p.per_split;
            const uint32_t r1 = min(p.red_count, r0 + p.per_split);
            for (uint32_t r = r0 + wave; r < r1; r += 4 * kUnroll) {
                Reg vals[kUnroll][V];
                bool ok[kUnroll];
#pragma unroll
                for (int u = 0; u < kUnroll; ++u) {
                    const uint32_t rr = r + 4 * u;
                    ok[u] = rr < r1;
                    if (ok[u]) {
                        uint32_t id2[kMaxRank];
#pragma unroll
                        for (int k = 0; k < kMaxRank; ++k) id2[k] = idx[k];
                        uint32_t t = rr;
                        for (int k = p.rank - 2; k >= p.n_kept; --k) {
                            uint32_t d = fastdiv(t, p.div[k]);
                            id2[k] = t - d * uint32_t(p.shape[k]);
                            t = d;
                        }
                        eval_chunk<V, NR, S, I64>(p.leaves, p.prog, p.rank, id2, j0, nvalid, vals[u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < kUnroll; ++u)
                    if (ok[u])
#pragma unroll
                        for (int v = 0; v < V; ++v) acc[v].add(vals[u][v]);
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v) lds[wave][lane][v] = acc[v].lds();
        __syncthreads();
        if (wave == 0 && active) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if (v < nvalid) {
                    A a;
                    a.from_lds(lds[0][lane][v]);
                    for (int w = 1; w < 4; ++w) {
                        A b;
                        b.from_lds(lds[w][lane][v]);
                        a.merge(b);
                    }
                    const uint32_t o = ko * p.lf + j0 + v;
                    if (p.splits == 1) finish(p, a, o);
                    else reinterpret_cast<Store*>(p.partials)[size_t(o) * p.splits + sp] = a.store();
                }
            }
        }
        __syncthreads();
    }
}

// combine within aligned groups of G (power of two <= 64) lanes of a wave,
// lower lane's value first; every lane of the wave must call it
template <class A> __device__ __forceinline__ A wave_group_combine(A a, int G) {
    const int lane = threadIdx.x & 63;
    for (int m = 1; m < G; m <<= 1) {
        A o = a.shfl_xor(m);
        if (lane & m) {
            o.merge(a);
            a = o;
        } else {
            a.merge(o);
        }
    }
    return a;
}

// ---------------------------------------------------------------- split combine
// A partial as written by another workgroup in the same launch: device-scope
// loads, which bypass the non-coherent per-XCD L2 (DEV), or plain loads
// (partials written by an earlier launch).
template <bool DEV> __device__ __forceinline__ Store load_partial(const Store* p) {
    if constexpr (DEV) {
        return Store{__hip_atomic_load(&p->a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                     __hip_atomic_load(&p->b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
    } else {
        return *p;
    }
}

// The split partials of output o merged by T (a power of two <= 64) lanes:
// lane t merges partials t, t + T, ... in order (loaded 8 at a time), then
// the T lanes combine in a fixed butterfly, lower lane first.  Every lane of
// the wave calls it; lanes with o past the outputs contribute the identity.
template <class A, bool DEV>
__device__ __forceinline__ A combine_partials(const RedPlan& p, uint32_t o, uint32_t T, uint32_t t) {
    A a;
    a.init(p.identity);
    if (o < p.n_out) {
        const size_t ps = p.split_major ? size_t(p.n_out) : 1;  // partial stride between splits
        const Store* part = reinterpret_cast<const Store*>(p.partials) + (p.split_major ? size_t(o) : size_t(o) * p.splits);
        uint32_t s = t;
        for (; s + 7 * T < uint32_t(p.splits); s += 8 * T) {
            Store st[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) st[k] = load_partial<DEV>(part + size_t(s + uint32_t(k) * T) * ps);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                A b;
                b.load(st[k]);
                a.merge(b);
            }
        }
        for (; s < uint32_t(p.splits); s += T) {
            A b;
            b.load(load_partial<DEV>(part + size_t(s) * ps));
            a.merge(b);
        }
    }
    return wave_group_combine(a, int(T));
}

// p.group (T) lanes per output, see combine_partials
template <int AK, int OP>
__global__ __launch_bounds__(256) void reduce_combine_kernel(const RedPlan p) {
    using A = Acc<AK, OP>;
    const uint32_t T = uint32_t(p.group);
    const uint32_t per_block = 256u / T;
    const uint32_t t = threadIdx.x & (T - 1);
    for (uint32_t ob = blockIdx.x * per_block; ob < p.n_out; ob += gridDim.x * per_block) {
        const uint32_t o = ob + threadIdx.x / T;  // uniform over each group of T lanes
        A a = combine_partials<A, false>(p, o, T, t);
        if (o < p.n_out && t == 0) finish(p, a, o);
    }
}

// a block per output, for many partials: thread t merges partials
// t, t + 256, ... in order (loaded 16 at a time, so a few thousand partials
// cost one or two memory round trips), then the block combines in a fixed
// tree
template <int AK, int OP>
__global__ __launch_bounds__(256) void reduce_combine_wide_kernel(const RedPlan p) {
    using A = Acc<AK, OP>;
    constexpr int B = 16;
    __shared__ Store lds[4];
    for (uint32_t o = blockIdx.x; o < p.n_out; o += gridDim.x) {
        const size_t ps = p.split_major ? size_t(p.n_out) : 1;  // partial stride between splits
        const Store* part = reinterpret_cast<const Store*>(p.partials) + (p.split_major ? size_t(o) : size_t(o) * p.splits);
        A a;
        a.init(p.identity);
        int s = int(threadIdx.x);
        for (; s + (B - 1) * 256 < p.splits; s += B * 256) {
            Store st[B];
#pragma unroll
            for (int k = 0; k < B; ++k) st[k] = part[size_t(s + k * 256) * ps];
#pragma unroll
            for (int k = 0; k < B; ++k) {
                A b;
                b.load(st[k]);
                a.merge(b);
            }
        }
        for (; s < p.splits; s += 256) {
            A b;
            b.load(part[size_t(s) * ps]);
            a.merge(b);
        }
        a = group_combine(a, 256, lds);
        if (threadIdx.x == 0) finish(p, a, o);
    }
}

// ---------------------------------------------------------------- compiled fast kernels
// For compact plans (normalised rank <= 2, 32-bit offsets) and a compiled
// program: leaf offsets are one multiply-add per dimension, U chunks (rows)
// are in flight per lane, and the rows kernel combines inside a wave only
// (no LDS, no barrier).  p.group / p.splits / p.per_split hold the fast
// split (f_group / f_splits / f_per_split of the plan).

// the non-temporal flags of a compiled program's leaves that are not
// broadcast: 1 if all are set, 0 if none is, 2 if they differ
template <class S>
__device__ __forceinline__ int spec_leaves_nt(const RedPlan& p) {
    bool all = true, none = true;
#pragma unroll
    for (int l = 0; l < S::nleaves; ++l) {
        if (lm_base(S::leaf_mode[l]) == LM_BCAST) continue;
        const bool nt = lm_nt(p.leaves[l].mode);
        all = all && nt;
        none = none && !nt;
    }
    return all ? 1 : none ? 0 : 2;
}

// chunks (rows) in flight per lane (the cols kernel measured 4.2 TB/s with
// 4 and 4.0 with 8 on a [8192, 4096] float32 column sum)
template <int NR, int V, bool COLS> constexpr int fast_unroll() {
    return NR * V <= 8 ? 4 : (NR * V <= 16 ? 2 : 1);
}

// rows: dims [outputs | fast] (RANK 2) or [fast] (RANK 1, one output);
// G lanes per (output, split), 64 / G of them per wave
template <int V, int NR, class S, int AK, int OP, int RANK>
__global__ __launch_bounds__(256) void reduce_rows_fast_kernel(const RedPlan p) {
    using A = Acc<AK, OP>;
    constexpr int U = fast_unroll<NR, V, false>();
    const int G = p.group;
    const int lane = threadIdx.x & 63;
    const uint32_t per_wave = 64u / uint32_t(G);
    const uint32_t wi = (blockIdx.x * 4u + uint32_t(threadIdx.x >> 6)) * per_wave + uint32_t(lane) / uint32_t(G);
    const uint32_t gl = uint32_t(lane) & uint32_t(G - 1);
    const uint32_t items = p.n_out * uint32_t(p.splits);
    A acc;
    acc.init(p.identity);
    const bool live = wi < items;
    const uint32_t o = live ? fastdiv(wi, p.div_splits) : 0u;
    const uint32_t sp = wi - o * uint32_t(p.splits);
    const uint32_t c0 = sp * p.per_split;
    const uint32_t c1 = live ? min(p.nchunks_f, c0 + p.per_split) : 0u;
    const uint32_t idx[4] = {RANK == 2 ? o : 0u, 0u, 0u, 0u};
    // FULL: every chunk of the lane's range is whole, so chunks past the
    // range's end are clamped to its last chunk and their values dropped,
    // and no load sits behind a divergent branch (U loads stay in flight)
    auto sweep = [&](auto full_c, auto ntm_c) {
        constexpr bool FULL = decltype(full_c)::value;
        constexpr int NTM = decltype(ntm_c)::value;
        for (uint32_t c = c0 + gl; c < c1; c += uint32_t(G) * U) {
            Reg R[U][NR][V];
            int nv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t cc = c + uint32_t(u * G);
                if constexpr (FULL) {
                    nv[u] = cc < c1 ? V : 0;
                    run_spec<S, 0, V, NR, RANK, true, NTM>(p, idx, min(cc, c1 - 1) * V, V, R[u]);
                } else {
                    nv[u] = cc < c1 ? min(V, int(p.lf - cc * V)) : 0;
                    if (nv[u] > 0) run_spec<S, 0, V, NR, RANK>(p, idx, cc * V, nv[u], R[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if constexpr (FULL) {  // a select, not a branch (see the cols kernel)
                    A t = acc;
#pragma unroll
                    for (int v = 0; v < V; ++v) t.add(R[u][S::result_reg][v]);
                    acc = nv[u] > 0 ? t : acc;
                } else {
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (v < nv[u]) acc.add(R[u][S::result_reg][v]);
                }
            }
        }
    };
    // whole chunks up to the lane's range end (or an empty range): uniform over the wave
    const bool full = __all(c1 <= c0 || c1 * uint32_t(V) <= p.lf);
    const int ntm = spec_leaves_nt<S>(p);
    if (full && ntm == 1) sweep(std::true_type{}, std::integral_constant<int, 1>{});
    else if (full && ntm == 0) sweep(std::true_type{}, std::integral_constant<int, 0>{});
    else sweep(std::false_type{}, std::integral_constant<int, 2>{});
    acc = wave_group_combine(acc, G);
    if (live && gl == 0) {
        if (p.splits == 1) finish(p, acc, o);
        else reinterpret_cast<Store*>(p.partials)[wi] = acc.store();
    }
}

// cols: dims [rows | fast outputs]; a block of 512 threads = one tile of
// TW * V adjacent outputs (TW = p.group lanes: 64, 128 or 256) and one split
// of the rows.  The block's 512 / TW row groups interleave rows, each lane
// with U rows in flight, and combine in LDS in row-group order.
template <int V, int NR, class S, int AK, int OP>
__global__ __launch_bounds__(512) void reduce_cols_fast_kernel(const RedPlan p) {
    using A = Acc<AK, OP>;
    constexpr int U = fast_unroll<NR, V, true>();
    __shared__ typename A::Lds lds[512][V];  // 8-byte pairs: a bf16 tile (V = 8) takes 32 KiB
    const uint32_t TW = uint32_t(p.group);
    const uint32_t RG = 512u / TW;
    const uint32_t tl = threadIdx.x & (TW - 1);
    const uint32_t rg = threadIdx.x / TW;
    // split fastest: the splits of one column strip run side by side (measured
    // faster than tile-fastest, which reads whole rows per group of blocks:
    // 3.6 vs 3.1 TB/s on a [8192, 4096] float32 column sum)
    const uint32_t tile = fastdiv(blockIdx.x, p.div_splits);
    const uint32_t sp = blockIdx.x - tile * uint32_t(p.splits);
    const uint32_t cc = tile * TW + tl;
    const bool active = cc < p.nchunks_f;
    const uint32_t j0 = cc * V;
    const int nvalid = active ? min(V, int(p.lf - j0)) : 0;
    A acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v].init(p.identity);
    const uint32_t r0 = sp * p.per_split;
    const uint32_t r1 = min(p.red_count, r0 + p.per_split);
    // the rows of this block's split; FULL: the whole tile holds whole
    // chunks, so rows past the split's end are clamped to its last row and
    // their values dropped, and no load sits behind a divergent branch
    auto sweep = [&](auto full_c, auto ntm_c) {
        constexpr bool FULL = decltype(full_c)::value;
        constexpr int NTM = decltype(ntm_c)::value;
        for (uint32_t r = r0 + rg; r < r1; r += RG * U) {
            Reg R[U][NR][V];
            bool ok[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t rr = r + RG * uint32_t(u);
                ok[u] = rr < r1;
                const uint32_t idx[4] = {FULL ? min(rr, r1 - 1) : rr, 0u, 0u, 0u};
                if (FULL || ok[u]) run_spec<S, 0, V, NR, 2, FULL, NTM>(p, idx, j0, nvalid, R[u]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if constexpr (FULL) {
                    // a select, not a branch: a branch would let the compiler
                    // sink the clamped row's load into it and wait there
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        A t = acc[v];
                        t.add(R[u][S::result_reg][v]);
                        acc[v] = ok[u] ? t : acc[v];
                    }
                } else if (ok[u]) {
#pragma unroll
                    for (int v = 0; v < V; ++v)
                        if (v < nvalid) acc[v].add(R[u][S::result_reg][v]);
                }
            }
        }
    };
    const bool full_tile = (tile + 1) * TW <= p.nchunks_f && p.lf == p.nchunks_f * uint32_t(V);  // uniform
    const int ntm = spec_leaves_nt<S>(p);
    if (full_tile && ntm == 1) sweep(std::true_type{}, std::integral_constant<int, 1>{});
    else if (full_tile && ntm == 0) sweep(std::true_type{}, std::integral_constant<int, 0>{});
    else if (active) sweep(std::false_type{}, std::integral_constant<int, 2>{});
#pragma unroll
    for (int v = 0; v < V; ++v) lds[threadIdx.x][v] = acc[v].lds();
    __syncthreads();
    const bool merge_here = p.splits > 1 && p.counters != nullptr;  // uniform
    if (rg == 0 && active) {
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (v < nvalid) {
                A a;
                a.from_lds(lds[tl][v]);
                for (uint32_t g = 1; g < RG; ++g) {
                    A b;
                    b.from_lds(lds[g * TW + tl][v]);
                    a.merge(b);
                }
                const uint32_t o = j0 + uint32_t(v);
                if (p.splits == 1) {
                    finish(p, a, o);
                } else if (merge_here) {  // split-major, written through to memory
                    const Store st = a.store();
                    Store* dst = reinterpret_cast<Store*>(p.partials) + size_t(sp) * p.n_out + o;
                    __hip_atomic_store(&dst->a, st.a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(&dst->b, st.b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else if (p.split_major) {
                    reinterpret_cast<Store*>(p.partials)[size_t(sp) * p.n_out + o] = a.store();
                } else {
                    reinterpret_cast<Store*>(p.partials)[size_t(o) * p.splits + sp] = a.store();
                }
            }
        }
    }
    if (!merge_here) return;
    // the last of the tile's splits to arrive merges its partials: every
    // thread's partial stores complete, then one arrival count per block
    __shared__ uint32_t s_last;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    if (threadIdx.x == 0)
        s_last = __hip_atomic_fetch_add(p.counters + tile, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                 uint32_t(p.splits) - 1u;
    __syncthreads();
    if (!s_last) return;  // uniform in the block
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const uint32_t T = uint32_t(p.combine_group);
    const uint32_t per_pass = 512u / T;
    const uint32_t o0 = tile * TW * uint32_t(V);  // the tile's first output
    const uint32_t o1 = min(p.n_out, o0 + TW * uint32_t(V));
    for (uint32_t ob = o0; ob < o1; ob += per_pass) {
        const uint32_t o = ob + threadIdx.x / T;  // uniform over each group of T lanes
        A a = combine_partials<A, true>(p, o < o1 ? o : p.n_out, T, threadIdx.x & (T - 1));
        if (o < o1 && (threadIdx.x & (T - 1)) == 0) finish(p, a, o);
    }
    if (threadIdx.x == 0) __hip_atomic_store(p.counters + tile, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace xg

namespace xg {

using KFn = void (*)(const RedPlan);

struct RedKernels {
    KFn rows32, rows64, cols32, cols64;
    KFn rows_fast1, rows_fast2, cols_fast;  // compiled fast kernels (compact plans), or null
};

// specialised programs: unrolled x4, 32-bit indexing only (64-bit plans run the generic kernels)
template <int V, int NR, class S, int AK, int OP> RedKernels spec_kernels_for() {
    return RedKernels{reduce_rows_kernel<V, NR, S, false, AK, OP, 4>, nullptr,
                      reduce_cols_kernel<V, NR, S, false, AK, OP, 4>, nullptr,
                      reduce_rows_fast_kernel<V, NR, S, AK, OP, 1>, reduce_rows_fast_kernel<V, NR, S, AK, OP, 2>,
                      reduce_cols_fast_kernel<V, NR, S, AK, OP>};
}

// generic interpreter: V = 4, no unrolling
template <int AK, int OP> RedKernels dyn_kernels_for() {
    return RedKernels{reduce_rows_kernel<4, kMaxRegs, DynSpec, false, AK, OP, 1>,
                      reduce_rows_kernel<4, kMaxRegs, DynSpec, true, AK, OP, 1>,
                      reduce_cols_kernel<4, kMaxRegs, DynSpec, false, AK, OP, 1>,
                      reduce_cols_kernel<4, kMaxRegs, DynSpec, true, AK, OP, 1>,
                      nullptr, nullptr, nullptr};
}

// generic kernels per accumulator family (defined in reduce_dyn_*.hip)
bool dyn_reduce_float(int ak, int op, RedKernels* k);
bool dyn_reduce_int32(int ak, int op, RedKernels* k);
bool dyn_reduce_int64(int ak, int op, RedKernels* k);
// split combine: a thread per output, or (wide) a block per output
KFn combine_kernel(int ak, int op, bool wide);

}  // namespace xg
