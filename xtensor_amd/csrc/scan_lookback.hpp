// scan_lookback.hpp — single-pass scans of long contiguous rows with
// decoupled look-back (Merrill & Garland, "Single-pass Parallel Prefix Scan
// with Decoupled Look-back", NVIDIA NVR-2016-002).
//
// accumulator_impl (reducers/xaccumulator.hpp:215-297, the flattened form
// :343-367) runs res[k] = res[k-1] + res[k] along a row.  A row of millions
// of elements (the flattened cumsum of a large array) gives one workgroup to
// the direct rows kernel; here the row is cut into chunks of TILE elements
// and every chunk is a workgroup:
//   * a ticket from an atomic counter orders the chunks, so a workgroup only
//     ever waits for chunks whose workgroups started before it;
//   * the workgroup loads its chunk, scans it locally (the tile scheme of the
//     direct rows kernel) and publishes the chunk total as an aggregate;
//   * one wave looks back over the 64 preceding chunks at once, waits until
//     those nearer than the nearest published inclusive prefix have published
//     their totals, and merges them back to that prefix (the row's first
//     chunk always publishes one); then it publishes its own inclusive prefix;
//   * every element's result is that exclusive prefix merged with its local
//     prefix.
// Publication: for accumulators that fit 64 bits with a flag (float32 sums
// with compensation, 32-bit values) the value and its flag are one 64-bit
// word written and read with single atomic accesses at device scope, so no
// fence is needed; for the others the value goes first (device-scope atomic
// stores), then a wait for those stores and a release fence, then the flag,
// and the reader acquires after seeing the flag.  Device-scope atomic loads
// and stores bypass the non-coherent per-XCD L2.  Every wait is bounded:
// after kSpinLimit polls the chunk counts as published with the identity
// and the kernel records the timeout in ctrl[1] (a wrong result, never a
// hang).  The look-back merges in an order that depends on timing, so float
// results can differ in the last bits between runs (integers are exact).
//
// Included by scan_kernels.hpp, after ScanIn / ScanInDtype / the DPP helpers.
#pragma once

#include <hip/hip_runtime.h>

#include "acc.hpp"
#include "scan.hpp"

namespace xg {

constexpr uint32_t kLbAggregate = 1, kLbPrefix = 2;
constexpr int kSpinLimit = 1 << 20;

template <class T> __device__ __forceinline__ T lb_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T> __device__ __forceinline__ void lb_store(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// How an accumulator is published: packed with its flag in one 64-bit word,
// or as a Store beside a separate flag word.
template <int AK> struct LbPacked { static constexpr bool value = false; };
template <> struct LbPacked<AK_F32C> { static constexpr bool value = true; };
template <> struct LbPacked<AK_F32> { static constexpr bool value = true; };
template <> struct LbPacked<AK_I32> { static constexpr bool value = true; };
template <> struct LbPacked<AK_U32> { static constexpr bool value = true; };

// packed word: high 32 bits = the value (or a float32 sum), low 32 bits =
// flag, or for compensated float32 sums the compensation with its two low
// bits replaced by the flag (a relative change of 2^-22 of a term that is
// itself below half an ulp of the sum)
template <int AK, int OP> __device__ __forceinline__ uint64_t lb_pack(const Acc<AK, OP>& a, uint32_t flag) {
    const Store st = a.store();
    const uint32_t hi = uint32_t(st.a);
    const uint32_t lo = AK == AK_F32C ? (uint32_t(st.b) & ~3u) | flag : flag;
    return (uint64_t(hi) << 32) | lo;
}
template <int AK, int OP> __device__ __forceinline__ Acc<AK, OP> lb_unpack(uint64_t w) {
    Store st{w >> 32, AK == AK_F32C ? uint64_t(uint32_t(w) & ~3u) : 0};
    Acc<AK, OP> a;
    a.load(st);
    return a;
}

template <int AK, int OP>
__device__ __forceinline__ void lb_publish(const LbScan& p, uint32_t chunk, uint32_t flag, const Acc<AK, OP>& a) {
    if constexpr (LbPacked<AK>::value) {
        lb_store(p.words + chunk, lb_pack(a, flag));
    } else {
        const Store st = a.store();
        Store* dst = (flag == kLbPrefix ? p.prefix : p.aggregate) + chunk;
        lb_store(&dst->a, st.a);
        lb_store(&dst->b, st.b);
        __builtin_amdgcn_s_waitcnt(0);  // the value's stores have completed
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        lb_store(p.flags + chunk, flag);
    }
}

// Poll chunk `chunk`'s publication: returns its flag (0 = nothing yet) and
// the value that goes with it.
template <int AK, int OP>
__device__ __forceinline__ uint32_t lb_poll(const LbScan& p, uint32_t chunk, Acc<AK, OP>* v) {
    if constexpr (LbPacked<AK>::value) {
        const uint64_t w = lb_load(p.words + chunk);
        const uint32_t f = uint32_t(w) & 3u;
        if (f) *v = lb_unpack<AK, OP>(w);
        return f;
    } else {
        const uint32_t f = lb_load(p.flags + chunk);
        if (f) {
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const Store* src = (f == kLbPrefix ? p.prefix : p.aggregate) + chunk;
            Store st{lb_load(&src->a), lb_load(&src->b)};
            v->load(st);
        }
        return f;
    }
}

// Exclusive prefix of chunk `c` (> 0) of the row whose first chunk is
// `first`, in every lane of the calling wave.
template <int AK, int OP>
__device__ __forceinline__ Acc<AK, OP> lb_lookback(const LbScan& p, uint32_t first, uint32_t c, const Acc<AK, OP>& id) {
    using A = Acc<AK, OP>;
    const int lane = threadIdx.x & 63;
    A excl = id;
    int64_t nearest = int64_t(c) - 1;  // the window's nearest chunk (row-relative)
    while (true) {
        const int64_t j = nearest - lane;
        const bool valid = j >= 0;
        A v = id;
        uint32_t f = valid ? 0u : kLbPrefix;
        int spins = 0;
        while (true) {
            if (f == 0) f = lb_poll<AK, OP>(p, first + uint32_t(j), &v);
            // wait only for the chunks nearer than the nearest published prefix
            const uint64_t pm0 = __ballot(valid && f == kLbPrefix);
            const int lim = pm0 ? __builtin_ctzll(pm0) : 64;
            if (__ballot(f == 0 && lane < lim) == 0) break;
            if (++spins >= kSpinLimit) {  // bounded: count the timeout, take the identity
                if (f == 0) {
                    v = id;
                    f = kLbPrefix;
                }
                if (lane == 0) atomicAdd(p.ctrl + 1, 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        const uint64_t pm = __ballot(valid && f == kLbPrefix);
        // lanes up to the nearest published prefix take part (all if none)
        const int stop = pm ? __builtin_ctzll(pm) : 63;
        A t = (valid && lane <= stop) ? v : id;
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            A o = t.shfl_xor(m);
            t.merge(o);
        }
        t.merge(excl);  // this window lies before what was merged so far
        excl = t;
        if (pm || nearest < 64) break;
        nearest -= 64;
    }
    return excl;
}

// One chunk of TILE = 4 waves x U segments x 64 * V elements per workgroup
// (U = kLbSegments: 16384 float32 elements; with 4096-element chunks the
// chain of look-backs and the ticket counter measured slower on MI355X).
// MODE kLbOnePass: ticket, look-back, results.  The same chunks also run the
// reduce-then-scan form without tickets or waits, in three launches:
// kLbTotals writes each chunk's total to p.aggregate (chunk = blockIdx.x),
// scan_lookback_carry_kernel turns them into exclusive carries in place, and
// scan_chunks_apply_kernel scans each chunk from its carry in tiles (the
// second read of the input is served by the last-level cache when it fits).
constexpr int kLbOnePass = 0, kLbTotals = 1;
template <int AK, int OP, bool NATIVE, int MODE>
__global__ __launch_bounds__(256) void scan_lookback_kernel(const LbScan p) {
    using A = Acc<AK, OP>;
    using T = typename ScanIn<AK>::T;
    constexpr int V = int(16 / sizeof(T));
    constexpr int U = kLbSegments;
    constexpr uint32_t SEG = 64u * V;
    __shared__ uint32_t s_ticket;
    __shared__ Store wtot[4];
    __shared__ Store s_excl;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if constexpr (MODE == kLbOnePass) {
        if (threadIdx.x == 0) s_ticket = atomicAdd(p.ctrl, 1u);
        __syncthreads();
    }
    const uint32_t ticket = MODE == kLbOnePass ? s_ticket : blockIdx.x;
    if (ticket >= p.total) return;  // uniform in the block (the grid holds exactly `total`)
    const uint32_t row = ticket / p.cpr;
    const uint32_t c = ticket - row * p.cpr;
    const T* in = reinterpret_cast<const T*>(p.in) + int64_t(row) * p.in_row_stride;
    const bool nt_in = p.nt_in != 0, nt_out = p.nt_out != 0;
    const int odt = NATIVE ? int(ScanInDtype<T>::value) : p.out_dtype;
    const int rdt = NATIVE ? int(ScanInDtype<T>::value) : p.rdt;
    const int oesz = dtype_size(odt);
    char* out = p.out + int64_t(row) * p.out_row_stride * oesz;
    A id;
    id.init(p.identity);
    const uint32_t wbase = c * p.tile + uint32_t(wave) * SEG * U;
    // a whole chunk (all but a row's last) loads without a branch between the loads
    const bool whole = (c + 1u) * p.tile <= p.len;  // uniform in the block
    if constexpr (MODE == kLbTotals) {
        // the chunk's total only, streaming G segments a step (fewer registers
        // than holding the chunk: more waves in flight); lanes in order, then
        // the waves; default cache policy, so the second read can come from
        // the last-level cache
        constexpr int G = 8;
        A t = id;
#pragma unroll 1
        for (int g = 0; g < U; g += G) {
            T xs[G][V];
            int nvs[G];
#pragma unroll
            for (int k = 0; k < G; ++k) {
                const uint32_t j = wbase + uint32_t(g + k) * SEG + uint32_t(lane * V);
                nvs[k] = j < p.len ? min(V, int(p.len - j)) : 0;
                if (whole || nvs[k] == V) load_row<V, false>(in + j, xs[k]);
                else load_row_partial<V>(in + j, true, nvs[k], xs[k]);
            }
#pragma unroll
            for (int k = 0; k < G; ++k)
#pragma unroll
                for (int v = 0; v < V; ++v)
                    if (v < nvs[k]) t.add(reg_put<T>(xs[k][v]));
        }
        t = lane_value<63>(wave_inclusive_scan(t, id));
        if (lane == 0) wtot[wave] = t.store();
        __syncthreads();
        if (threadIdx.x == 0) {
            A agg = id;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                A o;
                o.load(wtot[w]);
                agg.merge(o);
            }
            p.aggregate[ticket] = agg.store();
        }
        return;
    }
    T x[U][V];
    int nv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const uint32_t j = wbase + uint32_t(u) * SEG + uint32_t(lane * V);
        nv[u] = j < p.len ? min(V, int(p.len - j)) : 0;
    }
    if (whole && nt_in) {
#pragma unroll
        for (int u = 0; u < U; ++u) load_row<V, true>(in + wbase + uint32_t(u) * SEG + uint32_t(lane * V), x[u]);
    } else if (whole) {
#pragma unroll
        for (int u = 0; u < U; ++u) load_row<V, false>(in + wbase + uint32_t(u) * SEG + uint32_t(lane * V), x[u]);
    } else {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t j = wbase + uint32_t(u) * SEG + uint32_t(lane * V);
            if (nv[u] == V) load_row<V, false>(in + j, x[u]);
            else load_row_partial<V>(in + j, true, nv[u], x[u]);
        }
    }
    // wave-local exclusive prefixes of the segments, and the wave's total
    A local = id;
    A pre[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        pre[u] = local;
        if (wbase + uint32_t(u) * SEG >= p.len) continue;  // uniform in the wave
        A t = id;
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (v < nv[u]) t.add(reg_put<T>(x[u][v]));
        const A incl = wave_inclusive_scan(t, id);
        pre[u].merge(dpp_move<0x138, 0xf>(incl, id));  // wave_shr:1, lane 0 takes the identity
        local.merge(lane_value<63>(incl));
    }
    if (lane == 0) wtot[wave] = local.store();
    __syncthreads();
    A before = id;  // the waves before this one in the chunk
    A agg = id;     // the chunk's total
#pragma unroll
    for (int w = 0; w < 4; ++w) {
        A o;
        o.load(wtot[w]);
        if (w == wave) before = agg;
        agg.merge(o);
    }
    if (wave == 0) {
        const uint32_t first = row * p.cpr;
        A excl = id;
        if (c == 0) {
            if (lane == 0) lb_publish<AK, OP>(p, ticket, kLbPrefix, agg);
        } else {
            if (lane == 0) lb_publish<AK, OP>(p, ticket, kLbAggregate, agg);
            excl = lb_lookback<AK, OP>(p, first, c, id);
            if (lane == 0) {
                A incl = excl;
                incl.merge(agg);
                lb_publish<AK, OP>(p, ticket, kLbPrefix, incl);
            }
        }
        if (lane == 0) s_excl = excl.store();
    }
    __syncthreads();
    A start;
    start.load(s_excl);
    start.merge(before);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        if (wbase + uint32_t(u) * SEG >= p.len) continue;  // uniform in the wave
        A t = start;
        t.merge(pre[u]);
        Reg o[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            if (v < nv[u]) t.add(reg_put<T>(x[u][v]));
            const Reg val = t.value(rdt);
            if constexpr (NATIVE) o[v] = val;
            else o[v] = cvt(val, rdt, odt);
        }
        const uint32_t j = wbase + uint32_t(u) * SEG + uint32_t(lane * V);
        char* ptr = out + int64_t(j) * oesz;
        if (nv[u] == V) {
            store_vec_rt(nt_out, ptr, oesz, odt, o);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v)
                if (v < nv[u]) store_raw(ptr + int64_t(v) * oesz, oesz, o[v]);
        }
    }
}

// In place: the chunk totals of every row (p.aggregate[row * cpr + c])
// become exclusive carries.  A workgroup per row walks its totals in tiles of
// 256, a thread each: a wave scan, the wave totals through LDS, and the tile
// total carried into the next tile.
template <int AK, int OP>
__global__ __launch_bounds__(256) void scan_lookback_carry_kernel(const LbScan p) {
    using A = Acc<AK, OP>;
    __shared__ Store wtot[4];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    Store* t = p.aggregate + uint64_t(blockIdx.x) * p.cpr;
    A id;
    id.init(p.identity);
    A carry = id;
    for (uint32_t base = 0; base < p.cpr; base += 256u) {  // uniform in the block
        const uint32_t j = base + threadIdx.x;
        A x = id;
        if (j < p.cpr) x.load(t[j]);
        const A incl = wave_inclusive_scan(x, id);
        A excl = dpp_move<0x138, 0xf>(incl, id);  // wave_shr:1, lane 0 takes the identity
        if (lane == 63) wtot[wave] = incl.store();
        __syncthreads();
        A before = carry;
        A next = carry;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            A o;
            o.load(wtot[w]);
            if (w == wave) before = next;
            next.merge(o);
        }
        before.merge(excl);
        if (j < p.cpr) t[j] = before.store();
        carry = next;
        __syncthreads();  // wtot is written again for the next tile
    }
}

// Each chunk (a workgroup) scanned from its exclusive carry, streaming in
// tiles of 4 waves x 4 segments as the direct rows kernel does (only a tile's
// values in registers at a time).  SUM: the carry is the merge of the chunk
// totals before it in its row (p.aggregate, read from the L2: at most
// kLbSumCarries of them; thread t takes totals t, t + 256, ..., then a wave
// scan and the four waves in order); else p.aggregate already holds the
// exclusive carries (scan_lookback_carry_kernel).
constexpr uint32_t kLbSumCarries = 4096;
template <int AK, int OP, bool NATIVE, bool SUM>
__global__ __launch_bounds__(256) void scan_chunks_apply_kernel(const LbScan p) {
    using A = Acc<AK, OP>;
    using T = typename ScanIn<AK>::T;
    __shared__ Store wtot[4];
    const uint32_t chunk = blockIdx.x;
    if (chunk >= p.total) return;  // uniform in the block
    const uint32_t row = chunk / p.cpr;
    const uint32_t c = chunk - row * p.cpr;
    const int odt = NATIVE ? int(ScanInDtype<T>::value) : p.out_dtype;
    const int rdt = NATIVE ? int(ScanInDtype<T>::value) : p.rdt;
    const int oesz = dtype_size(odt);
    const int64_t off = int64_t(c) * p.tile;
    const T* in = reinterpret_cast<const T*>(p.in) + int64_t(row) * p.in_row_stride + off;
    char* out = p.out + (int64_t(row) * p.out_row_stride + off) * oesz;
    const uint32_t len = min(p.tile, p.len - c * p.tile);
    A id;
    id.init(p.identity);
    A carry;
    if constexpr (SUM) {
        const int lane = threadIdx.x & 63;
        const int wave = threadIdx.x >> 6;
        const Store* tot = p.aggregate + uint64_t(row) * p.cpr;
        A acc = id;
#pragma unroll 4
        for (uint32_t j = threadIdx.x; j < c; j += 256u) {
            A o;
            o.load(tot[j]);
            acc.merge(o);
        }
        acc = lane_value<63>(wave_inclusive_scan(acc, id));
        if (lane == 0) wtot[wave] = acc.store();
        __syncthreads();
        carry = id;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            A o;
            o.load(wtot[w]);
            carry.merge(o);
        }
        __syncthreads();  // wtot is used again by the tiles
    } else {
        carry.load(p.aggregate[chunk]);
    }
    scan_row_tiles<AK, OP, NATIVE, 4>(in, out, len, carry, id, false, p.nt_out != 0, odt, rdt,
                                      int(threadIdx.x >> 6), wtot);
}

// clears the ticket counter, the timeout count and every chunk's flag (a
// kernel, not a memset, so that the pair can be captured in a HIP graph)
template <bool PACKED>
__global__ __launch_bounds__(256) void scan_lookback_init_kernel(const LbScan p) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < 2) p.ctrl[i] = 0;
    if (i < p.total) {
        if constexpr (PACKED) p.words[i] = 0;
        else p.flags[i] = 0;
    }
}

}  // namespace xg
