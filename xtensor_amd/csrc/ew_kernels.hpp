// ew_kernels.hpp — the compiled-program elementwise kernel templates.
// Included by ew.hip (ahead-of-time specialisations) and by the sources
// that the run-time compiler builds for programs outside that table.
#pragma once

#include "chunk.hpp"
#include "spec_eval.hpp"

namespace xg {

// ---------------------------------------------------------------- compiled programs
// Kernels instantiated on a constexpr program, for normalised rank RANK <= 4
// and 32-bit offsets.  Every loop is unrolled at compile time, leaf types
// and access modes are constants, so each leaf read is one address
// computation and one (vector) load.

// Evaluate work unit `u` (V consecutive elements of one inner row) of a
// compiled plan and store it.
template <int V, int NR, class S, int RANK, class P>
__device__ __forceinline__ void spec_unit(const P& plan, uint32_t u) {
    const uint32_t row = fastdiv(u, plan.div_chunks);
    const uint32_t j0 = (u - row * plan.nchunks) * V;
    uint32_t idx[4] = {0, 0, 0, 0};
    uint32_t r = row;
#pragma unroll
    for (int k = RANK - 2; k >= 1; --k) {
        const uint32_t q = fastdiv(r, plan.div[k]);
        idx[k] = r - q * plan.shape32[k];
        r = q;
    }
    if constexpr (RANK >= 2) idx[0] = r;
    const int nvalid = min(V, int(plan.inner - j0));
    constexpr int odt = S::out_dtype;
    constexpr int esz = dtype_size(odt);
    const int32_t off = coff<RANK, true>(plan.cst[kMaxLeaves], idx, j0);
    char* p = plan.out_ptr + int64_t(off) * esz;
    uint32_t fill = 0;
    if constexpr (S::out_mode == LM_CONTIG) fill = fill_read(plan, p, j0, nvalid, esz);
    Reg R[NR][V];
    run_spec<S, 0, V, NR, RANK>(plan, idx, j0, nvalid, R);
    R[S::result_reg][0] |= fill_tie(plan, fill);
    if (S::out_mode == LM_CONTIG && nvalid == V) {
        store_vec_rt(out_nt(plan), p, esz, odt, R[S::result_reg]);
    } else {
        const int32_t os = plan.cst[kMaxLeaves][RANK - 1];
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (v < nvalid) store_raw(p + int64_t(v) * os * esz, esz, to_storage(R[S::result_reg][v], odt));
    }
}

template <int V, int NR, class S, int RANK>
__global__ __launch_bounds__(256) void ew_spec_kernel(const EwPlan plan) {
    const uint32_t nunits = uint32_t(plan.units);
    const uint32_t step = gridDim.x * 256u;
    for (uint32_t u = blockIdx.x * 256u + threadIdx.x; u < nunits; u += step) spec_unit<V, NR, S, RANK>(plan, u);
}

// UPT units per thread (u, u + 256, ...: every wave access stays one
// contiguous run), all their loads issued before the first store, for more
// bytes in flight per wave on large launches.
template <int V, int NR, class S, int RANK, int UPT>
__global__ __launch_bounds__(256) void ew_spec_multi_unit_kernel(const EwPlan plan) {
    const uint32_t nunits = uint32_t(plan.units);
    const uint32_t step = gridDim.x * 256u * UPT;
    for (uint32_t b = blockIdx.x * 256u * UPT + threadIdx.x; b < nunits; b += step) {
        Reg R[UPT][NR][V];
        uint32_t idx[UPT][4];
        uint32_t j0[UPT];
        int nvalid[UPT];
        uint32_t fill[UPT];
        constexpr int odt = S::out_dtype;
        constexpr int esz = dtype_size(odt);
#pragma unroll
        for (int k = 0; k < UPT; ++k) {
            const uint32_t u = b + uint32_t(k) * 256u;
            nvalid[k] = 0;
            if (u < nunits) {
                const uint32_t row = fastdiv(u, plan.div_chunks);
                j0[k] = (u - row * plan.nchunks) * V;
                idx[k][0] = idx[k][1] = idx[k][2] = idx[k][3] = 0;
                uint32_t r = row;
#pragma unroll
                for (int d = RANK - 2; d >= 1; --d) {
                    const uint32_t q = fastdiv(r, plan.div[d]);
                    idx[k][d] = r - q * plan.shape32[d];
                    r = q;
                }
                if constexpr (RANK >= 2) idx[k][0] = r;
                nvalid[k] = min(V, int(plan.inner - j0[k]));
                fill[k] = 0;
                if constexpr (S::out_mode == LM_CONTIG)
                    fill[k] = fill_read(plan, plan.out_ptr + int64_t(coff<RANK, true>(plan.cst[kMaxLeaves], idx[k], j0[k])) * esz,
                                        j0[k], nvalid[k], esz);
                run_spec<S, 0, V, NR, RANK>(plan, idx[k], j0[k], nvalid[k], R[k]);
                R[k][S::result_reg][0] |= fill_tie(plan, fill[k]);
            }
        }
#pragma unroll
        for (int k = 0; k < UPT; ++k) {
            if (nvalid[k] <= 0) continue;
            const int32_t off = coff<RANK, true>(plan.cst[kMaxLeaves], idx[k], j0[k]);
            char* p = plan.out_ptr + int64_t(off) * esz;
            if (S::out_mode == LM_CONTIG && nvalid[k] == V) {
                store_vec_rt(out_nt(plan), p, esz, odt, R[k][S::result_reg]);
            } else {
                const int32_t os = plan.cst[kMaxLeaves][RANK - 1];
#pragma unroll
                for (int v = 0; v < V; ++v)
                    if (v < nvalid[k]) store_raw(p + int64_t(v) * os * esz, esz, to_storage(R[k][S::result_reg][v], odt));
            }
        }
    }
}

// Many expressions of one compiled program in one launch: the work units of
// all plans are numbered consecutively (prefix[i] = first unit of plan i,
// prefix[n] = total); a thread finds its plan by binary search.
template <int V, int NR, class S, int RANK>
__global__ __launch_bounds__(256) void ew_multi_kernel(const CompactPlan* plans, const uint64_t* prefix, uint32_t n) {
    const uint64_t total = prefix[n];
    const uint64_t step = uint64_t(gridDim.x) * 256u;
    for (uint64_t u = uint64_t(blockIdx.x) * 256u + threadIdx.x; u < total; u += step) {
        uint32_t lo = 0, hi = n;  // prefix[lo] <= u < prefix[hi]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (prefix[mid] <= u) lo = mid;
            else hi = mid;
        }
        spec_unit<V, NR, S, RANK>(plans[lo], uint32_t(u - prefix[lo]));
    }
}

}  // namespace xg
