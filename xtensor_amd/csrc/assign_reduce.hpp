// assign_reduce.hpp — out = e over a 2-D row-major contiguous output, with a
// reduction of every row and / or every column of the stored values in the
// same pass (no counterpart in xtensor, where `c = e; s = sum(c, {1});
// m = amax(c, {0})` traverses c three times: core/xassign.hpp:208-226 then
// reducers/xreducer.hpp:1778-1828 twice).
//
// A workgroup owns a band of whole rows, its eight waves a row each at a
// time.  A wave's lanes take the row's work units (V consecutive elements),
// evaluate and store them as the elementwise kernel does, and fold the values
// into a row accumulator and into one accumulator per (column unit, element)
// they own; a wave-wide combine in lane order finishes the row.  After the
// band, the waves' column accumulators are merged in LDS and go to the
// workspace as the band's partials ([band][column], 16 bytes each), and the
// split-combine kernel merges them in band order.  Compiled at run time for each program (the expression's
// micro-ops are constants, as in the elementwise kernels).
#pragma once

#include "assign_reduce_plan.hpp"
#include "ew_kernels.hpp"
#include "reduce_kernels.hpp"

namespace xg {

template <class A> __device__ void finish_row(const FusedRC& f, A acc, uint32_t row) {
    if (f.row_has_initial) {
        A ini;
        ini.init(f.row_initial);
        ini.merge(acc);
        acc = ini;
    }
    const Reg v = acc.value(f.row_vdt);
    const Reg ov = cvt(v, f.row_rdt, f.row_dtype);
    const int esz = dtype_size(f.row_dtype);
    store_raw(f.row_out + int64_t(row) * f.row_stride * esz, esz, to_storage(ov, f.row_dtype));
}

// RAK / CAK < 0: no row / column reduction.  A wave takes whole rows (the
// band's rows round-robin over the eight waves), UL work units per lane, all
// of a row's loads in flight at once; the row's value is a wave-wide combine
// (shuffles only).  The eight waves' column accumulators meet in LDS as a
// tree (w0 w1) (w2 w3) ..., then pairs of pairs, and wave 0 writes the
// band's partials.
template <int V, int NR, class S, int RANK, int RAK, int ROP, int CAK, int COP, int UL>
__global__ __launch_bounds__(512) void assign_rc_kernel(const AssignRC a) {
    constexpr bool ROWS = RAK >= 0, COLS = CAK >= 0;
    using RA = Acc<(ROWS ? RAK : int(AK_F64C)), (ROWS ? ROP : int(RO_ADD))>;
    using CA = Acc<(COLS ? CAK : int(AK_F64C)), (COLS ? COP : int(RO_ADD))>;
    const EwPlan& plan = a.plan;
    const FusedRC& f = a.f;
    __shared__ Store xw[4][COLS ? 64 * UL * V : 1];
    const int lane = int(threadIdx.x & 63);
    const int wave = int(threadIdx.x >> 6);
    const uint32_t r0 = blockIdx.x * f.rows_per_band;
    const uint32_t r1 = min(f.rows, r0 + f.rows_per_band);
    constexpr int odt = S::out_dtype;
    constexpr int esz = dtype_size(odt);
    CA cacc[UL][V];
    if constexpr (COLS) {
#pragma unroll
        for (int k = 0; k < UL; ++k)
#pragma unroll
            for (int v = 0; v < V; ++v) cacc[k][v].init(f.col_identity);
    }
    for (uint32_t row = r0 + uint32_t(wave); row < r1; row += 8) {  // uniform in a wave
        Reg R[UL][NR][V];
        int nvalid[UL];
        uint32_t j0s[UL];
        uint32_t idxs[UL][4];
#pragma unroll
        for (int k = 0; k < UL; ++k) {  // every unit's loads first
            const uint32_t cu = uint32_t(lane) + uint32_t(k) * 64u;
            nvalid[k] = 0;
            if (cu < f.upr) {
                // the plan's unit decomposition, as spec_unit
                const uint32_t u = row * f.upr + cu;
                const uint32_t prow = fastdiv(u, plan.div_chunks);
                const uint32_t j0 = (u - prow * plan.nchunks) * V;
                uint32_t* idx = idxs[k];
                idx[0] = idx[1] = idx[2] = idx[3] = 0;
                uint32_t r = prow;
#pragma unroll
                for (int d = RANK - 2; d >= 1; --d) {
                    const uint32_t q = fastdiv(r, plan.div[d]);
                    idx[d] = r - q * plan.shape32[d];
                    r = q;
                }
                if constexpr (RANK >= 2) idx[0] = r;
                j0s[k] = j0;
                nvalid[k] = min(V, int(plan.inner - j0));
                run_spec<S, 0, V, NR, RANK>(plan, idxs[k], j0, nvalid[k], R[k]);
            }
        }
        RA racc;
        if constexpr (ROWS) racc.init(f.row_identity);
#pragma unroll
        for (int k = 0; k < UL; ++k) {
            if (nvalid[k] <= 0) continue;
            char* p = plan.out_ptr + int64_t(coff<RANK, true>(plan.cst[kMaxLeaves], idxs[k], j0s[k])) * esz;
            if (S::out_mode == LM_CONTIG && nvalid[k] == V) {  // rows 16-byte aligned: vector stores
                store_vec_rt(out_nt(plan), p, esz, odt, R[k][S::result_reg]);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v)
                    if (v < nvalid[k]) store_raw(p + int64_t(v) * esz, esz, to_storage(R[k][S::result_reg][v], odt));
            }
#pragma unroll
            for (int v = 0; v < V; ++v) {
                if (v < nvalid[k]) {
                    if constexpr (ROWS) racc.add(R[k][S::result_reg][v]);
                    if constexpr (COLS) cacc[k][v].add(R[k][S::result_reg][v]);
                }
            }
        }
        if constexpr (ROWS) {
            const RA t = group_combine(racc, 64, nullptr);  // one wave: shuffles in lane order
            if (lane == 0) finish_row(f, t, row);
        }
    }
    if constexpr (COLS) {
        constexpr int NC = 64 * UL * V;
        // slot of (unit k, element v) of this lane: the same for every wave
        auto slot = [&](int k, int v) { return (k * 64 + lane) * V + v; };
        // level m: waves with bit m set (and the lower bits clear) hand their
        // accumulators to the wave 2^m below them
#pragma unroll
        for (int m = 1; m < 8; m <<= 1) {
            if ((wave & (2 * m - 1)) == m) {
#pragma unroll
                for (int k = 0; k < UL; ++k)
#pragma unroll
                    for (int v = 0; v < V; ++v) xw[wave / (2 * m)][slot(k, v)] = cacc[k][v].store();
            }
            __syncthreads();
            if ((wave & (2 * m - 1)) == 0) {
#pragma unroll
                for (int k = 0; k < UL; ++k)
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        CA o;
                        o.load(xw[wave / (2 * m)][slot(k, v)]);
                        cacc[k][v].merge(o);
                    }
            }
            __syncthreads();
        }
        if (wave == 0) {
            Store* dst = f.col_partials + uint64_t(blockIdx.x) * f.cols;
#pragma unroll
            for (int k = 0; k < UL; ++k) {
                const uint32_t cu = uint32_t(lane) + uint32_t(k) * 64u;
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    const uint32_t col = cu * uint32_t(V) + uint32_t(v);
                    if (cu < f.upr && col < f.cols) dst[col] = cacc[k][v].store();
                }
            }
        }
        (void)NC;
    }
}

}  // namespace xg
