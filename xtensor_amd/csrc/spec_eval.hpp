// spec_eval.hpp — evaluation of a compiled (constexpr) micro-op program on V
// consecutive elements, shared by the elementwise and reduction kernels.
//
// A plan type P provides `cst[l][4]` (int32 element strides of leaf l over a
// normalised rank <= 4 iteration space, innermost last), `leaf_ptr(p, l)`
// and `plan_consts(p)`.  Every loop unrolls at compile time and leaf types
// and access modes are constants, so each leaf read is one address
// computation and one (vector) load.
#pragma once

#include "ew.hpp"
#include "interp.hpp"
#include "reduce.hpp"

namespace xg {

template <int RANK, bool WITH_J>
__device__ __forceinline__ int32_t coff(const int32_t (&st)[4], const uint32_t (&idx)[4], uint32_t j0) {
    int32_t off = WITH_J ? int32_t(j0) * st[RANK - 1] : 0;
#pragma unroll
    for (int k = 0; k < RANK - 1; ++k) off += int32_t(idx[k]) * st[k];
    return off;
}

// FULL: every lane's V elements are in range (no per-lane tail test, so the
// loads of several chunks carry no divergent branch and stay in flight
// together)
template <int V, bool NT = false, bool FULL = false>
__device__ __forceinline__ void load_at(const char* p, int mode, int dt, int32_t is, int nvalid, Reg (&out)[V]) {
    const int esz = dtype_size(dt);
    if (mode == LM_BCAST) {
        Reg x = from_storage(load_raw(p, esz), dt);
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = x;
    } else if (mode == LM_CONTIG && (FULL || nvalid == V)) {
        load_vec<V, NT>(p, esz, dt, out);
    } else {
#pragma unroll
        for (int v = 0; v < V; ++v)
            out[v] = (FULL || v < nvalid) ? from_storage(load_raw<NT>(p + int64_t(v) * is * esz, esz), dt) : 0;
    }
}

// Plan accessors: single-expression kernels read an EwPlan, fused launches
// one CompactPlan per expression, reductions their RedPlan.
__device__ __forceinline__ const char* leaf_ptr(const EwPlan& p, int l) { return p.leaves[l].ptr; }
__device__ __forceinline__ const uint64_t* plan_consts(const EwPlan& p) { return p.prog.consts; }
__device__ __forceinline__ const char* leaf_ptr(const CompactPlan& p, int l) { return p.leaf[l]; }
__device__ __forceinline__ const uint64_t* plan_consts(const CompactPlan& p) { return p.consts; }
__device__ __forceinline__ const char* leaf_ptr(const RedPlan& p, int l) { return p.leaves[l].ptr; }
__device__ __forceinline__ const uint64_t* plan_consts(const RedPlan& p) { return p.prog.consts; }
// non-temporal loads for leaf l (a run-time flag of the plan: one compiled
// program serves streaming and cache-resident operands)
__device__ __forceinline__ bool leaf_nt(const EwPlan& p, int l) { return lm_nt(p.leaves[l].mode); }
__device__ __forceinline__ bool leaf_nt(const CompactPlan&, int) { return false; }
__device__ __forceinline__ bool leaf_nt(const RedPlan& p, int l) { return lm_nt(p.leaves[l].mode); }
// non-temporal stores of the output
__device__ __forceinline__ bool out_nt(const EwPlan& p) { return lm_nt(p.out_mode); }
__device__ __forceinline__ bool out_nt(const CompactPlan&) { return false; }

// HBM writes whole 64-byte granules: a store that covers part of a granule
// which is not in the caches makes the memory controller read the granule
// back and merge.  Measured on the [64,256,56,56] float32 view
// y[:, :, 0:56:2, :] (rows of 224 bytes every 448, non-temporal stores): a
// copy from a dense array into the view 54.2 us, 47.3 us when the rest of
// each row's last granule is read first; the affine clip from the same view
// of x into it 61.2 us, 54.4 us.  For an output
// flagged LM_FILL the lane storing a row's last (first) unit reads one word
// past the row's end (before its start) in the same granule, and ties it into
// the stored value with plan.fill_zero, which is always 0: the read completes
// before the store and the stored bytes are unchanged.  The read is issued
// before the unit's operand loads and its value is used only after them, so
// it adds no wait of its own.  The word shares a 64-byte granule, hence a
// mapped page, with a byte of the output; its value is never written
// anywhere.  One read per unit: a row that is a single unit reads past its end.
__device__ __forceinline__ uint32_t fill_read(const CompactPlan&, const char*, uint32_t, int, int) { return 0; }
__device__ __forceinline__ uint32_t fill_read(const EwPlan& plan, const char* p, uint32_t j0, int nvalid, int esz) {
    if (!lm_fill(plan.out_mode)) return 0;
    const uintptr_t a = reinterpret_cast<uintptr_t>(p);
    const uintptr_t e = a + uintptr_t(nvalid) * uintptr_t(esz);
    uintptr_t q = 0;
    if (j0 + uint32_t(nvalid) >= plan.inner && (e & 63)) q = e & ~uintptr_t(3);
    else if (j0 == 0 && (a & 63)) q = (a - 1) & ~uintptr_t(3);
    if (!q) return 0;
    return *reinterpret_cast<const __attribute__((address_space(1))) uint32_t*>(q);
}
// the read's value as a tie into the stored value: always 0
__device__ __forceinline__ Reg fill_tie(const CompactPlan&, uint32_t) { return 0; }
__device__ __forceinline__ Reg fill_tie(const EwPlan& plan, uint32_t w) { return Reg(w & plan.fill_zero); }

// FULL as for load_at; NTM: non-temporal leaf loads never (0), always for
// leaves that are not broadcast (1), or per the plan's leaf flags (2)
template <class S, int I, int V, int NR, int RANK, bool FULL = false, int NTM = 2, class P>
__device__ __forceinline__ void run_spec(const P& plan, const uint32_t (&idx)[4], uint32_t j0, int nvalid,
                                         Reg (&R)[NR][V]) {
    if constexpr (I < S::nops) {
        constexpr MicroOp m = S::ops[I];
        if constexpr (m.op == MO_LOAD) {
            constexpr int dt = S::leaf_dtype[m.aux];
            constexpr int mode = lm_base(S::leaf_mode[m.aux]);
            const int32_t off = coff<RANK, mode != LM_BCAST>(plan.cst[m.aux], idx, j0);
            const char* p = leaf_ptr(plan, m.aux) + int64_t(off) * dtype_size(dt);
            const bool nt = NTM == 2 ? leaf_nt(plan, m.aux) : NTM == 1;
            if (mode != LM_BCAST && nt)
                load_at<V, true, FULL>(p, mode, dt, plan.cst[m.aux][RANK - 1], nvalid, R[m.dst]);
            else
                load_at<V, false, FULL>(p, mode, dt, plan.cst[m.aux][RANK - 1], nvalid, R[m.dst]);
        } else if constexpr (m.op == MO_GATHER) {
            gather_vals<V>(leaf_ptr(plan, m.aux), S::leaf_dtype[m.aux], plan.leaves[m.aux].stride[0],
                           plan.leaves[m.aux].stride[1], nvalid, R[m.a], R[m.dst]);
        } else {
            Reg tmp[V];
#pragma unroll
            for (int v = 0; v < V; ++v) tmp[v] = exec_scalar(m, R[m.a][v], R[m.b][v], R[m.c][v], plan_consts(plan));
#pragma unroll
            for (int v = 0; v < V; ++v) R[m.dst][v] = tmp[v];
        }
        run_spec<S, I + 1, V, NR, RANK, FULL, NTM>(plan, idx, j0, nvalid, R);
    }
}

}  // namespace xg
