// chunk.hpp — evaluate a micro-op program on V consecutive elements along the
// last dimension of a normalised iteration space.  Shared by the
// elementwise, reduction and scan kernels, so an expression fused into a
// reduction or scan is evaluated by exactly the code that assigns it.
#pragma once

#include "interp.hpp"

namespace xg {

struct DynSpec {
    static constexpr bool is_static = false;
};

template <bool I64> struct IdxT { using type = int32_t; };
template <> struct IdxT<true> { using type = int64_t; };

// element offset for multi-index idx[0..rank-2] and position j0 of the last dim
template <bool I64>
__device__ __forceinline__ typename IdxT<I64>::type unit_offset(const int64_t* stride, const uint32_t* idx, int rank,
                                                                uint32_t j0) {
    using I = typename IdxT<I64>::type;
    I off = I(j0) * I(stride[rank - 1]);
#pragma unroll
    for (int k = 0; k < kMaxRank - 1; ++k)
        if (k < rank - 1) off += I(idx[k]) * I(stride[k]);
    return off;
}

template <int V, bool I64>
__device__ __forceinline__ void load_leaf(const LeafView& lv, int mode, int dt, const uint32_t* idx, int rank,
                                          uint32_t j0, int nvalid, Reg (&out)[V]) {
    const int esz = dtype_size(dt);
    auto off = unit_offset<I64>(lv.stride, idx, rank, j0);
    const char* p = lv.ptr + int64_t(off) * esz;
    const bool nt = lm_nt(mode);
    mode = lm_base(mode);
    if (mode == LM_BCAST) {
        Reg x = from_storage(load_raw(p, esz), dt);
#pragma unroll
        for (int v = 0; v < V; ++v) out[v] = x;
    } else if (mode == LM_CONTIG && nvalid == V) {
        if (nt) load_vec<V, true>(p, esz, dt, out);
        else load_vec<V>(p, esz, dt, out);
    } else {
        const int64_t is = lv.stride[rank - 1];
        if (nt) {
#pragma unroll
            for (int v = 0; v < V; ++v)
                out[v] = v < nvalid ? from_storage(load_raw<true>(p + int64_t(v) * is * esz, esz), dt) : 0;
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) out[v] = v < nvalid ? from_storage(load_raw(p + int64_t(v) * is * esz, esz), dt) : 0;
        }
    }
}

// compile-time walk over a specialised program
template <class S, int I, int V, int NR, bool I64>
__device__ __forceinline__ void run_static(const LeafView* leaves, const uint64_t* kconst, int rank,
                                           const uint32_t* idx, uint32_t j0, int nvalid, Reg (&R)[NR][V]) {
    if constexpr (I < S::nops) {
        constexpr MicroOp m = S::ops[I];
        if constexpr (m.op == MO_LOAD) {
            load_leaf<V, I64>(leaves[m.aux], S::leaf_mode[m.aux] | (leaves[m.aux].mode & LM_NT), S::leaf_dtype[m.aux], idx, rank, j0, nvalid,
                              R[m.dst]);
        } else if constexpr (m.op == MO_GATHER) {
            const LeafView& lv = leaves[m.aux];
            gather_vals<V>(lv.ptr, S::leaf_dtype[m.aux], lv.stride[0], lv.stride[1], nvalid, R[m.a], R[m.dst]);
        } else {
            Reg tmp[V];
#pragma unroll
            for (int v = 0; v < V; ++v) tmp[v] = exec_scalar(m, R[m.a][v], R[m.b][v], R[m.c][v], kconst);
#pragma unroll
            for (int v = 0; v < V; ++v) R[m.dst][v] = tmp[v];
        }
        run_static<S, I + 1, V, NR, I64>(leaves, kconst, rank, idx, j0, nvalid, R);
    }
}

// A kernel's by-value plan argument, read where the launch put it (the
// kernel-argument segment; the first argument sits at offset 0).  Indexing
// a by-value struct argument with run-time operands (leaves[m.aux],
// ops[pc]) makes the compiler copy the whole struct into scratch memory
// first, ~3 KB per lane for these plans; reads through the segment pointer
// stay scalar loads.  Kernels instantiated on compiled programs index the
// plan with constants only and keep the argument itself.
template <class S, class P>
__device__ __forceinline__ const P& kernel_plan(const P& arg) {
    if constexpr (S::is_static) return arg;
    else return *(const P*)__builtin_amdgcn_kernarg_segment_ptr();
}

// The dynamic interpreter calls one out-of-line function per micro-op so
// that kernels instantiated on it stay small.
static __device__ __noinline__ Reg exec_scalar_dyn(MicroOp m, Reg a, Reg b, Reg c, const uint64_t* kconst) {
    return exec_scalar(m, a, b, c, kconst);
}

// Evaluates the program; returns the result values in `res`.
template <int V, int NR, class S, bool I64>
__device__ __forceinline__ void eval_chunk(const LeafView* leaves, const Program& prog, int rank, const uint32_t* idx,
                                           uint32_t j0, int nvalid, Reg (&res)[V]) {
    Reg R[NR][V];
    if constexpr (S::is_static) {
        run_static<S, 0, V, NR, I64>(leaves, prog.consts, rank, idx, j0, nvalid, R);
#pragma unroll
        for (int v = 0; v < V; ++v) res[v] = R[S::result_reg][v];
    } else {
        for (int pc = 0; pc < prog.nops; ++pc) {
            const MicroOp m = prog.ops[pc];
            Reg tmp[V];
            if (m.op == MO_LOAD) {
                const LeafView& lv = leaves[m.aux];
                load_leaf<V, I64>(lv, lv.mode, lv.dtype, idx, rank, j0, nvalid, tmp);
            } else if (m.op == MO_GATHER) {
                const LeafView& lv = leaves[m.aux];
                gather_vals<V>(lv.ptr, lv.dtype, lv.stride[0], lv.stride[1], nvalid, R[m.a], tmp);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) tmp[v] = exec_scalar_dyn(m, R[m.a][v], R[m.b][v], R[m.c][v], prog.consts);
            }
#pragma unroll
            for (int v = 0; v < V; ++v) R[m.dst][v] = tmp[v];
        }
#pragma unroll
        for (int v = 0; v < V; ++v) res[v] = R[prog.result_reg][v];
    }
}

}  // namespace xg
