// ew.hip — elementwise assignment kernels (out = f(leaves...)) for gfx950.
//
// Replaces xtensor's linear / strided-loop / stepper assigners
// (core/xassign.hpp:644-695, 701-849, 855-1358): instead of a host loop
// over a stepper, the planner hands over a normalised iteration space
// (outer dims x one inner row) and every work-item evaluates the micro-op
// program on V consecutive elements of one inner row.  Leaves contiguous
// along the row are read with 16-byte vector loads, leaves broadcast along
// the row (stride 0) with one load per work-item, others element by element;
// the output is written the same way.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "chunk.hpp"
#include "ew.hpp"
#include "planner.hpp"

namespace xg {

template <int V, bool I64>
__device__ __forceinline__ void store_out(const EwPlan& plan, int mode, int dt, const uint32_t* idx, uint32_t j0,
                                          int nvalid, const Reg (&val)[V]) {
    const int esz = dtype_size(dt);
    auto off = unit_offset<I64>(plan.out_stride, idx, plan.rank, j0);
    char* p = plan.out_ptr + int64_t(off) * esz;
    if (mode == LM_CONTIG && nvalid == V) {
        store_vec<V>(p, esz, dt, val);
    } else {
        const int64_t os = plan.out_stride[plan.rank - 1];
#pragma unroll
        for (int v = 0; v < V; ++v)
            if (v < nvalid) store_raw(p + int64_t(v) * os * esz, esz, to_storage(val[v], dt));
    }
}

template <int V, int NR, class S, bool I64>
__global__ __launch_bounds__(256) void ew_kernel(const EwPlan plan) {
    const uint32_t nunits = uint32_t(plan.units);
    const uint32_t step = gridDim.x * 256u;
    for (uint32_t u = blockIdx.x * 256u + threadIdx.x; u < nunits; u += step) {
        const uint32_t row = fastdiv(u, plan.div_chunks);
        const uint32_t j0 = (u - row * plan.nchunks) * V;
        uint32_t idx[kMaxRank];
        uint32_t r = row;
#pragma unroll
        for (int k = kMaxRank - 2; k >= 0; --k) {
            if (k <= plan.rank - 2) {
                uint32_t q = fastdiv(r, plan.div[k]);
                idx[k] = r - q * uint32_t(plan.shape[k]);
                r = q;
            } else {
                idx[k] = 0;
            }
        }
        const int nvalid = min(V, int(plan.inner - j0));
        Reg res[V];
        eval_chunk<V, NR, S, I64>(plan.leaves, plan.prog, plan.rank, idx, j0, nvalid, res);
        if constexpr (S::is_static) store_out<V, I64>(plan, S::out_mode, S::out_dtype, idx, j0, nvalid, res);
        else store_out<V, I64>(plan, plan.out_mode, plan.out_dtype, idx, j0, nvalid, res);
    }
}

}  // namespace xg

// compiled specialisations (generated from specs_list.cpp at build time)
#include "specs.inc"

namespace xg {

namespace {

struct SpecEntry {
    const char* name;
    int V;
    SpecKey key;
    void (*k32)(const EwPlan);
    void (*k64)(const EwPlan);
};

template <class S> SpecKey key_of() {
    SpecKey k{};
    k.nops = S::nops;
    for (int i = 0; i < S::nops; ++i) k.ops[i] = S::ops[i];
    k.nleaves = S::nleaves;
    for (int l = 0; l < S::nleaves; ++l) {
        k.leaf_dtype[l] = S::leaf_dtype[l];
        k.leaf_mode[l] = S::leaf_mode[l];
    }
    k.out_dtype = uint8_t(S::out_dtype);
    k.out_mode = uint8_t(S::out_mode);
    k.result_reg = S::result_reg;
    return k;
}

std::vector<SpecEntry>& spec_table() {
    static std::vector<SpecEntry> t;
    static std::once_flag once;
    std::call_once(once, [] {
#define XG_REGISTER_SPEC(S) \
    t.push_back(SpecEntry{#S, S::V, key_of<S>(), ew_kernel<S::V, S::nregs, S, false>, ew_kernel<S::V, S::nregs, S, true>});
        XG_SPEC_LIST(XG_REGISTER_SPEC)
#undef XG_REGISTER_SPEC
    });
    return t;
}

int g_force_dynamic = 0;

}  // namespace

void ew_set_force_dynamic(int on) { g_force_dynamic = on; }

const char* ew_find_spec(const EwPlan& p, int V) {
    if (g_force_dynamic) return nullptr;
    SpecKey k = spec_key(p);
    for (auto& s : spec_table())
        if (s.V == V && spec_equal(s.key, k)) return s.name;
    return nullptr;
}

int ew_grid(uint64_t units, int cus) {
    uint64_t blocks = (units + 255) / 256;
    uint64_t cap = uint64_t(cus) * 16;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    return int(blocks);
}

hipError_t launch_ew(const EwPlan& p, int V, int cus, hipStream_t stream, const char** used) {
    if (p.units == 0) return hipSuccess;
    dim3 grid(ew_grid(p.units, cus)), block(256);
    if (!g_force_dynamic) {
        SpecKey k = spec_key(p);
        for (auto& s : spec_table()) {
            if (s.V == V && spec_equal(s.key, k)) {
                if (used) *used = s.name;
                hipLaunchKernelGGL(p.index64 ? s.k64 : s.k32, grid, block, 0, stream, p);
                return hipGetLastError();
            }
        }
    }
    if (used) *used = "dynamic";
    if (V == 8) {
        if (p.index64) hipLaunchKernelGGL((ew_kernel<8, kMaxRegs, DynSpec, true>), grid, block, 0, stream, p);
        else hipLaunchKernelGGL((ew_kernel<8, kMaxRegs, DynSpec, false>), grid, block, 0, stream, p);
    } else {
        if (p.index64) hipLaunchKernelGGL((ew_kernel<4, kMaxRegs, DynSpec, true>), grid, block, 0, stream, p);
        else hipLaunchKernelGGL((ew_kernel<4, kMaxRegs, DynSpec, false>), grid, block, 0, stream, p);
    }
    return hipGetLastError();
}

}  // namespace xg
