// noalloc.hip — xg_reduce_noalloc: xg_reduce for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops).
//
// A reduction's temporaries (split partials, the kept-dims mean of a
// variance, the evaluated expression of an immediate reduction, the pieces
// of a reduction past 32-bit indexing) come from `scratch`, one fixed array
// per device that lives as long as the process, instead of the library's
// per-stream workspace.  The call plans twice as xg_reduce does (a sizing
// pass, then the real one) and rejects a reduction whose temporaries exceed
// the array.  Kernels are the ones xg_reduce launches, except that a
// run-time compiled kernel is used only when this process has already loaded
// it (no compile, no module load); otherwise the ahead-of-time or generic
// kernels run.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>
#include <vector>

#include "jit.hpp"
#include "xg_api_internal.hpp"

namespace xg {
namespace {

constexpr size_t kScratchBytes = size_t(16) << 20;  // 16 MiB per device

// Device memory only; host code never names it except through
// hipGetSymbolAddress (its host-side name is a stand-in, not an address).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
std::vector<char*> g_ptr;  // device address of `scratch`, per device

}  // namespace

// Device address of the current device's scratch array, resolved once per
// device with that device current.
char* scratch_setup() {
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_mu);
    if (int(g_ptr.size()) <= dev) g_ptr.resize(size_t(dev) + 1, nullptr);
    if (!g_ptr[size_t(dev)]) {
        void* p = nullptr;
        check_hip(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) raise(XG_DEVICE_ERROR, "hipGetSymbolAddress(scratch) returned no address");
        g_ptr[size_t(dev)] = static_cast<char*>(p);
    }
    return g_ptr[size_t(dev)];
}

namespace {

char* scratch_block(size_t bytes, hipStream_t, void*) {
    if (bytes > kScratchBytes)
        raise(XG_OUT_OF_MEMORY, "xg_reduce_noalloc: the reduction needs " + std::to_string(bytes) +
                                    " bytes of scratch, more than its " + std::to_string(kScratchBytes) + " bytes");
    return scratch_setup();
}

}  // namespace
}  // namespace xg

using namespace xg;

extern "C" {

xg_status xg_reduce_noalloc(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, xg_stream_t stream) {
    return guard([&] {
        scratch_setup();
        const int prev = jit_set_thread_loaded_only(1);
        try {
            reduce_call(e, r, out, reinterpret_cast<hipStream_t>(stream), scratch_block, nullptr, false);
        } catch (...) {
            jit_set_thread_loaded_only(prev);
            throw;
        }
        jit_set_thread_loaded_only(prev);
    });
}

size_t xg_reduce_noalloc_max_bytes(void) { return kScratchBytes; }

}  // extern "C"
