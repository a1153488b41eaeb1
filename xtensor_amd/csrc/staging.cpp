// staging.cpp — pinned upload buffers for fused launches.
#include "staging.hpp"

#include <map>
#include <memory>
#include <mutex>

#include "xg_api_internal.hpp"

namespace xg {

Staging::~Staging() {
    if (copied_) (void)hipEventSynchronize(copied_);
    if (h_) (void)hipHostFree(h_);
    if (d_) (void)hipFree(d_);
    if (copied_) (void)hipEventDestroy(copied_);
}

char* Staging::host(size_t bytes, hipStream_t stream) {
    if (!copied_) check_hip(hipEventCreateWithFlags(&copied_, hipEventDisableTiming), "hipEventCreate");
    // the previous upload must have read the host buffer before it is rewritten
    if (pending_) check_hip(hipEventSynchronize(copied_), "hipEventSynchronize");
    pending_ = false;
    if (bytes > cap_) {
        // kernels still queued on `stream` may read the old device buffer
        check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
        if (h_) check_hip(hipHostFree(h_), "hipHostFree");
        if (d_) check_hip(hipFree(d_), "hipFree");
        h_ = nullptr;
        d_ = nullptr;
        cap_ = 0;
        size_t n = size_t(64) << 10;
        while (n < bytes) n <<= 1;
        if (hipHostMalloc(reinterpret_cast<void**>(&h_), n, hipHostMallocDefault) != hipSuccess)
            raise(XG_OUT_OF_MEMORY, "pinned staging allocation failed");
        if (hipMalloc(reinterpret_cast<void**>(&d_), n) != hipSuccess) raise(XG_OUT_OF_MEMORY, "staging allocation failed");
        cap_ = n;
    }
    return h_;
}

char* Staging::upload(size_t bytes, hipStream_t stream) {
    if (bytes > cap_) raise(XG_INVALID_ARGUMENT, "staging upload larger than the reserved buffer");
    if (bytes) check_hip(hipMemcpyAsync(d_, h_, bytes, hipMemcpyHostToDevice, stream), "hipMemcpyAsync");
    check_hip(hipEventRecord(copied_, stream), "hipEventRecord");
    pending_ = true;
    return d_;
}

namespace {
struct Key {
    int dev;
    void* stream;
    bool operator<(const Key& o) const { return dev != o.dev ? dev < o.dev : stream < o.stream; }
};
std::mutex g_mu;
std::map<Key, std::unique_ptr<Staging>>& table() {
    static auto* t = new std::map<Key, std::unique_ptr<Staging>>();  // lives until process exit
    return *t;
}
}  // namespace

Staging& staging_for(hipStream_t stream) {
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    std::lock_guard<std::mutex> lk(g_mu);
    auto& p = table()[Key{dev, reinterpret_cast<void*>(stream)}];
    if (!p) p = std::make_unique<Staging>();
    return *p;
}

}  // namespace xg
