// multi.cpp — many elementwise expressions in few launches.
//
// Plans that run the same compiled program at the same normalised rank are
// fused: their compact plans and the prefix sums of their work units go to
// the device in one asynchronous upload, and one ew_multi_kernel launch
// covers all of them.  Plans without a compiled program run one by one on
// the generic interpreter.
#include <map>
#include <utility>
#include <vector>

#include "ew.hpp"
#include "staging.hpp"
#include "xg_api_internal.hpp"

namespace xg {

namespace {
size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }
}  // namespace

int launch_ew_many(const EwPlan* const* plans, const int* V, size_t n, hipStream_t stream) {
    int launches = 0;
    std::map<std::pair<int, int>, std::vector<size_t>> groups;  // (program, rank) -> plans
    const int cus = device_cus();
    for (size_t i = 0; i < n; ++i) {
        const EwPlan& p = *plans[i];
        if (p.units == 0) continue;
        const int si = ew_spec_index(p, V[i]);
        if (si < 0) {
            check_hip(launch_ew(p, V[i], cus, stream, nullptr), "elementwise launch");
            ++launches;
        } else {
            groups[{si, p.rank}].push_back(i);
        }
    }
    // one upload for every fused group: [prefix (n+1) x u64][n x CompactPlan] per group
    size_t bytes = 0;
    for (auto& g : groups) {
        if (g.second.size() < 2) continue;
        bytes = align_up(bytes, 256) + align_up((g.second.size() + 1) * sizeof(uint64_t), 256) +
                g.second.size() * sizeof(CompactPlan);
    }
    Staging* st = nullptr;
    char* h = nullptr;
    if (bytes) {
        st = &staging_for(stream);
        h = st->host(bytes, stream);
    }
    struct Pending {
        int si, rank;
        size_t off_prefix, off_plans;
        uint32_t count;
        uint64_t total;
    };
    std::vector<Pending> todo;
    size_t off = 0;
    for (auto& g : groups) {
        const auto& idx = g.second;
        if (idx.size() < 2) {
            check_hip(launch_ew(*plans[idx[0]], V[idx[0]], cus, stream, nullptr), "elementwise launch");
            ++launches;
            continue;
        }
        off = align_up(off, 256);
        const size_t op = off;
        const size_t opl = op + align_up((idx.size() + 1) * sizeof(uint64_t), 256);
        auto* prefix = reinterpret_cast<uint64_t*>(h + op);
        auto* cp = reinterpret_cast<CompactPlan*>(h + opl);
        uint64_t total = 0;
        for (size_t k = 0; k < idx.size(); ++k) {
            prefix[k] = total;
            cp[k] = make_compact(*plans[idx[k]]);
            total += plans[idx[k]]->units;
        }
        prefix[idx.size()] = total;
        off = opl + idx.size() * sizeof(CompactPlan);
        todo.push_back(Pending{g.first.first, g.first.second, op, opl, uint32_t(idx.size()), total});
    }
    if (bytes) {
        char* d = st->upload(bytes, stream);
        for (const Pending& t : todo) {
            check_hip(launch_ew_multi(t.si, t.rank, reinterpret_cast<const CompactPlan*>(d + t.off_plans),
                                      reinterpret_cast<const uint64_t*>(d + t.off_prefix), t.count, t.total, stream),
                      "fused elementwise launch");
            ++launches;
        }
    }
    return launches;
}

}  // namespace xg
