// server.cpp — the request loop behind xg_server_* (continuous batching).
//
// One loop thread per server.  Each turn it
//   1. retires finished groups (an event per stream) and releases their
//      bytes from the budget,
//   2. admits queued requests in arrival order while the bytes of admitted,
//      unfinished requests stay within the budget, preparing each (the
//      elementwise plan, or the row-kernel item) and filing it under its
//      class: template, dtype and, for elementwise templates, the compiled
//      program and normalised rank its plan runs,
//   3. gives every idle stream the class whose oldest request waited
//      longest, up to max_batch requests of it, as one fused launch:
//      elementwise classes through the many-plan kernel, row classes through
//      the persistent row kernel with a work queue.
// Requests that arrive while groups run are admitted as soon as bytes and
// streams free up, so work from different arrivals overlaps on the device.
// In plain mode every request runs alone through the same entry points a
// caller of the C ABI uses per expression (xg_assign / xg_reduce), which is
// the per-request baseline.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "batch.hpp"
#include "builder.hpp"
#include "ew.hpp"
#include "staging.hpp"
#include "xg_api_internal.hpp"

namespace xg {
namespace {

int64_t steady_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Job {
    xg_request r{};
    int64_t bytes = 0;
    int64_t admit_ns = 0;
    uint64_t seq = 0;  // admission order
    // elementwise templates
    EwPlan ew{};
    int V = 4;
    // row templates
    BItem item{};
    // EXPR / REDUCE: owned copies of the caller's descriptors
    std::vector<xg_node> nodes;
    std::vector<xg_tensor> leaves;
    std::vector<xg_scalar> consts;
    xg_tensor out{};
    xg_reduce_desc rd{};
    int64_t out_elems = 0;

    xg_expr expr() const {
        xg_expr e{};
        e.n_nodes = int32_t(nodes.size());
        e.n_leaves = int32_t(leaves.size());
        e.n_consts = int32_t(consts.size());
        e.nodes = nodes.data();
        e.leaves = leaves.data();
        e.consts = consts.data();
        return e;
    }
};

bool expr_kind(int tpl) { return tpl == XG_TPL_EXPR || tpl == XG_TPL_REDUCE; }

// bytes of device memory a strided view spans
int64_t view_bytes(const xg_tensor& t) {
    int64_t lo = 0, hi = 0;
    for (int k = 0; k < t.ndim; ++k) {
        if (t.shape[k] <= 0) return 0;
        const int64_t ext = (t.shape[k] - 1) * t.strides[k];
        (ext < 0 ? lo : hi) += ext;
    }
    return (hi - lo + 1) * dtype_size(t.dtype);
}

// EXPR / REDUCE: copy the caller's descriptors into the job and check them;
// returns the device bytes the request holds (its views' spans)
int64_t take_expr_request(Job& j) {
    const xg_request& r = j.r;
    if (!r.expr || !r.out_tensor || (r.tpl == XG_TPL_REDUCE && !r.reduce))
        raise(XG_INVALID_ARGUMENT, "expression requests need expr, out_tensor (and reduce)");
    const xg_expr& e = *r.expr;
    if (e.n_nodes < 1 || e.n_leaves < 0 || e.n_consts < 0 || !e.nodes || (e.n_leaves && !e.leaves) ||
        (e.n_consts && !e.consts))
        raise(XG_INVALID_ARGUMENT, "malformed expression");
    j.nodes.assign(e.nodes, e.nodes + e.n_nodes);
    j.leaves.assign(e.leaves, e.leaves + e.n_leaves);
    j.consts.assign(e.consts, e.consts + e.n_consts);
    j.out = *r.out_tensor;
    if (j.out.ndim < 0 || j.out.ndim > XG_MAX_RANK || j.out.dtype < 0 || j.out.dtype >= XG_NUM_DTYPES)
        raise(XG_INVALID_ARGUMENT, "malformed output tensor");
    if (r.tpl == XG_TPL_REDUCE) j.rd = *r.reduce;
    j.r.expr = nullptr;  // the copies are the request's from here on
    j.r.out_tensor = nullptr;
    j.r.reduce = nullptr;
    const xg_expr ce = j.expr();
    const ExprInfo info = analyse(&ce);  // throws on a malformed or unbroadcastable tree
    (void)info;
    j.out_elems = 1;
    for (int k = 0; k < j.out.ndim; ++k) j.out_elems *= j.out.shape[k];
    int64_t bytes = view_bytes(j.out);
    for (const xg_tensor& t : j.leaves) bytes += view_bytes(t);
    return bytes;
}

char* slot_workspace(size_t bytes, hipStream_t st, void*) { return static_cast<char*>(workspace(bytes, st)); }

// template, dtype, program, rank, V, solo (0: may be fused; else a unique id)
using ClassKey = std::tuple<int, int, int, int, int, uint64_t>;

// Elementwise requests above this many work units run as their own launch
// of the single-expression kernel (plan in kernel arguments, no table
// lookups); fusion pays off for the small ones, whose launch overhead
// dominates.
constexpr uint64_t kSoloUnits = uint64_t(1) << 18;

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    std::vector<Job> jobs;
    int64_t bytes = 0;
    int64_t launch_ns = 0;
    void* tmp = nullptr;  // plain-mode softmax temporaries (row maxima and sums)
    size_t tmp_bytes = 0;
};

int64_t request_elems(const xg_request& r) {
    switch (r.tpl) {
        case XG_TPL_AFFINE_CLIP: return r.shape[0] * r.shape[1] * r.shape[2] * r.shape[3];
        case XG_TPL_ROW_SUM:
        case XG_TPL_ROW_SOFTMAX: return r.shape[0] * r.shape[1];
        default: return r.shape[0];
    }
}

// validate a fixed template; return the device bytes the request holds while admitted
int64_t check_request(const xg_request& r) {
    const int dt = r.dtype;
    const bool f = dt == XG_FLOAT32 || dt == XG_FLOAT64;
    int nd = 1;
    switch (r.tpl) {
        case XG_TPL_AFFINE_CLIP:
            if (!f) raise(XG_TYPE_ERROR, "affine_clip requests are float32 or float64");
            nd = 4;
            break;
        case XG_TPL_ROW_SUM:
            if (!f && dt != XG_INT32) raise(XG_TYPE_ERROR, "row_sum requests are float32, float64 or int32");
            nd = 2;
            break;
        case XG_TPL_ROW_SOFTMAX:
            if (!f) raise(XG_TYPE_ERROR, "row_softmax requests are float32 or float64");
            nd = 2;
            break;
        case XG_TPL_BITMIX:
            if (dt != XG_INT32) raise(XG_TYPE_ERROR, "bitmix requests are int32");
            nd = 1;
            break;
        default: raise(XG_INVALID_ARGUMENT, "unknown request template");
    }
    for (int k = 0; k < nd; ++k)
        if (r.shape[k] < 0 || r.shape[k] >= (int64_t(1) << 31)) raise(XG_INVALID_ARGUMENT, "bad request shape");
    const int64_t n = request_elems(r);
    if (n >= (int64_t(1) << 31)) raise(XG_UNSUPPORTED, "request above 2^31 elements");
    const int64_t esz = dtype_size(dt);
    if (n > 0 && (!r.in || !r.out)) raise(XG_INVALID_ARGUMENT, "null operand");
    if (r.tpl == XG_TPL_AFFINE_CLIP) {
        for (int k = 0; k < 4; ++k)
            if (n > 0 && !r.params[k]) raise(XG_INVALID_ARGUMENT, "null affine parameter");
        return 2 * n * esz + 4 * r.shape[1] * esz;
    }
    if (r.tpl == XG_TPL_ROW_SUM) return n * esz + r.shape[0] * esz;
    return 2 * n * esz;
}

// The expression trees of the elementwise templates (the same trees a
// front end builds for the xtensor expressions).
void build_affine(Builder& b, const xg_request& r) {
    const int dt = r.dtype;
    const int64_t N = r.shape[0], C = r.shape[1], H = r.shape[2], W = r.shape[3];
    const int x = b.leaf(Builder::tensor(dt, {N, C, H, W}, {}, r.in));
    int p[4];
    for (int k = 0; k < 4; ++k) p[k] = b.leaf(Builder::tensor(dt, {1, C, 1, 1}, {}, r.params[k]));
    const int d = b.bin(XG_OP_MINUS, x, p[0]);
    const int s = b.un(XG_OP_SQRT, b.bin(XG_OP_PLUS, p[1], b.f64(1e-5)));
    const int y = b.bin(XG_OP_PLUS, b.bin(XG_OP_MULTIPLIES, b.bin(XG_OP_DIVIDES, d, s), p[2]), p[3]);
    b.node(XG_OP_CLIP, {y, b.i32(-3), b.i32(3)});
}

void build_bitmix(Builder& b, const xg_request& r) {
    const int x = b.leaf(Builder::tensor(XG_INT32, {r.shape[0]}, {}, r.in));
    const int m = b.bin(XG_OP_MULTIPLIES, x, b.i32(3));
    const int s = b.bin(XG_OP_RIGHT_SHIFT, x, b.i32(2));
    b.bin(XG_OP_BITWISE_AND, b.bin(XG_OP_PLUS, m, s), b.i32(0xFFFF));
}

xg_tensor out_tensor(const xg_request& r) {
    if (r.tpl == XG_TPL_AFFINE_CLIP)
        return Builder::tensor(r.dtype, {r.shape[0], r.shape[1], r.shape[2], r.shape[3]}, {}, r.out);
    if (r.tpl == XG_TPL_BITMIX) return Builder::tensor(XG_INT32, {r.shape[0]}, {}, r.out);
    if (r.tpl == XG_TPL_ROW_SUM) return Builder::tensor(r.dtype, {r.shape[0]}, {}, r.out);
    return Builder::tensor(r.dtype, {r.shape[0], r.shape[1]}, {}, r.out);
}

void status_check(xg_status st, const char* what) {
    if (st != XG_OK) raise(st, std::string(what) + ": " + xg_last_error_message());
}

}  // namespace
}  // namespace xg

using namespace xg;

struct xg_server_s {
    xg_server_config cfg{};
    int device = 0;
    int cus = 0;
    std::vector<Slot> slots;
    std::thread th;

    std::mutex mu;
    std::condition_variable cv_in;    // new submissions / stop
    std::condition_variable cv_done;  // completions
    std::deque<Job> inbox;
    std::vector<xg_completion> done;
    bool stop = false;
    std::atomic<bool> hold{false};
    int64_t submitted = 0, completed = 0, rejected = 0, launches = 0;
    int64_t inflight = 0, peak = 0, queued = 0, admitted = 0;

    // loop-thread state
    std::deque<Job> pending;
    std::map<ClassKey, std::deque<Job>> ready;
    uint64_t seq = 0;

    ~xg_server_s() {
        for (Slot& s : slots) {
            if (s.stream) (void)hipStreamSynchronize(s.stream);
            if (s.tmp) (void)hipFree(s.tmp);
            if (s.done) (void)hipEventDestroy(s.done);
            if (s.stream) (void)hipStreamDestroy(s.stream);
        }
    }

    ClassKey prepare(Job& j) {
        const xg_request& r = j.r;
        if (r.tpl == XG_TPL_EXPR) {
            const xg_expr e = j.expr();
            j.ew = plan_assign_best(&e, &j.out, &j.V);
            const int si = ew_spec_index(j.ew, j.V);
            const uint64_t solo = (si < 0 || j.ew.units > kSoloUnits) ? j.seq + 1 : 0;
            return ClassKey{r.tpl, j.out.dtype, si, j.ew.rank, j.V, solo};
        }
        if (r.tpl == XG_TPL_REDUCE) {  // axes and result type checked here; each runs alone
            const xg_expr e = j.expr();
            int dt = 0;
            Shape rs;
            reduce_result(&e, &j.rd, &dt, &rs);
            if (int(rs.size()) != j.out.ndim) raise(XG_BROADCAST_ERROR, "reduction output rank mismatch");
            for (size_t k = 0; k < rs.size(); ++k)
                if (rs[k] != j.out.shape[k]) raise(XG_BROADCAST_ERROR, "reduction output shape mismatch");
            return ClassKey{r.tpl, j.out.dtype, -1, 0, 0, j.seq + 1};
        }
        if (r.tpl == XG_TPL_AFFINE_CLIP || r.tpl == XG_TPL_BITMIX) {
            Builder b;
            if (r.tpl == XG_TPL_AFFINE_CLIP) build_affine(b, r);
            else build_bitmix(b, r);
            xg_expr e = b.expr();
            xg_tensor out = out_tensor(r);
            j.ew = plan_assign_best(&e, &out, &j.V);
            const int si = ew_spec_index(j.ew, j.V);
            const uint64_t solo = (si < 0 || j.ew.units > kSoloUnits) ? j.seq + 1 : 0;
            return ClassKey{r.tpl, r.dtype, si, j.ew.rank, j.V, solo};
        }
        const int op = r.tpl == XG_TPL_ROW_SUM ? BO_SUM : BO_SOFTMAX;
        j.item = batch_item(r.in, r.out, r.shape[0], r.shape[1], r.shape[1], op == BO_SUM ? 1 : r.shape[1], op, r.dtype);
        return ClassKey{r.tpl, r.dtype, -1, 0, 0, 0};
    }

    void complete(Slot& s, int status) {
        const int64_t t = steady_ns();
        std::lock_guard<std::mutex> lk(mu);
        for (const Job& j : s.jobs) {
            xg_completion c{};
            c.tag = j.r.tag;
            c.status = status;
            c.batch_size = int32_t(s.jobs.size());
            c.arrival_ns = j.r.arrival_ns;
            c.admit_ns = j.admit_ns;
            c.launch_ns = s.launch_ns;
            c.done_ns = t;
            done.push_back(c);
        }
        completed += int64_t(s.jobs.size());
        inflight -= s.bytes;
        admitted -= int64_t(s.jobs.size());
        s.jobs.clear();
        s.bytes = 0;
        s.busy = false;
        cv_done.notify_all();
    }

    // fused launch of one class group on slot s
    void launch_fused(Slot& s) {
        const xg_request& r0 = s.jobs[0].r;
        if (r0.tpl == XG_TPL_REDUCE) {
            for (const Job& j : s.jobs) {
                const xg_expr e = j.expr();
                reduce_call(&e, &j.rd, &j.out, s.stream, slot_workspace, nullptr, true);
            }
            std::lock_guard<std::mutex> lk(mu);
            ++launches;
            return;
        }
        if (r0.tpl == XG_TPL_AFFINE_CLIP || r0.tpl == XG_TPL_BITMIX || r0.tpl == XG_TPL_EXPR) {
            std::vector<const EwPlan*> plans;
            std::vector<int> V;
            for (const Job& j : s.jobs) {
                plans.push_back(&j.ew);
                V.push_back(j.V);
            }
            launch_ew_many(plans.data(), V.data(), plans.size(), s.stream);
            std::lock_guard<std::mutex> lk(mu);
            ++launches;
            return;
        }
        const int op = r0.tpl == XG_TPL_ROW_SUM ? BO_SUM : BO_SOFTMAX;
        // largest instances first, so the queue hands out the long work early;
        // tasks sized for a few per wave of a full grid
        std::vector<BItem> items;
        int64_t total = 0;
        for (const Job& j : s.jobs) {
            items.push_back(j.item);
            total += int64_t(j.item.rows) * j.item.cols;
        }
        std::stable_sort(items.begin(), items.end(), [](const BItem& a, const BItem& b) {
            return int64_t(a.rows) * a.cols > int64_t(b.rows) * b.cols;
        });
        const int64_t full_waves = int64_t(cus) * 16;
        const int task_elems = int(std::clamp<int64_t>(total / (full_waves * 4), 4096, 65536));
        std::vector<BTask> tasks;
        batch_split_tasks(items.data(), items.size(), op, r0.dtype, task_elems, tasks);
        const int blocks = int(std::max<size_t>(1, std::min<size_t>(size_t(cus) * 4, (tasks.size() + 3) / 4)));
        const size_t W = size_t(blocks) * 4;
        auto al = [](size_t x) { return (x + 255) / 256 * 256; };
        const size_t o_w = 256, o_items = al(o_w + (W + 1) * 4), o_tasks = al(o_items + items.size() * sizeof(BItem));
        const size_t bytes = o_tasks + tasks.size() * sizeof(BTask);
        Staging& st = staging_for(s.stream);
        char* h = st.host(bytes, s.stream);
        std::memset(h, 0, o_items);  // counter and wave lists (all work goes through the queue)
        std::memcpy(h + o_items, items.data(), items.size() * sizeof(BItem));
        std::memcpy(h + o_tasks, tasks.data(), tasks.size() * sizeof(BTask));
        char* d = st.upload(bytes, s.stream);
        BatchArgs a{reinterpret_cast<const BItem*>(d + o_items), reinterpret_cast<const BTask*>(d + o_tasks),
                    reinterpret_cast<const uint32_t*>(d + o_w), 0u, uint32_t(tasks.size()),
                    reinterpret_cast<uint32_t*>(d)};
        check_hip(launch_batch_rows(op, r0.dtype, a, blocks, 256, s.stream), "row kernel launch");
        std::lock_guard<std::mutex> lk(mu);
        ++launches;
    }

    // the per-request baseline: the plain per-expression entry points
    void launch_plain(Slot& s) {
        xg_stream_t st = reinterpret_cast<xg_stream_t>(s.stream);
        int n = 0;
        for (const Job& j : s.jobs) {
            const xg_request& r = j.r;
            if (expr_kind(r.tpl)) {
                const xg_expr e = j.expr();
                if (r.tpl == XG_TPL_EXPR) status_check(xg_assign(&e, &j.out, st), "assign");
                else status_check(xg_reduce(&e, &j.rd, &j.out, st), "reduce");
                ++n;
                continue;
            }
            if (request_elems(r) == 0 && r.tpl != XG_TPL_ROW_SUM) continue;
            if (r.tpl == XG_TPL_AFFINE_CLIP || r.tpl == XG_TPL_BITMIX) {
                Builder b;
                if (r.tpl == XG_TPL_AFFINE_CLIP) build_affine(b, r);
                else build_bitmix(b, r);
                xg_expr e = b.expr();
                xg_tensor out = out_tensor(r);
                status_check(xg_assign(&e, &out, st), "assign");
                ++n;
                continue;
            }
            const int64_t R = r.shape[0], Cc = r.shape[1];
            xg_reduce_desc rd{};
            rd.acc_dtype = -1;
            rd.n_axes = 1;
            rd.axes[0] = 1;
            if (r.tpl == XG_TPL_ROW_SUM) {
                Builder b;
                b.leaf(Builder::tensor(r.dtype, {R, Cc}, {}, r.in));
                xg_expr e = b.expr();
                xg_tensor out = out_tensor(r);
                rd.op = XG_RED_SUM;
                status_check(xg_reduce(&e, &rd, &out, st), "reduce");
                ++n;
                continue;
            }
            // softmax: m = amax(x, {1}, keep_dims); s = sum(exp(x - m), {1}, keep_dims); out = exp(x - m) / s
            const size_t esz = size_t(dtype_size(r.dtype));
            const size_t need = 2 * size_t(R) * esz;
            if (need > s.tmp_bytes) {
                if (s.tmp) check_hip(hipFree(s.tmp), "hipFree");  // the slot is idle: nothing reads it
                s.tmp = nullptr;
                s.tmp_bytes = 0;
                if (hipMalloc(&s.tmp, need) != hipSuccess) raise(XG_OUT_OF_MEMORY, "temporary allocation failed");
                s.tmp_bytes = need;
            }
            char* mp = static_cast<char*>(s.tmp);
            char* sp = mp + size_t(R) * esz;
            const xg_tensor mt = Builder::tensor(r.dtype, {R, 1}, {}, mp);
            const xg_tensor stt = Builder::tensor(r.dtype, {R, 1}, {}, sp);
            rd.keep_dims = 1;
            {
                Builder b;
                b.leaf(Builder::tensor(r.dtype, {R, Cc}, {}, r.in));
                xg_expr e = b.expr();
                rd.op = XG_RED_AMAX;
                status_check(xg_reduce(&e, &rd, &mt, st), "reduce");
            }
            {
                Builder b;
                const int x = b.leaf(Builder::tensor(r.dtype, {R, Cc}, {}, r.in));
                const int m = b.leaf(mt);
                b.un(XG_OP_EXP, b.bin(XG_OP_MINUS, x, m));
                xg_expr e = b.expr();
                rd.op = XG_RED_SUM;
                status_check(xg_reduce(&e, &rd, &stt, st), "reduce");
            }
            {
                Builder b;
                const int x = b.leaf(Builder::tensor(r.dtype, {R, Cc}, {}, r.in));
                const int m = b.leaf(mt);
                const int sv = b.leaf(stt);
                b.bin(XG_OP_DIVIDES, b.un(XG_OP_EXP, b.bin(XG_OP_MINUS, x, m)), sv);
                xg_expr e = b.expr();
                xg_tensor out = out_tensor(r);
                status_check(xg_assign(&e, &out, st), "assign");
            }
            n += 3;
        }
        std::lock_guard<std::mutex> lk(mu);
        launches += n;
    }

    void loop() {
        (void)hipSetDevice(device);
        for (;;) {
            bool progress = false;
            {
                std::lock_guard<std::mutex> lk(mu);
                while (!inbox.empty()) {
                    pending.push_back(std::move(inbox.front()));
                    inbox.pop_front();
                }
                bool idle = pending.empty() && ready.empty();
                for (const Slot& s : slots) idle = idle && !s.busy;
                if (stop && idle) return;
            }
            // 1. retire
            for (Slot& s : slots) {
                if (!s.busy) continue;
                const hipError_t q = hipEventQuery(s.done);
                if (q == hipErrorNotReady) continue;
                complete(s, q == hipSuccess ? XG_OK : XG_DEVICE_ERROR);
                progress = true;
            }
            // 2. admit in arrival order within the budget
            while (!pending.empty()) {
                Job& j = pending.front();
                {
                    std::lock_guard<std::mutex> lk(mu);
                    if (inflight + j.bytes > cfg.hbm_budget_bytes) break;
                    inflight += j.bytes;
                    peak = std::max(peak, inflight);
                    --queued;
                    ++admitted;
                }
                j.admit_ns = steady_ns();
                j.seq = seq++;
                ClassKey k{};
                bool ok = true;
                try {
                    k = cfg.plain ? ClassKey{0, 0, 0, 0, 0, 0} : prepare(j);
                } catch (...) {
                    ok = false;
                }
                if (!ok) {  // planning failed: complete it with an error right away
                    Slot tmp;
                    tmp.jobs.push_back(std::move(j));
                    tmp.bytes = tmp.jobs[0].bytes;
                    pending.pop_front();
                    complete(tmp, XG_INVALID_ARGUMENT);
                    progress = true;
                    continue;
                }
                ready[k].push_back(std::move(j));
                pending.pop_front();
                progress = true;
            }
            // 3. dispatch: idle streams take the class with the oldest request
            for (Slot& s : slots) {
                if (s.busy || ready.empty() || hold.load()) continue;
                auto best = ready.begin();
                for (auto it = ready.begin(); it != ready.end(); ++it)
                    if (it->second.front().seq < best->second.front().seq) best = it;
                auto& q = best->second;
                const size_t take = std::min<size_t>(q.size(), size_t(cfg.plain ? 1 : cfg.max_batch));
                for (size_t i = 0; i < take; ++i) {
                    s.bytes += q.front().bytes;
                    s.jobs.push_back(std::move(q.front()));
                    q.pop_front();
                }
                if (q.empty()) ready.erase(best);
                s.launch_ns = steady_ns();
                s.busy = true;
                int status = XG_OK;
                try {
                    if (cfg.plain) launch_plain(s);
                    else launch_fused(s);
                    check_hip(hipEventRecord(s.done, s.stream), "hipEventRecord");
                } catch (const Error& e) {
                    status = e.code;
                } catch (...) {
                    status = XG_DEVICE_ERROR;
                }
                if (status != XG_OK) {
                    (void)hipStreamSynchronize(s.stream);
                    complete(s, status);
                }
                progress = true;
            }
            if (!progress) {
                bool busy = false;
                for (const Slot& s : slots) busy = busy || s.busy;
                if (busy) {
                    std::this_thread::yield();
                } else {
                    std::unique_lock<std::mutex> lk(mu);
                    cv_in.wait_for(lk, std::chrono::microseconds(200), [&] { return !inbox.empty() || stop; });
                }
            }
        }
    }
};

extern "C" {

int64_t xg_server_now_ns(void) { return steady_ns(); }

xg_status xg_server_create(const xg_server_config* cfg, xg_server_t* out) {
    return guard([&] {
        if (!cfg || !out) raise(XG_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        if (cfg->hbm_budget_bytes <= 0) raise(XG_INVALID_ARGUMENT, "budget must be positive");
        if (cfg->streams < 1 || cfg->streams > 16) raise(XG_INVALID_ARGUMENT, "streams must be in [1, 16]");
        if (cfg->max_batch < 1) raise(XG_INVALID_ARGUMENT, "max_batch must be >= 1");
        if (cfg->queue_capacity < 1) raise(XG_INVALID_ARGUMENT, "queue_capacity must be >= 1");
        auto s = std::make_unique<xg_server_s>();
        s->cfg = *cfg;
        check_hip(hipGetDevice(&s->device), "hipGetDevice");
        s->cus = device_cus();
        s->slots.resize(size_t(cfg->streams));
        for (Slot& sl : s->slots) {
            // default priority: every stream of this library runs at the same priority
            check_hip(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreate");
            check_hip(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
        }
        xg_server_s* raw = s.get();
        s->th = std::thread([raw] { raw->loop(); });
        *out = s.release();
    });
}

xg_status xg_server_submit(xg_server_t s, const xg_request* req) {
    return guard([&] {
        if (!s || !req) raise(XG_INVALID_ARGUMENT, "null argument");
        Job j;
        j.r = *req;
        try {
            j.bytes = expr_kind(j.r.tpl) ? take_expr_request(j) : check_request(j.r);
        } catch (...) {
            std::lock_guard<std::mutex> lk(s->mu);
            ++s->rejected;
            throw;
        }
        if (j.r.arrival_ns == 0) j.r.arrival_ns = steady_ns();
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) raise(XG_INVALID_ARGUMENT, "server is stopping");
        if (j.bytes > s->cfg.hbm_budget_bytes) {
            ++s->rejected;
            raise(XG_OUT_OF_MEMORY, "request larger than the server's budget");
        }
        if (s->queued >= s->cfg.queue_capacity) {
            ++s->rejected;
            raise(XG_BUSY, "server queue full");
        }
        s->inbox.push_back(std::move(j));
        ++s->submitted;
        ++s->queued;
        s->cv_in.notify_one();
    });
}

xg_status xg_server_poll(xg_server_t s, xg_completion* out, int32_t cap, int32_t* n) {
    return guard([&] {
        if (!s || !n || (cap > 0 && !out)) raise(XG_INVALID_ARGUMENT, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        const size_t k = std::min<size_t>(s->done.size(), size_t(std::max(cap, 0)));
        std::copy(s->done.begin(), s->done.begin() + std::ptrdiff_t(k), out);
        s->done.erase(s->done.begin(), s->done.begin() + std::ptrdiff_t(k));
        *n = int32_t(k);
    });
}

xg_status xg_server_drain(xg_server_t s, int64_t timeout_ms) {
    return guard([&] {
        if (!s) raise(XG_INVALID_ARGUMENT, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto all_done = [&] { return s->completed >= s->submitted; };
        if (timeout_ms < 0) {
            s->cv_done.wait(lk, all_done);
        } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), all_done)) {
            raise(XG_BUSY, "drain timed out");
        }
    });
}

xg_status xg_server_hold(xg_server_t s, int32_t on) {
    return guard([&] {
        if (!s) raise(XG_INVALID_ARGUMENT, "null server");
        s->hold.store(on != 0);
        std::lock_guard<std::mutex> lk(s->mu);
        s->cv_in.notify_all();
    });
}

xg_status xg_server_stats_get(xg_server_t s, xg_server_stats* st) {
    return guard([&] {
        if (!s || !st) raise(XG_INVALID_ARGUMENT, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        st->submitted = s->submitted;
        st->completed = s->completed;
        st->rejected = s->rejected;
        st->launches = s->launches;
        st->inflight_bytes = s->inflight;
        st->peak_inflight_bytes = s->peak;
        st->queued = s->queued;
        st->admitted = s->admitted;
    });
}

void xg_server_destroy(xg_server_t s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_in.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    delete s;
}

}  // extern "C"
