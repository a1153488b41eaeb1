// npy.cpp — .npy files straight to and from device memory (io/xnpy.hpp).
//
// Format rules follow xtensor's reader and writer: magic "\x93NUMPY", version
// 1.0 (2.0 when the header would not fit a 16-bit length), a Python-dict
// header with the keys 'descr', 'fortran_order', 'shape' in that order
// (io/xnpy.hpp:236-393), written as
// "{'descr': '<f8', 'fortran_order': False, 'shape': (3, 3), }" padded with
// spaces to a multiple of 64 bytes and ended by '\n' (io/xnpy.hpp:396-473).
// Loading requires the file's type string to be this library's type string
// for the dtype, as load_npy<T> does (io/xnpy.hpp:580-596); a Fortran-order
// file is read as column-major data.  The payload moves through two pinned
// staging buffers in chunks; a destination or source that is not contiguous
// in the file's order goes through a device temporary and the copy kernel.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "ew.hpp"
#include "planner.hpp"
#include "xg_api_internal.hpp"

namespace xg {

namespace {

const char kMagic[] = "\x93NUMPY";
constexpr size_t kMagicLen = 6;
constexpr size_t kChunk = size_t(32) << 20;

// type string of a dtype (build_typestring: '|' for one-byte types, '<' otherwise on this host)
std::string typestring(int dt) {
    switch (dt) {
    case XG_BOOL: return "|b1";
    case XG_INT8: return "|i1";
    case XG_UINT8: return "|u1";
    case XG_INT16: return "<i2";
    case XG_UINT16: return "<u2";
    case XG_INT32: return "<i4";
    case XG_UINT32: return "<u4";
    case XG_INT64: return "<i8";
    case XG_UINT64: return "<u8";
    case XG_FLOAT32: return "<f4";
    case XG_FLOAT64: return "<f8";
    default: raise(XG_UNSUPPORTED, "dtype has no .npy type string");
    }
}

int dtype_of_typestring(const std::string& t) {
    for (int dt = 0; dt < XG_NUM_DTYPES; ++dt) {
        if (dt == XG_BFLOAT16) continue;
        if (typestring(dt) == t) return dt;
    }
    raise(XG_INVALID_ARGUMENT, "Cast error: formats not matching " + t + " (no dtype of this host)");
}

struct Header {
    int dtype = -1;
    bool fortran = false;
    Shape shape;
    size_t data_offset = 0;
};

std::string unwrap(const std::string& s, char f, char b) {
    if (s.size() >= 2 && s.front() == f && s.back() == b) return s.substr(1, s.size() - 2);
    raise(XG_INVALID_ARGUMENT, "invalid .npy header");
}

std::string value_of(const std::string& kv) {
    size_t c = kv.find(':');
    if (c == std::string::npos) raise(XG_INVALID_ARGUMENT, "invalid .npy header");
    return kv.substr(c + 1);
}

// parse_header restated (io/xnpy.hpp:236-393)
void parse(std::string h, Header* out) {
    if (h.empty() || h.back() != '\n') raise(XG_INVALID_ARGUMENT, "invalid header");
    h.pop_back();
    h.erase(std::remove(h.begin(), h.end(), ' '), h.end());
    h = unwrap(h, '{', '}');
    const size_t kd = h.find("'descr'"), kf = h.find("'fortran_order'"), ks = h.find("'shape'");
    if (kd == std::string::npos) raise(XG_INVALID_ARGUMENT, "missing 'descr' key");
    if (kf == std::string::npos) raise(XG_INVALID_ARGUMENT, "missing 'fortran_order' key");
    if (ks == std::string::npos) raise(XG_INVALID_ARGUMENT, "missing 'shape' key");
    if (kd >= kf || kf >= ks) raise(XG_INVALID_ARGUMENT, "header keys in wrong order");
    auto strip = [](std::string s) {
        if (!s.empty() && s.back() == ',') s.pop_back();
        return s;
    };
    const std::string d = value_of(strip(h.substr(kd, kf - kd)));
    const std::string f = value_of(strip(h.substr(kf, ks - kf)));
    const std::string sh = value_of(strip(h.substr(ks)));
    const std::string t = unwrap(d, '\'', '\'');
    if (t.size() < 3 || std::string("<>|").find(t[0]) == std::string::npos ||
        std::string("ifucb").find(t[1]) == std::string::npos)
        raise(XG_INVALID_ARGUMENT, "invalid typestring");
    out->dtype = dtype_of_typestring(t);
    if (f == "True") out->fortran = true;
    else if (f == "False") out->fortran = false;
    else raise(XG_INVALID_ARGUMENT, "invalid fortran_order value");
    const std::string dims = unwrap(sh, '(', ')');
    out->shape.clear();
    size_t pos = 0;
    for (;;) {
        const size_t next = dims.find(',', pos);
        const std::string tok = dims.substr(pos, next == std::string::npos ? std::string::npos : next - pos);
        if (tok.empty()) {
            if (next != std::string::npos) raise(XG_INVALID_ARGUMENT, "invalid shape");
        } else {
            size_t used = 0;
            long long v = std::stoll(tok, &used);
            if (used != tok.size() || v < 0) raise(XG_INVALID_ARGUMENT, "invalid shape");
            out->shape.push_back(v);
        }
        if (next == std::string::npos) break;
        pos = next + 1;
    }
    if (out->shape.size() > size_t(XG_MAX_RANK)) raise(XG_UNSUPPORTED, "rank too large");
}

Header read_header(std::ifstream& f, const char* path) {
    if (!f) raise(XG_INVALID_ARGUMENT, std::string("cannot open ") + (path ? path : "(null)"));
    char m[kMagicLen + 2];
    f.read(m, sizeof m);
    if (!f || std::memcmp(m, kMagic, kMagicLen) != 0) raise(XG_INVALID_ARGUMENT, "the magic string is not valid");
    const unsigned char vmaj = static_cast<unsigned char>(m[kMagicLen]), vmin = static_cast<unsigned char>(m[kMagicLen + 1]);
    size_t hlen = 0, lenbytes = 0;
    if (vmaj == 1 && vmin == 0) {
        unsigned char b[2];
        f.read(reinterpret_cast<char*>(b), 2);
        hlen = size_t(b[0]) | (size_t(b[1]) << 8);
        lenbytes = 2;
    } else if (vmaj == 2 && vmin == 0) {
        unsigned char b[4];
        f.read(reinterpret_cast<char*>(b), 4);
        hlen = size_t(b[0]) | (size_t(b[1]) << 8) | (size_t(b[2]) << 16) | (size_t(b[3]) << 24);
        lenbytes = 4;
    } else {
        raise(XG_INVALID_ARGUMENT, "unsupported file format version");
    }
    if (!f || (kMagicLen + 2 + lenbytes + hlen) % 16 != 0) raise(XG_INVALID_ARGUMENT, "invalid header length");
    std::string h(hlen, '\0');
    f.read(h.data(), std::streamsize(hlen));
    if (!f) raise(XG_INVALID_ARGUMENT, "truncated header");
    Header out;
    parse(h, &out);
    out.data_offset = kMagicLen + 2 + lenbytes + hlen;
    return out;
}

// write_header restated (io/xnpy.hpp:396-473)
std::string header_bytes(int dt, bool fortran, const Shape& shape) {
    std::string sh = "(";
    for (size_t k = 0; k < shape.size(); ++k) sh += std::to_string(shape[k]) + ", ";
    if (shape.size() > 1) sh.erase(sh.size() - 2);
    else if (shape.size() == 1) sh.erase(sh.size() - 1);
    sh += ")";
    std::string dict = "{'descr': '" + typestring(dt) + "', 'fortran_order': " + (fortran ? "True" : "False") +
                       ", 'shape': " + sh + ", }";
    const size_t pre = dict.size() + 1;
    size_t meta = kMagicLen + 2 + 2 + pre;
    unsigned char ver[2] = {1, 0};
    if (meta >= 255 * 255) {
        meta = kMagicLen + 2 + 4 + pre;
        ver[0] = 2;
    }
    dict += std::string(64 - meta % 64, ' ');
    dict += '\n';
    std::string out(kMagic, kMagicLen);
    out += char(ver[0]);
    out += char(ver[1]);
    const uint32_t n = uint32_t(dict.size());
    out += char(n & 0xff);
    out += char((n >> 8) & 0xff);
    if (ver[0] == 2) {
        out += char((n >> 16) & 0xff);
        out += char((n >> 24) & 0xff);
    }
    return out + dict;
}

// two pinned bounce buffers shared by all calls (serialised)
std::mutex g_io_mu;
void* g_bounce[2] = {nullptr, nullptr};

void bounce_init() {
    for (auto& b : g_bounce)
        if (!b) check_hip(hipHostMalloc(&b, kChunk, hipHostMallocDefault), "hipHostMalloc");
}

// strides of a dense array of `shape` in C or Fortran order
void dense_strides(const Shape& shape, bool fortran, int64_t* st) {
    const int n = int(shape.size());
    int64_t s = 1;
    if (fortran) {
        for (int k = 0; k < n; ++k) { st[k] = s; s *= std::max<int64_t>(shape[size_t(k)], 1); }
    } else {
        for (int k = n - 1; k >= 0; --k) { st[k] = s; s *= std::max<int64_t>(shape[size_t(k)], 1); }
    }
}

bool dense_like(const xg_tensor& t, const int64_t* st) {
    for (int k = 0; k < t.ndim; ++k)
        if (t.shape[k] > 1 && t.strides[k] != st[k]) return false;
    return true;
}

// dst (device, any strides) = src (device, dense with strides st)
void copy_view(const xg_tensor& dst, const xg_tensor& src, hipStream_t st) {
    xg_node n{};
    n.op = XG_OP_LEAF;
    xg_expr e{};
    e.n_nodes = 1;
    e.n_leaves = 1;
    e.nodes = &n;
    e.leaves = &src;
    int V = 4;
    EwPlan p = plan_assign_best(&e, &dst, &V);
    if (p.units) check_hip(launch_ew(p, V, device_cus(), st, nullptr), "copy launch");
}

}  // namespace

}  // namespace xg

using namespace xg;

extern "C" {

xg_status xg_npy_info(const char* path, int32_t* dtype, int32_t* ndim, int64_t* shape, int32_t* fortran_order) {
    return guard([&] {
        if (!path || !dtype || !ndim || !shape || !fortran_order) raise(XG_INVALID_ARGUMENT, "null argument");
        std::ifstream f(path, std::ios::binary);
        Header h = read_header(f, path);
        *dtype = h.dtype;
        *ndim = int32_t(h.shape.size());
        for (size_t k = 0; k < h.shape.size(); ++k) shape[k] = h.shape[k];
        *fortran_order = h.fortran ? 1 : 0;
    });
}

xg_status xg_npy_load(const char* path, const xg_tensor* out, xg_stream_t stream) {
    return guard([&] {
        if (!path || !out) raise(XG_INVALID_ARGUMENT, "null argument");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        device_cus();  // the runtime's current device, before the pinned buffers
        std::ifstream f(path, std::ios::binary);
        Header h = read_header(f, path);
        if (out->dtype != h.dtype)
            raise(XG_INVALID_ARGUMENT, "Cast error: formats not matching " + typestring(h.dtype) + " vs " +
                                           (out->dtype == XG_BFLOAT16 ? std::string("bfloat16") : typestring(out->dtype)));
        Shape os(out->shape, out->shape + std::max(out->ndim, 0));
        if (os != h.shape) raise(XG_BROADCAST_ERROR, "output shape " + shape_str(os) + " differs from the file's " +
                                                     shape_str(h.shape));
        const int64_t n = numel(h.shape);
        const size_t bytes = size_t(n) * size_t(dtype_size(h.dtype));
        if (bytes == 0) return;
        int64_t dst[XG_MAX_RANK];
        dense_strides(h.shape, h.fortran, dst);
        const bool direct = dense_like(*out, dst);
        const int esz = dtype_size(h.dtype);
        char* dev = direct ? static_cast<char*>(out->data) + out->offset * esz
                           : static_cast<char*>(workspace(bytes, st));
        {
            std::lock_guard<std::mutex> lk(g_io_mu);
            bounce_init();
            hipEvent_t done[2] = {nullptr, nullptr};
            check_hip(hipEventCreateWithFlags(&done[0], hipEventDisableTiming), "hipEventCreate");
            check_hip(hipEventCreateWithFlags(&done[1], hipEventDisableTiming), "hipEventCreate");
            size_t off = 0;
            int b = 0;
            bool used[2] = {false, false};
            while (off < bytes) {
                const size_t c = std::min(kChunk, bytes - off);
                if (used[b]) check_hip(hipEventSynchronize(done[b]), "hipEventSynchronize");
                f.read(static_cast<char*>(g_bounce[b]), std::streamsize(c));
                if (!f) raise(XG_INVALID_ARGUMENT, "truncated .npy data");
                check_hip(hipMemcpyAsync(dev + off, g_bounce[b], c, hipMemcpyHostToDevice, st), "hipMemcpyAsync");
                check_hip(hipEventRecord(done[b], st), "hipEventRecord");
                used[b] = true;
                off += c;
                b ^= 1;
            }
            for (int k = 0; k < 2; ++k)
                if (used[k]) check_hip(hipEventSynchronize(done[k]), "hipEventSynchronize");
            hipEventDestroy(done[0]);
            hipEventDestroy(done[1]);
        }
        if (!direct) {
            xg_tensor src{};
            src.data = dev;
            src.dtype = h.dtype;
            src.ndim = int32_t(h.shape.size());
            for (int k = 0; k < src.ndim; ++k) {
                src.shape[k] = h.shape[size_t(k)];
                src.strides[k] = dst[k];
            }
            copy_view(*out, src, st);
        }
    });
}

xg_status xg_npy_dump(const char* path, const xg_tensor* in, xg_stream_t stream) {
    return guard([&] {
        if (!path || !in) raise(XG_INVALID_ARGUMENT, "null argument");
        if (in->ndim < 0 || in->ndim > XG_MAX_RANK) raise(XG_INVALID_ARGUMENT, "malformed tensor");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        device_cus();
        Shape shape(in->shape, in->shape + in->ndim);
        const std::string head = header_bytes(in->dtype, false, shape);
        const int esz = dtype_size(in->dtype);
        const size_t bytes = size_t(numel(shape)) * size_t(esz);
        int64_t dst[XG_MAX_RANK];
        dense_strides(shape, false, dst);
        const char* dev = static_cast<const char*>(in->data) + in->offset * esz;
        if (bytes && !dense_like(*in, dst)) {  // evaluate into a C-order temporary first
            xg_tensor tmp{};
            tmp.data = workspace(bytes, st);
            tmp.dtype = in->dtype;
            tmp.ndim = in->ndim;
            for (int k = 0; k < in->ndim; ++k) {
                tmp.shape[k] = in->shape[k];
                tmp.strides[k] = dst[k];
            }
            copy_view(tmp, *in, st);
            dev = static_cast<const char*>(tmp.data);
        }
        std::ofstream f(path, std::ios::binary | std::ios::trunc);
        if (!f) raise(XG_INVALID_ARGUMENT, std::string("cannot open ") + path + " for writing");
        f.write(head.data(), std::streamsize(head.size()));
        std::lock_guard<std::mutex> lk(g_io_mu);
        bounce_init();
        size_t off = 0;
        while (off < bytes) {
            const size_t c = std::min(kChunk, bytes - off);
            check_hip(hipMemcpyAsync(g_bounce[0], dev + off, c, hipMemcpyDeviceToHost, st), "hipMemcpyAsync");
            check_hip(hipStreamSynchronize(st), "hipStreamSynchronize");
            f.write(static_cast<const char*>(g_bounce[0]), std::streamsize(c));
            off += c;
        }
        if (!f) raise(XG_INVALID_ARGUMENT, std::string("write to ") + path + " failed");
    });
}

}  // extern "C"
