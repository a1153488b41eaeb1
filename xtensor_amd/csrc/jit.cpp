// jit.cpp — run-time compiled specialisations (see jit.hpp).
//
// A plan's program is written out as the same constexpr struct the
// ahead-of-time table holds (spec_emit.hpp) and the kernel template is
// instantiated on it through a hipRTC name expression.  The compiler runs in
// this process (no child process is started).  Include paths: the library's
// own headers next to the loaded libxg.so, ROCm's, and the host C++ headers
// that hipcc itself uses for device compilation.
#include "jit.hpp"

#include <hip/hiprtc.h>

#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

#include "spec_emit.hpp"

namespace fs = std::filesystem;

namespace xg {

namespace {

std::mutex g_mu;
int g_mode = -1;  // -1: not yet read from the environment
uint64_t g_min_elements = uint64_t(1) << 20;
JitStats g_stats{};
std::string g_last_error;

struct Entry {
    bool failed = false;
    hipModule_t mod = nullptr;
    JitKernel k;
};
std::map<std::string, Entry> g_entries;  // key: device | name expression | source

struct Env {
    bool ok = false;
    std::vector<std::string> opts;
    uint64_t salt = 0;  // hash of the headers and options: part of every cache key
    std::string cache_dir;
};

uint64_t fnv1a(const void* data, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char* p = static_cast<const unsigned char*>(data);
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 1099511628211ull;
    }
    return h;
}
uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) { return fnv1a(s.data(), s.size(), h); }

std::string hex(uint64_t v) {
    char b[17];
    std::snprintf(b, sizeof b, "%016llx", static_cast<unsigned long long>(v));
    return b;
}

// the subdirectory of `dir` with the highest leading version number that holds `file`
std::string newest_with(const fs::path& dir, const char* file) {
    std::string best;
    long bv = -1;
    std::error_code ec;
    for (const auto& d : fs::directory_iterator(dir, ec)) {
        const std::string n = d.path().filename().string();
        long v = std::strtol(n.c_str(), nullptr, 10);
        if (v > bv && fs::exists(d.path() / file, ec)) {
            bv = v;
            best = d.path().string();
        }
    }
    return best;
}

const Env& env() {
    static Env e;
    static std::once_flag once;
    std::call_once(once, [] {
        Dl_info info{};
        if (!dladdr(reinterpret_cast<void*>(&jit_stats), &info) || !info.dli_fname) {
            g_last_error = "cannot locate the library";
            return;
        }
        const fs::path pkg = fs::absolute(fs::path(info.dli_fname)).parent_path().parent_path();
        const fs::path csrc = pkg / "csrc", build = pkg / "_build", inc = pkg.parent_path() / "include";
        std::error_code ec;
        if (!fs::exists(csrc / "ew_kernels.hpp", ec) || !fs::exists(inc / "xg.h", ec)) {
            g_last_error = "kernel headers not found next to the library";
            return;
        }
        const char* rp = std::getenv("ROCM_PATH");
        const fs::path rocm = rp && *rp ? fs::path(rp) : fs::path("/opt/rocm");
        const std::string cxx = newest_with("/usr/include/c++", "type_traits");
        const std::string clang = newest_with(rocm / "lib" / "llvm" / "lib" / "clang", "include/stddef.h");
        if (cxx.empty() || clang.empty()) {
            g_last_error = "host C++ headers or the compiler's resource headers not found";
            return;
        }
        const std::string cxx_ver = fs::path(cxx).filename().string();
        e.opts = {"--offload-arch=gfx950", "-O3", "-std=c++20", "-ffp-contract=off",
                  "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result",
                  "-I" + csrc.string(), "-I" + build.string(), "-I" + inc.string(), "-I" + (rocm / "include").string(),
                  "-isystem" + cxx, "-isystem/usr/include/x86_64-linux-gnu/c++/" + cxx_ver,
                  "-isystem" + clang + "/include", "-isystem/usr/include/x86_64-linux-gnu", "-isystem/usr/include"};
        // headers and options salt every cache key, so a rebuilt library never
        // picks up kernels compiled from older headers
        uint64_t h = fnv1a(std::string("xg-jit-1"));
        std::vector<fs::path> hdrs;
        for (const auto& d : fs::directory_iterator(csrc, ec))
            if (d.path().extension() == ".hpp") hdrs.push_back(d.path());
        hdrs.push_back(inc / "xg.h");
        std::sort(hdrs.begin(), hdrs.end());
        for (const auto& p : hdrs) {
            std::ifstream f(p, std::ios::binary);
            std::stringstream ss;
            ss << f.rdbuf();
            h = fnv1a(p.filename().string(), h);
            h = fnv1a(ss.str(), h);
        }
        for (const auto& o : e.opts) h = fnv1a(o, h);
        e.salt = h;
        const char* cd = std::getenv("XG_JIT_CACHE");
        const char* home = std::getenv("HOME");
        if (cd && *cd) e.cache_dir = cd;
        else if (home && *home) e.cache_dir = (fs::path(home) / ".cache" / "xtensor_amd" / "jit").string();
        e.ok = true;
    });
    return e;
}

bool read_cache(const std::string& path, std::string* name, std::vector<char>* code) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::string magic, nm;
    if (!std::getline(f, magic) || magic != "XGJIT1" || !std::getline(f, nm) || nm.empty()) return false;
    std::vector<char> c((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    if (c.empty()) return false;
    *name = nm;
    *code = std::move(c);
    return true;
}

void write_cache(const std::string& dir, const std::string& path, const std::string& name, const std::vector<char>& code) {
    std::error_code ec;
    fs::create_directories(dir, ec);
    if (ec) return;
    const std::string tmp = path + ".tmp" + std::to_string(reinterpret_cast<uintptr_t>(&code));
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f) return;
        f << "XGJIT1\n" << name << "\n";
        f.write(code.data(), std::streamsize(code.size()));
        if (!f) return;
    }
    fs::rename(tmp, path, ec);
    if (ec) fs::remove(tmp, ec);
}

bool compile(const Env& e, const std::string& src, const std::string& expr, std::string* lowered,
             std::vector<char>* code, std::string* err) {
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "xg_jit.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS) {
        *err = "hiprtcCreateProgram failed";
        return false;
    }
    bool ok = false;
    if (hiprtcAddNameExpression(prog, expr.c_str()) == HIPRTC_SUCCESS) {
        std::vector<const char*> o;
        for (const auto& s : e.opts) o.push_back(s.c_str());
        const hiprtcResult r = hiprtcCompileProgram(prog, int(o.size()), o.data());
        if (r == HIPRTC_SUCCESS) {
            const char* nm = nullptr;
            size_t n = 0;
            if (hiprtcGetLoweredName(prog, expr.c_str(), &nm) == HIPRTC_SUCCESS && nm &&
                hiprtcGetCodeSize(prog, &n) == HIPRTC_SUCCESS && n > 0) {
                *lowered = nm;
                code->resize(n);
                ok = hiprtcGetCode(prog, code->data()) == HIPRTC_SUCCESS;
            }
            if (!ok) *err = "compiled code unavailable";
        } else {
            size_t ls = 0;
            hiprtcGetProgramLogSize(prog, &ls);
            std::string log(ls, '\0');
            if (ls) hiprtcGetProgramLog(prog, log.data());
            *err = std::string("hipRTC: ") + hiprtcGetErrorString(r) + ": " + log.substr(0, 2000);
        }
    } else {
        *err = "hiprtcAddNameExpression failed";
    }
    hiprtcDestroyProgram(&prog);
    return ok;
}

thread_local int t_loaded_only = 0;

// compiled (or cached) kernel `expr` of source `src` on the current device
const JitKernel* get(const std::string& src, const std::string& expr) {
    const Env& e = env();
    if (!e.ok) return nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    const std::string key = std::to_string(dev) + "|" + expr + "|" + src;
    Entry& en = g_entries[key];
    if (en.failed) return nullptr;
    if (en.k.fn) {
        ++g_stats.memory_hits;
        return &en.k;
    }
    if (t_loaded_only) {  // not loaded yet: leave the entry untouched for later callers
        if (!en.mod) g_entries.erase(key);
        return nullptr;
    }
    const uint64_t h = fnv1a(expr + "\n" + src, e.salt);
    const std::string path = e.cache_dir.empty() ? "" : (fs::path(e.cache_dir) / (hex(h) + ".co")).string();
    std::string lowered, err;
    std::vector<char> code;
    bool have = !path.empty() && read_cache(path, &lowered, &code);
    if (have) {
        ++g_stats.disk_hits;
    } else if (compile(e, src, expr, &lowered, &code, &err)) {
        ++g_stats.compiles;
        have = true;
        if (!path.empty()) write_cache(e.cache_dir, path, lowered, code);
    }
    if (have && hipModuleLoadData(&en.mod, code.data()) == hipSuccess) {
        if (hipModuleGetFunction(&en.k.fn, en.mod, lowered.c_str()) != hipSuccess) {
            en.k.fn = nullptr;
            err = "kernel not found in the compiled module";
        }
    } else if (have) {
        err = "hipModuleLoadData failed";
    }
    if (!en.k.fn) {
        en.failed = true;
        ++g_stats.failures;
        g_last_error = err;
        if (std::getenv("XG_JIT_VERBOSE")) std::fprintf(stderr, "xtensor_amd jit: %s\n", err.c_str());
        return nullptr;
    }
    en.k.name = "jit:" + hex(h).substr(0, 8);
    return &en.k;
}

}  // namespace

void jit_set_mode(int mode) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_mode = mode < 0 ? 0 : (mode > 2 ? 2 : mode);
}

int jit_get_mode() {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_mode < 0) {
        const char* m = std::getenv("XG_JIT");
        g_mode = m && *m ? std::max(0, std::min(2, std::atoi(m))) : 1;
        const char* n = std::getenv("XG_JIT_MIN_ELEMENTS");
        if (n && *n) g_min_elements = std::strtoull(n, nullptr, 10);
    }
    return g_mode;
}

int jit_set_thread_loaded_only(int on) {
    const int prev = t_loaded_only;
    t_loaded_only = on ? 1 : 0;
    return prev;
}

bool jit_wanted(uint64_t elements) {
    const int m = jit_get_mode();
    return m == 2 || (m == 1 && elements >= g_min_elements);
}

const JitKernel* jit_ew(const EwPlan& p, int V) {
    if (!p.compact || p.rank < 1 || p.rank > 4) return nullptr;
    const std::string src = "#include \"ew_kernels.hpp\"\n" + emit_ew_spec("JitSpec", p, V);
    const std::string expr = "&xg::ew_spec_kernel<" + std::to_string(V) + ", " + std::to_string(p.prog.nregs) +
                             ", JitSpec, " + std::to_string(p.rank) + ">";
    return get(src, expr);
}

const JitKernel* jit_reduce(const RedPlan& p, int V, int kind) {
    if (p.index64) return nullptr;
    const std::string src = "#include \"reduce_kernels.hpp\"\n" + emit_red_spec("JitSpec", p, V);
    const std::string head = std::to_string(V) + ", " + std::to_string(p.prog.nregs) + ", JitSpec, ";
    const std::string ao = std::to_string(p.acc) + ", " + std::to_string(p.op);
    std::string expr;
    switch (kind) {
    case JR_ROWS_FAST1: expr = "&xg::reduce_rows_fast_kernel<" + head + ao + ", 1>"; break;
    case JR_ROWS_FAST2: expr = "&xg::reduce_rows_fast_kernel<" + head + ao + ", 2>"; break;
    case JR_COLS_FAST: expr = "&xg::reduce_cols_fast_kernel<" + head + ao + ">"; break;
    case JR_ROWS: expr = "&xg::reduce_rows_kernel<" + head + "false, " + ao + ", 4>"; break;
    case JR_COLS: expr = "&xg::reduce_cols_kernel<" + head + "false, " + ao + ", 4>"; break;
    default: return nullptr;
    }
    return get(src, expr);
}

const JitKernel* jit_assign_rc(const EwPlan& p, int V, int rak, int rop, int cak, int cop, int ul) {
    if (!p.compact || p.rank < 1 || p.rank > 4) return nullptr;
    const std::string src = "#include \"assign_reduce.hpp\"\n" + emit_ew_spec("JitSpec", p, V);
    const std::string expr = "&xg::assign_rc_kernel<" + std::to_string(V) + ", " + std::to_string(p.prog.nregs) +
                             ", JitSpec, " + std::to_string(p.rank) + ", " + std::to_string(rak) + ", " +
                             std::to_string(rop) + ", " + std::to_string(cak) + ", " + std::to_string(cop) + ", " +
                             std::to_string(ul) + ">";
    return get(src, expr);
}

hipError_t jit_launch(const JitKernel* k, uint32_t grid, uint32_t block, const void* arg, size_t arg_bytes,
                      hipStream_t stream) {
    size_t bytes = arg_bytes;
    void* config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void*>(arg), HIP_LAUNCH_PARAM_BUFFER_SIZE, &bytes,
                      HIP_LAUNCH_PARAM_END};
    return hipModuleLaunchKernel(k->fn, grid, 1, 1, block, 1, 1, 0, stream, nullptr, config);
}

JitStats jit_stats() {
    std::lock_guard<std::mutex> lk(g_mu);
    return g_stats;
}

const char* jit_last_error() { return g_last_error.c_str(); }

}  // namespace xg
