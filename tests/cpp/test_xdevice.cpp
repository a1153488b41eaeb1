// test_xdevice.cpp — the C++ front end (xtensor_amd/cpp/xdevice.hpp) against
// the library.
//   test_xdevice cpu : descriptor flattening, shape / type deduction and
//                      error mapping (no device memory is touched)
//   test_xdevice gpu : evaluation on the device, checked against host loops
// Prints "OK" and exits 0 on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iterator>
#include <unistd.h>
#include <random>
#include <string>
#include <vector>

#include "xdevice.hpp"

namespace xt = xtensor_amd;

static int failures = 0;
#define EXPECT(c)                                                              \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);         \
            ++failures;                                                        \
        }                                                                      \
    } while (0)

static void cpu_tests() {
    // placeholder arrays: adapt() wraps an address that is never read here
    double* fake = reinterpret_cast<double*>(uintptr_t(1) << 20);
    auto a = xt::xdevice_array<double>::adapt(fake, {4, 3});
    auto b = xt::xdevice_array<double>::adapt(fake, {3});
    auto fa = xt::xdevice_array<float>::adapt(reinterpret_cast<float*>(fake), {2, 4, 3});
    auto ia = xt::xdevice_array<int32_t>::adapt(reinterpret_cast<int32_t*>(fake), {5});

    auto [dt, sh] = xt::expr_info(a * b + xt::exp(-a * a));  // test_xbroadcast.cpp: right-aligned
    EXPECT(dt == XG_FLOAT64 && sh == xt::shape_type({4, 3}));
    auto [dt2, sh2] = xt::expr_info(fa + 1);  // float + int -> float
    EXPECT(dt2 == XG_FLOAT32 && sh2 == xt::shape_type({2, 4, 3}));
    auto [dt3, sh3] = xt::expr_info(fa * 1e-5);  // float * double -> double
    EXPECT(dt3 == XG_FLOAT64 && sh3 == xt::shape_type({2, 4, 3}));
    auto [dt4, sh4] = xt::expr_info((ia * 3 + (ia >> 2)) & 0xFFFF);
    EXPECT(dt4 == XG_INT32 && sh4 == xt::shape_type({5}));
    auto [dt5, sh5] = xt::expr_info(xt::clip(fa, -3, 3));  // common type of float and int
    EXPECT(dt5 == XG_FLOAT32);
    auto [dt6, sh6] = xt::expr_info(a > b);
    EXPECT(dt6 == XG_BOOL && sh6 == xt::shape_type({4, 3}));

    // reducers: result types of core/xmath.hpp:367-423, 1825-1912
    auto [rt, rs] = xt::reduce_info(xt::mean(fa, {1}));
    EXPECT(rt == XG_FLOAT64 && rs == xt::shape_type({2, 3}));
    auto [rt2, rs2] = xt::reduce_info(xt::sum(ia, {0}));
    EXPECT(rt2 == XG_INT32 && rs2.empty());
    auto [rt3, rs3] = xt::reduce_info(xt::sum<int64_t>(ia, {0}));
    EXPECT(rt3 == XG_INT64);
    auto [rt4, rs4] = xt::reduce_info(xt::amax(fa, {0, 2}, xt::keep_dims));
    EXPECT(rt4 == XG_FLOAT32 && rs4 == xt::shape_type({1, 4, 1}));
    // variance / stddev / minmax (core/xmath.hpp:2031-2230)
    auto [rt5, rs5] = xt::reduce_info(xt::variance(fa, {0, 2}));
    EXPECT(rt5 == XG_FLOAT64 && rs5 == xt::shape_type({4}));
    auto [rt6, rs6] = xt::reduce_info(xt::stddev<float>(fa, {1}));
    EXPECT(rt6 == XG_FLOAT32 && rs6 == xt::shape_type({2, 3}));
    auto [rt7, rs7] = xt::reduce_info(xt::minmax(ia));
    EXPECT(rt7 == XG_INT32 && rs7 == xt::shape_type({2}));
    auto [rt8, rs8] = xt::reduce_info(xt::mean<int64_t>(ia, {0}));
    EXPECT(rt8 == XG_INT64 && rs8.empty());

    // views: strides and offsets of views/xview.hpp
    auto y = xt::xdevice_array<float>::adapt(reinterpret_cast<float*>(fake), {8, 6, 10});
    auto v = xt::view(y, xt::all(), xt::range(0, 6, 2), 3);
    EXPECT(v.shape() == xt::shape_type({8, 3}) && v.strides() == xt::shape_type({60, 20}) && v.offset() == 3);
    auto r = xt::view(y, xt::range(7, -9, -3));
    EXPECT(r.shape() == xt::shape_type({3, 6, 10}) && r.strides()[0] == -180 && r.offset() == 420);
    auto tr = xt::transpose(y);
    EXPECT(tr.shape() == xt::shape_type({10, 6, 8}) && tr.strides() == xt::shape_type({1, 10, 60}));

    // errors keep xtensor's exception types
    bool threw = false;
    try {
        xt::expr_info(a + fa);  // (4,3) vs (2,4,3) broadcast fine; (4,3) vs (5) not
        xt::expr_info(a + ia);
    } catch (const xt::broadcast_error&) {
        threw = true;
    }
    EXPECT(threw);
    threw = false;
    try {
        xt::reduce_info(xt::sum(fa, {2, 1}));  // unsorted axes (reducers/xreducer.hpp:1426-1440)
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    threw = false;
    try {
        xt::expr_info(fa & fa);  // bitwise and of floats: not a valid C++ expression
    } catch (const std::invalid_argument&) {
        threw = true;
    }
    EXPECT(threw);
}

static double ulp_rel(double got, double ref) {
    if (got == ref) return 0.0;
    return std::fabs(got - ref) / std::fabs(std::nextafter(ref, 2 * ref + 1) - ref);
}

// .npy files (io/xnpy.hpp) against the reference's fixture files in `dir`
static void npy_tests(const std::string& dir) {
    auto ia = xt::load_npy<int32_t>(dir + "/int.le.npy");  // test_xnpy.cpp:57, 80-81
    EXPECT(ia.to_host() == std::vector<int32_t>({3, 4, 5, 6, 7}));
    bool threw = false;
    try {
        (void)xt::load_npy<double>(dir + "/int.le.npy");
    } catch (const std::runtime_error&) {
        threw = true;
    }
    EXPECT(threw);
    auto u = xt::xdevice_array<uint64_t>::from_host({12, 14, 16, 18, 1234321}, {5});  // test_xnpy.cpp:144
    const std::string tmp = "/tmp/xg_test_dump_" + std::to_string(::getpid()) + ".npy";
    xt::dump_npy(tmp, u);
    auto slurp = [](const std::string& p) {
        std::ifstream f(p, std::ios::binary);
        return std::string((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    };
    EXPECT(!slurp(tmp).empty() && slurp(tmp) == slurp(dir + "/unsignedlong.le.npy"));
    std::remove(tmp.c_str());
}

static void gpu_tests() {
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    const int R = 64, Cc = 128;
    std::vector<double> ha(R * Cc), hb(Cc);
    for (auto& x : ha) x = u(rng);
    for (auto& x : hb) x = u(rng);
    auto a = xt::xdevice_array<double>::from_host(ha, {R, Cc});
    auto b = xt::xdevice_array<double>::from_host(hb, {Cc});

    // c = a * b + exp(-a * a)  (the broadcast expression of core/xassign.hpp's strided loop)
    xt::xdevice_array<double> c({R, Cc});
    xt::noalias(c) = a * b + xt::exp(-a * a);
    auto hc = c.to_host();
    double worst = 0;
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < Cc; ++j) {
            const double x = ha[size_t(i * Cc + j)];
            const double e = std::exp(-x * x);
            const double got_e = hc[size_t(i * Cc + j)] - x * hb[size_t(j)];
            (void)got_e;
            worst = std::max(worst, std::fabs(hc[size_t(i * Cc + j)] - (x * hb[size_t(j)] + e)) /
                                        std::fabs(x * hb[size_t(j)] + e));
        }
    EXPECT(worst < 1e-12);  // exp within 2 ulp, the rest exact given exp

    // s = sum(c, {1}), m = amax(c, {0})
    auto s = xt::eval<double>(xt::sum(c, {1}));
    auto m = xt::eval<double>(xt::amax(c, {0}));
    auto hs = s.to_host();
    auto hm = m.to_host();
    for (int i = 0; i < R; ++i) {
        long double acc = 0;
        for (int j = 0; j < Cc; ++j) acc += hc[size_t(i * Cc + j)];
        EXPECT(std::fabs(hs[size_t(i)] - double(acc)) <= 1e-14 * std::fabs(double(acc)) + 1e-300);
    }
    for (int j = 0; j < Cc; ++j) {
        double mx = -1e308;
        for (int i = 0; i < R; ++i) mx = hc[size_t(i * Cc + j)] > mx ? hc[size_t(i * Cc + j)] : mx;
        EXPECT(hm[size_t(j)] == mx);
    }

    // writing through a strided view: y[:, 0:128:2] = 2 * a[:, 0:128:2] + 1 (exact)
    xt::xdevice_array<double> y({R, Cc});
    xt::noalias(y) = a * 0.0;
    xt::noalias(xt::view(y, xt::all(), xt::range(0, Cc, 2))) = 2.0 * xt::view(a, xt::all(), xt::range(0, Cc, 2)) + 1;
    auto hy = y.to_host();
    for (int i = 0; i < R; ++i)
        for (int j = 0; j < Cc; ++j) {
            const double want = (j % 2 == 0) ? 2.0 * ha[size_t(i * Cc + j)] + 1 : 0.0;
            EXPECT(hy[size_t(i * Cc + j)] == want);
        }

    // integers: bit-exact, sum<int64_t>
    std::vector<int32_t> hi(1000);
    for (size_t k = 0; k < hi.size(); ++k) hi[k] = int32_t(k * 2654435761u);
    auto ia = xt::xdevice_array<int32_t>::from_host(hi, {1000});
    auto mix = xt::eval<int32_t>((ia * 3 + (ia >> 2)) & 0xFFFF);
    auto hmix = mix.to_host();
    int64_t tot = 0;
    for (size_t k = 0; k < hi.size(); ++k) {
        const int32_t want = int32_t((uint32_t(hi[k]) * 3u + uint32_t(hi[k] >> 2)) & 0xFFFFu);
        EXPECT(hmix[k] == want);
        tot += hi[k];
    }
    auto st = xt::eval<int64_t>(xt::sum<int64_t>(ia, {0}));
    EXPECT(st.to_host()[0] == tot);

    // variance with ddof and minmax on the device; small integers keep the
    // sums exact, so the values are exact too
    std::vector<double> hv = {1, 2, 3, 4, 6, 8};
    auto dv = xt::xdevice_array<double>::from_host(hv, {2, 3});
    auto var1 = xt::eval<double>(xt::variance(dv, {1}, 1)).to_host();  // rows {1,2,3}, {4,6,8}: 1, 4
    EXPECT(var1.size() == 2 && var1[0] == 1.0 && var1[1] == 4.0);
    auto sd = xt::eval<double>(xt::stddev(dv, {1})).to_host();
    EXPECT(std::fabs(sd[0] - std::sqrt(2.0 / 3.0)) < 1e-15 && std::fabs(sd[1] - std::sqrt(8.0 / 3.0)) < 1e-15);
    auto mm = xt::eval<int32_t>(xt::minmax(ia)).to_host();
    EXPECT(mm.size() == 2 && mm[0] == *std::min_element(hi.begin(), hi.end()) &&
           mm[1] == *std::max_element(hi.begin(), hi.end()));

    // index views, filter, filtration (test_xindex_view.cpp:24-57, 140-156)
    std::vector<double> he = {0.5, 1.5, 2.5, 3.5, 4.5, 5.5, 6.5, 7.5, 8.5};
    auto e = xt::xdevice_array<double>::from_host(he, {3, 3});
    auto iv = xt::index_view(e, {{1, 1}, {1, 2}, {2, 2}});
    auto ivh = xt::eval<double>(iv).to_host();
    EXPECT(ivh == std::vector<double>({4.5, 5.5, 8.5}));
    iv += 3;
    auto eh = e.to_host();
    EXPECT(eh[4] == 7.5 && eh[5] == 8.5 && eh[8] == 11.5 && eh[0] == 0.5);
    auto a5 = xt::xdevice_array<double>::from_host({1, 5, 3, 4, 5, 6}, {2, 3});
    xt::filter(a5, a5 >= 5) = 100;
    EXPECT(a5.to_host() == std::vector<double>({1, 100, 3, 4, 100, 100}));
    auto b5 = xt::xdevice_array<double>::from_host({1, 5, 3, 4, 5, 6}, {2, 3});
    xt::filtration(b5, b5 >= 5) += 2;
    EXPECT(b5.to_host() == std::vector<double>({1, 7, 3, 4, 7, 8}));

    // cumsum along an axis (reducers/xaccumulator.hpp:282-294): exact on small integers
    std::vector<double> hsmall(6 * 5);
    for (size_t k = 0; k < hsmall.size(); ++k) hsmall[k] = double(int(k % 7) - 3);
    auto sm = xt::xdevice_array<double>::from_host(hsmall, {6, 5});
    auto cs = xt::cumsum<double>(sm, 1);
    auto hcs = cs.to_host();
    for (int i = 0; i < 6; ++i) {
        double acc = 0;
        for (int j = 0; j < 5; ++j) {
            acc += hsmall[size_t(i * 5 + j)];
            EXPECT(hcs[size_t(i * 5 + j)] == acc);
        }
    }
    (void)ulp_rel;
}

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "cpu";
    try {
        cpu_tests();
        if (mode == "gpu") gpu_tests();
        if (mode == "gpu" && argc > 2) npy_tests(argv[2]);
    } catch (const std::exception& e) {
        std::printf("FAILED: exception %s\n", e.what());
        return 1;
    }
    if (failures) return 1;
    std::printf("OK\n");
    return 0;
}
