// specgen.cpp — build-time generator of compiled program specialisations.
//
// Runs the same planner the library uses on a list of expression
// signatures and writes their micro-op programs as constexpr tables
// (specs.inc).  The kernels are then instantiated on those tables, so for a
// listed signature the compiler folds the interpreter into straight-line
// code.  At run time a plan whose program, leaf types and access modes equal
// a table entry launches that kernel; every other plan runs the dynamic
// interpreter.  The tensors below carry placeholder addresses that are only
// used to derive alignment and are never dereferenced.
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "builder.hpp"
#include "spec_emit.hpp"
#include "planner.hpp"
#include "reduce.hpp"

using namespace xg;

namespace {

struct SpecDef {
    const char* name;
    std::function<xg_tensor(Builder&)> build;  // builds the tree, returns the output
};

// clip((x - mean) / sqrt(var + 1e-5) * gamma + beta, -3, 3) on NCHW with
// per-channel [1,C,1,1] parameters, written through a row-strided view
xg_tensor affine_clip(Builder& b, int dt) {
    std::vector<int64_t> vs{64, 256, 28, 56}, vst{802816, 3136, 112, 1};
    int x = b.leaf(Builder::tensor(dt, vs, vst));
    int mean = b.leaf(Builder::tensor(dt, {1, 256, 1, 1}));
    int var = b.leaf(Builder::tensor(dt, {1, 256, 1, 1}));
    int gamma = b.leaf(Builder::tensor(dt, {1, 256, 1, 1}));
    int beta = b.leaf(Builder::tensor(dt, {1, 256, 1, 1}));
    int d = b.bin(XG_OP_MINUS, x, mean);
    int s = b.un(XG_OP_SQRT, b.bin(XG_OP_PLUS, var, b.f64(1e-5)));
    int y = b.bin(XG_OP_PLUS, b.bin(XG_OP_MULTIPLIES, b.bin(XG_OP_DIVIDES, d, s), gamma), beta);
    b.node(XG_OP_CLIP, {y, b.i32(-3), b.i32(3)});
    return Builder::tensor(dt, vs, vst);
}

const std::vector<SpecDef>& spec_defs() {
    static const std::vector<SpecDef> defs = {
        {"mul_add_exp_f64",  // a * b + exp(-a * a), b broadcast over rows
         [](Builder& b) {
             int a = b.leaf(Builder::tensor(XG_FLOAT64, {2048, 1024}));
             int bb = b.leaf(Builder::tensor(XG_FLOAT64, {1024}));
             int ab = b.bin(XG_OP_MULTIPLIES, a, bb);
             int na = b.un(XG_OP_NEGATE, a);
             int e = b.un(XG_OP_EXP, b.bin(XG_OP_MULTIPLIES, na, a));
             b.bin(XG_OP_PLUS, ab, e);
             return Builder::tensor(XG_FLOAT64, {2048, 1024});
         }},
        {"affine_clip_f32", [](Builder& b) { return affine_clip(b, XG_FLOAT32); }},
        {"affine_clip_f64", [](Builder& b) { return affine_clip(b, XG_FLOAT64); }},
        {"bitmix_i32",  // (x * 3 + (x >> 2)) & 0xFFFF
         [](Builder& b) {
             int x = b.leaf(Builder::tensor(XG_INT32, {64, 256, 56, 56}));
             int m = b.bin(XG_OP_MULTIPLIES, x, b.i32(3));
             int s = b.bin(XG_OP_RIGHT_SHIFT, x, b.i32(2));
             b.bin(XG_OP_BITWISE_AND, b.bin(XG_OP_PLUS, m, s), b.i32(0xFFFF));
             return Builder::tensor(XG_INT32, {64, 256, 56, 56});
         }},
        {"softmax_normalize_f32",  // exp(x - m) / s with row maxima m and row sums s of shape [R, 1]
         [](Builder& b) {
             int x = b.leaf(Builder::tensor(XG_FLOAT32, {512, 2048}));
             int m = b.leaf(Builder::tensor(XG_FLOAT32, {512, 1}));
             int s = b.leaf(Builder::tensor(XG_FLOAT32, {512, 1}));
             b.bin(XG_OP_DIVIDES, b.un(XG_OP_EXP, b.bin(XG_OP_MINUS, x, m)), s);
             return Builder::tensor(XG_FLOAT32, {512, 2048});
         }},
        {"softmax_normalize_f64",
         [](Builder& b) {
             int x = b.leaf(Builder::tensor(XG_FLOAT64, {512, 2048}));
             int m = b.leaf(Builder::tensor(XG_FLOAT64, {512, 1}));
             int s = b.leaf(Builder::tensor(XG_FLOAT64, {512, 1}));
             b.bin(XG_OP_DIVIDES, b.un(XG_OP_EXP, b.bin(XG_OP_MINUS, x, m)), s);
             return Builder::tensor(XG_FLOAT64, {512, 2048});
         }},
        {"plus_f32",
         [](Builder& b) {
             int x = b.leaf(Builder::tensor(XG_FLOAT32, {1 << 20}));
             int y = b.leaf(Builder::tensor(XG_FLOAT32, {1 << 20}));
             b.bin(XG_OP_PLUS, x, y);
             return Builder::tensor(XG_FLOAT32, {1 << 20});
         }},
        {"multiplies_f32",
         [](Builder& b) {
             int x = b.leaf(Builder::tensor(XG_FLOAT32, {1 << 20}));
             int y = b.leaf(Builder::tensor(XG_FLOAT32, {1 << 20}));
             b.bin(XG_OP_MULTIPLIES, x, y);
             return Builder::tensor(XG_FLOAT32, {1 << 20});
         }},
        {"plus_f64",
         [](Builder& b) {
             int x = b.leaf(Builder::tensor(XG_FLOAT64, {1 << 20}));
             int y = b.leaf(Builder::tensor(XG_FLOAT64, {1 << 20}));
             b.bin(XG_OP_PLUS, x, y);
             return Builder::tensor(XG_FLOAT64, {1 << 20});
         }},
        {"copy_f32",
         [](Builder& b) {
             b.leaf(Builder::tensor(XG_FLOAT32, {1 << 20}));
             return Builder::tensor(XG_FLOAT32, {1 << 20});
         }},
    };
    return defs;
}

struct RSpecDef {
    const char* name;
    int op;  // xg_reduce_op
    std::vector<int> axes;
    std::function<Shape(Builder&)> build;  // builds the tree, returns its shape
};

Shape leaf_only(Builder& b, int dt, Shape s) {
    b.leaf(Builder::tensor(dt, s));
    return s;
}

const std::vector<RSpecDef>& rspec_defs() {
    static const std::vector<RSpecDef> defs = {
        {"sum_f32", XG_RED_SUM, {1}, [](Builder& b) { return leaf_only(b, XG_FLOAT32, {8192, 4096}); }},
        {"sum_bf16", XG_RED_SUM, {1}, [](Builder& b) { return leaf_only(b, XG_BFLOAT16, {8192, 4096}); }},
        {"sum_f64", XG_RED_SUM, {1}, [](Builder& b) { return leaf_only(b, XG_FLOAT64, {2048, 1024}); }},
        {"amax_f32", XG_RED_AMAX, {1}, [](Builder& b) { return leaf_only(b, XG_FLOAT32, {8192, 4096}); }},
        {"amax_bf16", XG_RED_AMAX, {1}, [](Builder& b) { return leaf_only(b, XG_BFLOAT16, {8192, 4096}); }},
        {"amax_f64", XG_RED_AMAX, {1}, [](Builder& b) { return leaf_only(b, XG_FLOAT64, {2048, 1024}); }},
        {"norm_sq_f32", XG_RED_NORM_SQ, {1}, [](Builder& b) { return leaf_only(b, XG_FLOAT32, {8192, 4096}); }},
        {"norm_sq_bf16", XG_RED_NORM_SQ, {1}, [](Builder& b) { return leaf_only(b, XG_BFLOAT16, {8192, 4096}); }},
        {"sum_i32_to_i64", XG_RED_SUM, {1}, [](Builder& b) { return leaf_only(b, XG_INT32, {8192, 4096}); }},
        {"sum_exp_shifted_f32", XG_RED_SUM, {1},  // sum(exp(x - m), {1}) with m of shape [R, 1]
         [](Builder& b) {
             int x = b.leaf(Builder::tensor(XG_FLOAT32, {512, 2048}));
             int m = b.leaf(Builder::tensor(XG_FLOAT32, {512, 1}));
             b.un(XG_OP_EXP, b.bin(XG_OP_MINUS, x, m));
             return Shape{512, 2048};
         }},
        {"sum_exp_shifted_f64", XG_RED_SUM, {1},
         [](Builder& b) {
             int x = b.leaf(Builder::tensor(XG_FLOAT64, {512, 2048}));
             int m = b.leaf(Builder::tensor(XG_FLOAT64, {512, 1}));
             b.un(XG_OP_EXP, b.bin(XG_OP_MINUS, x, m));
             return Shape{512, 2048};
         }},
    };
    return defs;
}

}  // namespace

int main(int argc, char** argv) {
    FILE* f = argc > 1 ? std::fopen(argv[1], "w") : stdout;
    if (!f) return 1;
    std::fprintf(f, "// generated by specgen.cpp: compiled program specialisations\n#pragma once\n\n");
    std::string list = "#define XG_SPEC_LIST(X)";
    for (const auto& d : spec_defs()) {
        Builder b;
        xg_tensor out = d.build(b);
        xg_expr e = b.expr();
        int V = choose_vector_width(&e, &out);
        EwPlan p;
        try {
            p = plan_assign(&e, &out, V);
        } catch (const Error& err) {
            std::fprintf(stderr, "specgen: %s: %s\n", d.name, err.msg.c_str());
            return 2;
        }
        std::string sname = std::string("Spec_") + d.name;
        std::fprintf(f, "%s\n", emit_ew_spec(sname, p, V).c_str());
        list += " X(" + sname + ")";
    }
    std::fprintf(f, "%s\n\n", list.c_str());
    std::string rlist = "#define XG_RSPEC_LIST(X)";
    for (const auto& d : rspec_defs()) {
        Builder b;
        Shape shape = d.build(b);
        xg_expr e = b.expr();
        xg_reduce_desc r{};
        r.op = d.op;
        r.acc_dtype = d.op == XG_RED_SUM && std::string(d.name) == "sum_i32_to_i64" ? XG_INT64 : -1;
        r.n_axes = int(d.axes.size());
        for (size_t k = 0; k < d.axes.size(); ++k) r.axes[k] = d.axes[k];
        int rdt = 0;
        Shape rs;
        reduce_result(&e, &r, &rdt, &rs);
        xg_tensor out = Builder::tensor(rdt, rs);
        int V = choose_vector_width(&e, nullptr);
        RedPlan p;
        try {
            p = plan_reduce(&e, &r, &out, V, 256, nullptr, nullptr);
        } catch (const Error& err) {
            std::fprintf(stderr, "specgen: %s: %s\n", d.name, err.msg.c_str());
            return 2;
        }
        std::string sname = std::string("RSpec_") + d.name;
        std::fprintf(f, "%s\n", emit_red_spec(sname, p, V).c_str());
        rlist += " X(" + sname + ")";
    }
    std::fprintf(f, "%s\n", rlist.c_str());
    if (f != stdout) std::fclose(f);
    return 0;
}
