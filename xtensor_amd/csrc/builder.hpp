// builder.hpp — host-side helper that assembles an xg_expr descriptor
// (nodes in post-order, leaves and constants numbered in order of creation),
// the same layout a front end hands to the C ABI.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/xg.h"

namespace xg {

struct Builder {
    std::vector<xg_node> nodes;
    std::vector<xg_tensor> leaves;
    std::vector<xg_scalar> consts;

    // row-major unless strides are given; `data` defaults to a placeholder
    // address that is only used to derive alignment
    static xg_tensor tensor(int dt, std::vector<int64_t> shape, std::vector<int64_t> strides = {},
                            const void* data = reinterpret_cast<const void*>(uintptr_t(1) << 20)) {
        xg_tensor t{};
        t.data = const_cast<void*>(data);
        t.dtype = dt;
        t.ndim = int(shape.size());
        if (strides.empty()) {
            strides.assign(shape.size(), 1);
            for (int k = int(shape.size()) - 2; k >= 0; --k) strides[size_t(k)] = strides[size_t(k + 1)] * shape[size_t(k + 1)];
        }
        for (size_t k = 0; k < shape.size(); ++k) {
            t.shape[k] = shape[k];
            t.strides[k] = strides[k];
        }
        return t;
    }
    int leaf(const xg_tensor& t) {
        leaves.push_back(t);
        return node(XG_OP_LEAF, {}, int(leaves.size()) - 1);
    }
    int node(int op, std::vector<int> args, int index = 0) {
        xg_node n{};
        n.op = op;
        n.nargs = int(args.size());
        for (size_t k = 0; k < args.size(); ++k) n.args[k] = args[k];
        n.index = index;
        nodes.push_back(n);
        return int(nodes.size()) - 1;
    }
    int f64(double v) {
        xg_scalar s{};
        s.dtype = XG_FLOAT64;
        s.v.f = v;
        consts.push_back(s);
        return node(XG_OP_CONST, {}, int(consts.size()) - 1);
    }
    int i32(int64_t v) {
        xg_scalar s{};
        s.dtype = XG_INT32;
        s.v.i = v;
        consts.push_back(s);
        return node(XG_OP_CONST, {}, int(consts.size()) - 1);
    }
    int un(int op, int a) { return node(op, {a}); }
    int bin(int op, int a, int b) { return node(op, {a, b}); }
    xg_expr expr() const {
        xg_expr e{};
        e.n_nodes = int(nodes.size());
        e.n_leaves = int(leaves.size());
        e.n_consts = int(consts.size());
        e.nodes = nodes.data();
        e.leaves = leaves.data();
        e.consts = consts.data();
        return e;
    }
};

}  // namespace xg
