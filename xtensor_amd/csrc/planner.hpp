// planner.hpp — host side of the evaluator: validates an xg_expr, derives
// xtensor's value types and broadcast shape, lowers the tree to a typed
// micro-op program and normalises iteration spaces for the kernels.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "xg_internal.hpp"

namespace xg {

struct Error {
    int code;
    std::string msg;
};

[[noreturn]] void raise(int code, const std::string& msg);

using Shape = std::vector<int64_t>;

struct ExprInfo {
    int dtype = 0;            // value type of the root
    Shape shape;              // broadcast shape (xfunction::broadcast_shape)
    std::vector<int> node_dt; // value type of every node
};

// Validate + derive types and broadcast shape.  Throws Error.
ExprInfo analyse(const xg_expr* e);

// Result dtype of binary / unary / ternary xtensor functors (C++ rules).
int result_dtype(int op, int nargs, const int* arg_dt, int index);

enum ElemMap { MAP_NONE = 0, MAP_SQUARE = 1, MAP_ABS = 2, MAP_NONZERO = 3 };

// Lower the expression into a micro-op program whose result register holds
// map(root value) converted to `final_dt` (or left in its own type when -1).
Program lower(const xg_expr* e, const ExprInfo& info, int final_dt, int map = MAP_NONE);

// Value type of map(x) for an element of type dt.
int mapped_dtype(int map, int dt);

// Per-leaf strides against an iteration shape (right-aligned, size-1 and
// missing dims -> 0).
std::vector<std::vector<int64_t>> leaf_strides(const xg_expr* e, const Shape& iter);

bool broadcastable(const Shape& src, const Shape& dst);

// Elements per work unit of compiled kernels: 16 bytes of the widest operand
// (the generic interpreter always uses 4).
int choose_vector_width(const xg_expr* e, const xg_tensor* out);

// Elementwise plan for out = e, vector width V.
EwPlan plan_assign(const xg_expr* e, const xg_tensor* out, int V);

// Program signature used to find a compiled specialisation.
struct SpecKey {
    int nops;
    MicroOp ops[kMaxOps];
    int nleaves;
    uint8_t leaf_dtype[kMaxLeaves];
    uint8_t leaf_mode[kMaxLeaves];
    uint8_t out_dtype;
    uint8_t out_mode;
    int result_reg;
};

SpecKey spec_key(const EwPlan& p);

// True when a leaf with these element strides over an iteration space of
// this shape reads every element once (no stride-0 dim of extent > 1) and
// spans at least kNtMinBytes: its loads then stream non-temporally.
bool streams(const int64_t* stride, const int64_t* shape, int rank, int esz);
// LeafView of a leaf read by MO_GATHER: base pointer (offset applied), mode
// LM_GATHER, stride[0] / stride[1] = the lowest / highest valid element offset.
void set_gather_leaf(const xg_tensor& t, LeafView* lv);
// Switch non-temporal leaf loads / output stores on or off (measurement aid).
void set_nontemporal(int loads, int stores);
// Switch the granule-fill reads of partly written output granules on or off (measurement aid).
void set_fill_granules(int on);
bool spec_equal(const SpecKey& a, const SpecKey& b);

int64_t numel(const Shape& s);
std::string shape_str(const Shape& s);

}  // namespace xg
