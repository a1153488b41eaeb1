/*
 * xg.h — C ABI of the MI355X (gfx950) evaluator for xtensor's lazy-expression
 * hot path: broadcast assignment, axis reductions and scans.
 *
 * Everything that crosses this boundary is a plain C type: pointers, sizes,
 * enums and fixed-size structs.  No C++ or torch types appear here.  A front
 * end (xtensor's expression-tag layer in C++, or the Python package shipped
 * beside this library) describes an expression tree with `xg_expr`, names the
 * output with `xg_tensor`, and calls one entry point.  All device buffers
 * belong to the caller; the library never allocates or frees them.  Calls are
 * stream-ordered: they enqueue work on `stream` and return without waiting.
 *
 * Reference interfaces that each entry point replaces (xtensor 0.27.1):
 *   xg_assign  <- xexpression_assigner<Tag>::assign_xexpression / assign_data
 *                 include/xtensor/core/xassign.hpp:208-226, 439-486
 *                 (reached from xt::noalias(out) = e, core/xnoalias.hpp:160-165,
 *                  and xview_semantic::assign_xexpression, core/xsemantic.hpp:724-731)
 *   xg_reduce  <- xt::reduce / xreducer + xreducer_stepper::aggregate
 *                 include/xtensor/reducers/xreducer.hpp:915-1079, 1404-1462, 1778-1828
 *                 and the sum/prod/amax/amin/mean entry points,
 *                 include/xtensor/core/xmath.hpp:367-423, 777-800, 1803-1912,
 *                 and norm_sq/norm_l2/norm_l1/norm_linf/norm_l0,
 *                 include/xtensor/reducers/xnorm.hpp:397-526
 *   xg_scan    <- xt::accumulate / accumulator_impl and cumsum/cumprod,
 *                 include/xtensor/reducers/xaccumulator.hpp:215-367,
 *                 include/xtensor/core/xmath.hpp:2247-2301
 *   xg_expr_info / xg_reduce_info / xg_scan_info
 *              <- xfunction::broadcast_shape + value_type deduction,
 *                 include/xtensor/core/xfunction.hpp:133-142, 537-544, 709-727,
 *                 include/xtensor/core/xstrides.hpp:737-794
 *   status codes <- broadcast_error (utils/xexception.hpp:63) and the
 *                 std::runtime_error throws of reducers/xreducer.hpp:1426-1440
 *                 and reducers/xaccumulator.hpp:231-234.
 */
#ifndef XG_H
#define XG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as HIP's hipStream_t (`struct ihipStream_t*`); NULL = null stream. */
typedef struct ihipStream_t* xg_stream_t;

#define XG_MAX_RANK 8
#define XG_MAX_NODES 96
#define XG_MAX_LEAVES 16
#define XG_MAX_CONSTS 32

/* ---- status codes ------------------------------------------------------ */
typedef enum xg_status {
    XG_OK = 0,
    XG_BROADCAST_ERROR = 1,  /* xt::broadcast_error: incompatible shapes             */
    XG_BAD_AXES = 2,         /* runtime_error: axes unsorted / duplicated / too big */
    XG_UNSUPPORTED = 3,      /* valid xtensor expression this library cannot run;
                                the C++ layer falls back to the host path         */
    XG_INVALID_ARGUMENT = 4, /* malformed descriptor                                */
    XG_TYPE_ERROR = 5,       /* operator not defined for the operand types
                                (the xtensor expression would not compile)          */
    XG_DEVICE_ERROR = 6,     /* a HIP call failed                                   */
    XG_OUT_OF_MEMORY = 7,    /* workspace or admission budget exhausted             */
    XG_BUSY = 8              /* server queue full                                   */
} xg_status;

/* ---- element types ------------------------------------------------------ */
/* C++ types used by xtensor containers; BFLOAT16 is a storage-only type that
 * loads as float (xtensor itself has no bf16). */
typedef enum xg_dtype {
    XG_BOOL = 0,
    XG_INT8 = 1,
    XG_UINT8 = 2,
    XG_INT16 = 3,
    XG_UINT16 = 4,
    XG_INT32 = 5,
    XG_UINT32 = 6,
    XG_INT64 = 7,
    XG_UINT64 = 8,
    XG_FLOAT32 = 9,
    XG_FLOAT64 = 10,
    XG_BFLOAT16 = 11,
    XG_NUM_DTYPES = 12
} xg_dtype;

/* ---- a strided array view ---------------------------------------------- */
/* Shape and strides are in elements, row-major or any other order.  As in
 * xtensor containers, the stride of a size-1 dimension is ignored (treated as
 * 0).  Element (i0..in-1) lives at data + (offset + sum_k i_k*strides[k]) *
 * sizeof(dtype).  Strides may be negative (reversed slices). */
typedef struct xg_tensor {
    void* data;
    int32_t dtype;
    int32_t ndim;
    int64_t offset;
    int64_t shape[XG_MAX_RANK];
    int64_t strides[XG_MAX_RANK];
} xg_tensor;

/* ---- a typed scalar (xt::xscalar leaf, or a reducer's `initial`) -------- */
typedef struct xg_scalar {
    int32_t dtype;
    int32_t pad_;
    union {
        int64_t i;   /* all signed integer types and bool */
        uint64_t u;  /* all unsigned integer types        */
        double f;    /* float32 (value exactly representable) and float64 */
    } v;
} xg_scalar;

/* ---- expression opcodes (xtensor functors) ------------------------------ */
typedef enum xg_op {
    XG_OP_LEAF = 0,   /* index = leaf number                                  */
    XG_OP_CONST = 1,  /* index = constant number                              */
    /* unary operators, core/xoperation.hpp:86-103 */
    XG_OP_NEGATE = 2,
    XG_OP_IDENTITY = 3,
    XG_OP_LOGICAL_NOT = 4,
    XG_OP_BITWISE_NOT = 5,
    XG_OP_CAST = 6,   /* xt::cast<T>, index = target dtype, core/xoperation.hpp:145-165 */
    /* binary operators, core/xoperation.hpp:86-109 */
    XG_OP_PLUS = 10,
    XG_OP_MINUS = 11,
    XG_OP_MULTIPLIES = 12,
    XG_OP_DIVIDES = 13,
    XG_OP_MODULUS = 14,
    XG_OP_LOGICAL_AND = 15,
    XG_OP_LOGICAL_OR = 16,
    XG_OP_BITWISE_AND = 17,
    XG_OP_BITWISE_OR = 18,
    XG_OP_BITWISE_XOR = 19,
    XG_OP_LEFT_SHIFT = 20,
    XG_OP_RIGHT_SHIFT = 21,
    XG_OP_LESS = 22,
    XG_OP_LESS_EQUAL = 23,
    XG_OP_GREATER = 24,
    XG_OP_GREATER_EQUAL = 25,
    XG_OP_EQUAL_TO = 26,
    XG_OP_NOT_EQUAL_TO = 27,
    XG_OP_MAXIMUM = 28, /* xt::maximum: select(a > b, a, b), core/xmath.hpp:589-602 */
    XG_OP_MINIMUM = 29, /* xt::minimum: select(a < b, a, b), core/xmath.hpp:573-587 */
    XG_OP_POW = 30,     /* std::pow, core/xmath.hpp:217 */
    XG_OP_FMOD = 31,    /* std::fmod */
    XG_OP_ATAN2 = 32,   /* std::atan2 */
    XG_OP_HYPOT = 33,   /* std::hypot */
    /* ternary */
    XG_OP_WHERE = 40,   /* xt::where(c, a, b) = select(c, a, b), core/xoperation.hpp:775-784 */
    XG_OP_CLIP = 41,    /* xt::clip(v, lo, hi), core/xmath.hpp:604-617, 814-823 */
    /* unary math functions (std:: overloads), core/xmath.hpp:152-217 */
    XG_OP_ABS = 50,
    XG_OP_FABS = 51,
    XG_OP_EXP = 52,
    XG_OP_EXP2 = 53,
    XG_OP_EXPM1 = 54,
    XG_OP_LOG = 55,
    XG_OP_LOG2 = 56,
    XG_OP_LOG10 = 57,
    XG_OP_LOG1P = 58,
    XG_OP_SQRT = 59,
    XG_OP_CBRT = 60,
    XG_OP_SIN = 61,
    XG_OP_COS = 62,
    XG_OP_TAN = 63,
    XG_OP_TANH = 64,
    XG_OP_SINH = 65,
    XG_OP_COSH = 66,
    XG_OP_ASIN = 67,
    XG_OP_ACOS = 68,
    XG_OP_ATAN = 69,
    XG_OP_ERF = 70,
    XG_OP_FLOOR = 71,
    XG_OP_CEIL = 72,
    XG_OP_TRUNC = 73,
    XG_OP_ROUND = 74,
    /* gather: the element of leaf `index` at element offset args[0] (an
     * integer expression) from the leaf's own offset, i.e. data[offset + v]:
     * xt::index_view / xt::filter read through a list of positions
     * (views/xindex_view.hpp:786-826).  The leaf does not take part in the
     * broadcast; offsets outside its extent read as 0. */
    XG_OP_GATHER = 80
} xg_op;

/* One node of an expression tree.  Nodes are listed so that every argument
 * precedes its user; the last node is the root.  A node may be used by
 * several parents (a DAG): that evaluates to the same values as repeating the
 * subtree, which is what xtensor would compute. */
typedef struct xg_node {
    int32_t op;       /* xg_op */
    int32_t nargs;    /* 0..3 */
    int32_t args[3];  /* node numbers of the arguments, in operator order */
    int32_t index;    /* leaf / constant number, or target dtype for CAST */
} xg_node;

typedef struct xg_expr {
    int32_t n_nodes;
    int32_t n_leaves;
    int32_t n_consts;
    int32_t flags;           /* reserved, 0 */
    const xg_node* nodes;
    const xg_tensor* leaves; /* device views, in their own shapes          */
    const xg_scalar* consts;
} xg_expr;

/* ---- reductions ---------------------------------------------------------- */
typedef enum xg_reduce_op {
    XG_RED_SUM = 0,       /* xt::sum<T>   (detail::plus, init 0)            */
    XG_RED_PROD = 1,      /* xt::prod<T>  (detail::multiplies, init 1)      */
    XG_RED_AMAX = 2,      /* xt::amax<T>  (maximum, init lowest())          */
    XG_RED_AMIN = 3,      /* xt::amin<T>  (minimum, init max())             */
    XG_RED_MEAN = 4,      /* xt::mean<T>: sum<T> / T(n), T = double if unset */
    XG_RED_NORM_SQ = 5,   /* xt::norm_sq: sum of v*v in big_promote type    */
    XG_RED_NORM_L2 = 6,   /* xt::norm_l2 = sqrt(norm_sq)                    */
    XG_RED_NORM_L1 = 7,   /* xt::norm_l1: sum of abs(v)                     */
    XG_RED_NORM_LINF = 8, /* xt::norm_linf: max of abs(v)                   */
    XG_RED_NORM_L0 = 9,   /* xt::norm_l0: count of v != 0                   */
    XG_RED_VARIANCE = 10, /* xt::variance<T>(e, axes, ddof): mean<T> of       */
                          /* square(e - mean<T>(e, axes)), core/xmath.hpp:2031-2113 */
    XG_RED_STDDEV = 11,   /* xt::stddev<T> = sqrt(variance<T>), core/xmath.hpp:2046, 2143 */
    XG_RED_MINMAX = 12    /* xt::minmax(e): {min, max} over all elements into */
                          /* a [2] output (std::min / std::max folds),        */
                          /* core/xmath.hpp:2195-2230                          */
} xg_reduce_op;

/* xt::evaluation_strategy::immediate (reducers/xreducer.hpp:35-161, 944-952):
 * the expression is evaluated into a temporary first (a lone container is
 * reduced as it is) and the temporary is reduced.  Values are those of the
 * default lazy form: integers exactly, floating-point sums within the same
 * error bound (the device's summation order is its own in both forms). */
#define XG_REDUCE_IMMEDIATE 1

typedef struct xg_reduce_desc {
    int32_t op;           /* xg_reduce_op                                    */
    int32_t acc_dtype;    /* the template argument T, or -1 for the default  */
    int32_t n_axes;       /* -1 = all axes (the no-axes overloads)           */
    int32_t axes[XG_MAX_RANK]; /* may be negative (normalised like xtensor) */
    int32_t keep_dims;    /* xt::keep_dims option                            */
    int32_t has_initial;  /* xt::initial(value) option                       */
    xg_scalar initial;
    int32_t flags;        /* 0, or XG_REDUCE_IMMEDIATE                        */
    int32_t ddof;         /* variance / stddev: delta degrees of freedom; the */
                          /* divisor is (size - ddof) / outputs in size_t     */
                          /* arithmetic, like detail::mean; 0 for other ops    */
} xg_reduce_desc;

/* ---- scans (accumulators) ------------------------------------------------ */
typedef enum xg_scan_op {
    XG_SCAN_CUMSUM = 0,
    XG_SCAN_CUMPROD = 1
} xg_scan_op;

#define XG_AXIS_NONE (-1000) /* flattened scan (cumsum(e) without axis)     */

typedef struct xg_scan_desc {
    int32_t op;        /* xg_scan_op                                         */
    int32_t acc_dtype; /* template argument T, or -1                          */
    int32_t axis;      /* may be negative; XG_AXIS_NONE = flattened           */
    int32_t flags;
} xg_scan_desc;

/* ---- library ------------------------------------------------------------- */

/* Version string of this library. */
const char* xg_version(void);

/* Message of the last error raised on the calling thread ("" if none). */
const char* xg_last_error_message(void);

/* Initialise the library on the current HIP device (queries the device,
 * resolves the device address of xg_reduce_noalloc's scratch array).
 * Optional: every entry point initialises lazily. */
xg_status xg_init(void);

/* Shape and value type of an expression: broadcast shape of all leaves
 * (xfunction::broadcast_shape), and the C++ value type of the root node
 * (decltype of the functor applied to the children's value types).  Pure
 * host computation, no device access. */
xg_status xg_expr_info(const xg_expr* e, int32_t* dtype, int32_t* ndim, int64_t* shape);

/* out = e   (xt::noalias(out) = e, or assignment into a strided view).
 * The broadcast shape of e must be broadcastable to out's shape
 * (xexpression_assigner::assert_compatible_shape).  Values are converted to
 * out's dtype with static_cast (conditional_cast).  Outputs may alias inputs
 * only element-for-element (out[i] read and written by the same element).
 * Arrays past 2^31 elements (or offsets) are accepted: the work is enqueued
 * as several launches along out's first dimension. */
xg_status xg_assign(const xg_expr* e, const xg_tensor* out, xg_stream_t stream);

/* Result dtype and shape of a reduction of e. */
xg_status xg_reduce_info(const xg_expr* e, const xg_reduce_desc* r, int32_t* dtype, int32_t* ndim,
                         int64_t* shape);

/* out = reduce(e, axes)   (xt::sum/prod/amax/amin/mean/norm_* assigned into out).
 * out must have exactly the reduction's shape; its dtype may differ from the
 * result dtype (then values are static_cast on store).  sum / prod / amax /
 * amin accept a reduced dimension past 2^31 elements: it is reduced in
 * pieces through a temporary of the result's size per piece. */
xg_status xg_reduce(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, xg_stream_t stream);

/* out = reduce(e, axes) exactly as xg_reduce, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in xtensor.  The call's
 * temporaries (split partials, a variance's kept-dims mean, an immediate
 * reduction's evaluated expression, the pieces of a reduction past 32-bit
 * indexing) live in one fixed array per device of
 * xg_reduce_noalloc_max_bytes() bytes (16 MiB), held until the process ends;
 * a reduction that needs more returns XG_OUT_OF_MEMORY before enqueueing
 * anything.  It calls no hipMalloc / hipFree and loads no module: a run-time
 * compiled kernel is used only if this process has already loaded it (else
 * the ahead-of-time or generic kernels run).  Every call on one device
 * shares that array: calls that overlap in time on one device (two streams,
 * or two threads) race and must be serialised by the caller. */
xg_status xg_reduce_noalloc(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, xg_stream_t stream);
/* Size in bytes of the per-device array xg_reduce_noalloc uses. */
size_t xg_reduce_noalloc_max_bytes(void);

/* Result dtype and shape of a scan of e. */
xg_status xg_scan_info(const xg_expr* e, const xg_scan_desc* s, int32_t* dtype, int32_t* ndim, int64_t* shape);

/* out = cumsum/cumprod(e, axis). */
xg_status xg_scan(const xg_expr* e, const xg_scan_desc* s, const xg_tensor* out, xg_stream_t stream);

/* out = e for a 2-D row-major output, and in the same pass over the stored
 * values the reduction `row_red` of every row (axes {1}) into `row_out` and
 * `col_red` of every column (axes {0}) into `col_out`; either pair may be NULL.
 * The results equal xg_assign(e, out) followed by xg_reduce of `out` with
 * each descriptor (within the reductions' stated error: the fold order
 * differs), and the call falls back to exactly those calls where the fused
 * kernel does not apply (see DESIGN).  No counterpart in xtensor, where
 * `c = e; s = sum(c, {1}); m = amax(c, {0})` (core/xassign.hpp:208-226,
 * reducers/xreducer.hpp:1778-1828) reads c back twice. */
xg_status xg_assign_reduce2d(const xg_expr* e, const xg_tensor* out, const xg_reduce_desc* row_red,
                             const xg_tensor* row_out, const xg_reduce_desc* col_red, const xg_tensor* col_out,
                             xg_stream_t stream);

/* Assignment into an index view: target's element at offset offsets[i] (from
 * the target's own offset) = values[i], for the 1-d integer expression
 * `offsets` and `values` broadcast to its length.  Replaces the assignment of
 * xt::index_view(e, indices) = values (views/xindex_view.hpp:786-797).  Both
 * are evaluated before any store; offsets outside the target are skipped;
 * repeated offsets leave one of their values. */
xg_status xg_assign_indexed(const xg_expr* values, const xg_expr* offsets, const xg_tensor* target,
                            xg_stream_t stream);

/* xt::argwhere<L>(cond) (core/xoperation.hpp:893-930; the indices of
 * xt::filter, views/xindex_view.hpp:820-826): out (int64 [count, ndim], any
 * strides) receives the multi-index of every element of cond that converts
 * to true, in row-major (column_major = 0) or column-major traversal order.
 * count, the number of such elements, is the caller's (for example from
 * xg_reduce of sum(cast<int64>(cast<bool>(cond)))); rows past out's first
 * dimension are not written.  Stream-ordered: a scan of the flags on the
 * device, no host copy of the condition. */
xg_status xg_argwhere(const xg_expr* cond, int32_t column_major, const xg_tensor* out, xg_stream_t stream);

/* ---- .npy files to and from device memory (io/xnpy.hpp) ------------------
 * Header of an .npy file: dtype, shape and order (xt::detail::load_npy_file's
 * header parse, io/xnpy.hpp:236-393, 646-673).  A type string without a
 * dtype of this host (big-endian, complex) is an error, as load_npy<T>'s
 * "Cast error" (io/xnpy.hpp:588-596). */
xg_status xg_npy_info(const char* path, int32_t* dtype, int32_t* ndim, int64_t* shape, int32_t* fortran_order);
/* Read the file's data into out: xt::load_npy<T>(path) (io/xnpy.hpp:775-800).
 * out must have the file's dtype and shape and may have any strides; a
 * Fortran-order file fills it in column-major element order.  Returns when
 * the data is on the device. */
xg_status xg_npy_load(const char* path, const xg_tensor* in_out, xg_stream_t stream);
/* Write a device tensor (any strides) as a C-order file with xtensor's
 * header layout: xt::dump_npy(path, e) (io/xnpy.hpp:396-473, 709-755).
 * Returns when the file is written. */
xg_status xg_npy_dump(const char* path, const xg_tensor* in, xg_stream_t stream);

/* ---- prepared plans --------------------------------------------------------
 * Plan once on the host, launch many times: a plan captures the normalised
 * iteration space, the lowered program and the kernel choice for the given
 * descriptors (including their device pointers).  Executing a plan only
 * enqueues its kernels, so it may be called inside HIP graph capture.  No
 * counterpart in xtensor (whose assignment is synchronous and re-derived on
 * every call); it serves the batching and serving layers. */
typedef struct xg_plan_s* xg_plan_t;

xg_status xg_plan_assign(const xg_expr* e, const xg_tensor* out, xg_plan_t* plan);
xg_status xg_plan_reduce(const xg_expr* e, const xg_reduce_desc* r, const xg_tensor* out, xg_plan_t* plan);
xg_status xg_plan_scan(const xg_expr* e, const xg_scan_desc* s, const xg_tensor* out, xg_plan_t* plan);
/* Reductions and scans the one-launch plans cannot hold (variance, stddev and
 * minmax, immediate evaluation, extents past 2^31 elements) are planned as
 * their descriptors and run as xg_reduce / xg_scan on every execution: the
 * host plans them each time, and a first execution may grow the stream's
 * workspace, so run such a plan once before capturing it in a graph. */
/* xg_assign_reduce2d prepared: XG_UNSUPPORTED when the single pass does not
 * apply to these descriptors (then plan the assignment and the reductions
 * apart). */
xg_status xg_plan_assign_reduce2d(const xg_expr* e, const xg_tensor* out, const xg_reduce_desc* row_red,
                                  const xg_tensor* row_out, const xg_reduce_desc* col_red, const xg_tensor* col_out,
                                  xg_plan_t* plan);
/* Bytes of device workspace the plan needs (0 for most elementwise plans). */
size_t xg_plan_workspace_size(xg_plan_t plan);
/* Enqueue the plan on `stream`; `workspace` must hold xg_plan_workspace_size
 * bytes (NULL: the library's per-stream workspace is used). */
xg_status xg_plan_execute(xg_plan_t plan, void* workspace, xg_stream_t stream);
/* Enqueue n plans that do not read each other's outputs.  Elementwise plans
 * that run the same compiled program at the same normalised rank are fused
 * into one launch (their descriptors go to the device in one asynchronous
 * upload through a pinned buffer of the stream); the rest run one launch
 * each, with the library's per-stream workspace.  No counterpart in xtensor:
 * this is the batching path for many small expressions. */
xg_status xg_plan_execute_many(const xg_plan_t* plans, int64_t n, xg_stream_t stream);
void xg_plan_destroy(xg_plan_t plan);

/* ---- batches of uneven instances -------------------------------------------
 * Many independent 2-D instances of one row-wise template, evaluated by ONE
 * persistent launch (a work queue over row blocks of all instances), which
 * can be replayed as a HIP graph.  Per instance this replaces the plain
 * expression sequence
 *   XG_BATCH_SOFTMAX: m = amax(x, {1}, keep_dims); e = exp(x - m);
 *                     out = e / sum(e, {1}, keep_dims)
 *                     (the row softmax of test/test_xmath.cpp:985-1027,
 *                      three assign/reduce calls of core/xmath.hpp:1803-1823)
 *   XG_BATCH_SUM:     out = sum(x, {1})        (core/xmath.hpp:1803)
 *   XG_BATCH_AMAX:    out = amax(x, {1})       (core/xmath.hpp:1823)
 * Floating-point sums accumulate with TwoSum compensation in the input's
 * precision; a row's result depends only on that row (its reduction tree is
 * a function of `cols` alone), so it is the same whatever batch it is in.
 * dtypes: float32 and float64 for all ops, int32 for SUM (wrapping, as
 * sum(int32) in xtensor) and AMAX. */
typedef enum xg_batch_op {
    XG_BATCH_SOFTMAX = 0,
    XG_BATCH_SUM = 1,
    XG_BATCH_AMAX = 2
} xg_batch_op;

typedef struct xg_batch_item {
    const void* in;          /* rows x cols, row stride in_row_stride          */
    void* out;               /* SOFTMAX: rows x cols; SUM/AMAX: rows values    */
    int64_t rows;
    int64_t cols;
    int64_t in_row_stride;   /* elements between consecutive input rows        */
    int64_t out_row_stride;  /* SOFTMAX: elements between output rows;
                                SUM/AMAX: elements between output values      */
} xg_batch_item;

/* Launch shape; chosen per shape class by xg_batch_autotune. */
typedef struct xg_batch_tuning {
    int32_t block;           /* threads per workgroup (64..256, multiple of 64) */
    int32_t blocks_per_cu;   /* persistent workgroups per compute unit          */
    int32_t task_elems;      /* target elements per task                        */
    int32_t dyn_percent;     /* share of the work (smallest tasks, percent)
                                handed out by the shared work queue; the rest
                                is assigned to waves on the host              */
} xg_batch_tuning;

typedef struct xg_batch_s* xg_batch_t;

/* Validate the items and upload the work tables (the library allocates
 * them; the caller keeps owning in/out).  Tuning starts from the cached
 * choice for the batch's shape class, or a default. */
xg_status xg_batch_create(int32_t op, int32_t dtype, const xg_batch_item* items, int64_t n_items,
                          xg_batch_t* batch);
/* Time the candidate launch shapes on `stream` (each trial writes the
 * outputs), keep the fastest and cache it for the shape class. */
xg_status xg_batch_autotune(xg_batch_t batch, xg_stream_t stream);
xg_status xg_batch_get_tuning(xg_batch_t batch, xg_batch_tuning* tuning);
xg_status xg_batch_set_tuning(xg_batch_t batch, const xg_batch_tuning* tuning);
/* Enqueue the batch as kernel launches (may be called under graph capture). */
xg_status xg_batch_launch(xg_batch_t batch, xg_stream_t stream);
/* Enqueue the batch by replaying its instantiated HIP graph.  One batch
 * must not run on two streams at once (its work queue is shared). */
xg_status xg_batch_execute(xg_batch_t batch, xg_stream_t stream);
/* Number of work-queue tasks and total elements of the batch. */
xg_status xg_batch_info(xg_batch_t batch, int64_t* n_tasks, int64_t* n_elements);
void xg_batch_destroy(xg_batch_t batch);

/* ---- request server ---------------------------------------------------------
 * A request loop for serving: callers submit requests whose operands already
 * live in device memory; a library thread admits them in arrival order while
 * the device bytes held by admitted, unfinished requests stay within a
 * budget, groups admitted requests by template, dtype and shape class, and
 * runs each group as one fused launch on one of a few streams, so groups
 * overlap on the device and new requests are admitted as earlier ones
 * finish (continuous batching).  No counterpart in xtensor, which evaluates
 * synchronously on the calling thread; each template is the expression the
 * comment names, with xtensor's semantics:
 *   XG_TPL_AFFINE_CLIP: out = clip((x - mean) / sqrt(var + 1e-5) * gamma + beta, -3, 3)
 *                       x, out [N, C, H, W]; mean/var/gamma/beta [C] broadcast
 *                       as [1, C, 1, 1]; float32 or float64
 *   XG_TPL_ROW_SUM:     out = sum(x, {1}), x [rows, cols]; float32/float64/int32
 *   XG_TPL_ROW_SOFTMAX: out = softmax of each row of x [rows, cols]; float32/float64
 *   XG_TPL_BITMIX:      out = (x * 3 + (x >> 2)) & 0xFFFF, x [n]; int32
 *   XG_TPL_EXPR:        *out_tensor = *expr, any expression xg_assign takes
 *                       (xt::noalias(out) = e); requests of one compiled
 *                       program and rank are fused like the templates above
 *   XG_TPL_REDUCE:      *out_tensor = reduce(*expr) with *reduce, any
 *                       reduction xg_reduce takes (any axes); each runs alone
 * The operands of the four fixed templates are contiguous row-major; EXPR and
 * REDUCE take strided views.  `expr`, `out_tensor` and `reduce` are copied at
 * submission (the device memory they point to must stay valid until the
 * request completes).  Results do not depend on what else is in flight: the
 * elementwise templates are exact functions of their element, and a
 * reduction's order depends only on its own shape. */
typedef enum xg_template {
    XG_TPL_AFFINE_CLIP = 0,
    XG_TPL_ROW_SUM = 1,
    XG_TPL_ROW_SOFTMAX = 2,
    XG_TPL_BITMIX = 3,
    XG_TPL_EXPR = 4,
    XG_TPL_REDUCE = 5
} xg_template;

typedef struct xg_request {
    int32_t tpl;             /* xg_template                                   */
    int32_t dtype;
    int64_t shape[4];        /* AFFINE_CLIP: N, C, H, W; ROW_*: rows, cols;
                                BITMIX: n                                      */
    const void* in;
    const void* params[4];   /* AFFINE_CLIP: mean, var, gamma, beta ([C] each) */
    void* out;
    uint64_t tag;            /* returned in the completion                    */
    int64_t arrival_ns;      /* xg_server_now_ns() time the request arrived;
                                0 = the time of submission                     */
    const xg_expr* expr;     /* EXPR / REDUCE: the expression                 */
    const xg_tensor* out_tensor; /* EXPR / REDUCE: the output view            */
    const xg_reduce_desc* reduce; /* REDUCE: the reduction                    */
} xg_request;

typedef struct xg_server_config {
    int64_t hbm_budget_bytes; /* operand + output bytes of admitted, unfinished
                                 requests (a request larger than this is
                                 rejected at submission)                       */
    int32_t streams;          /* groups in flight on the device (1..16)        */
    int32_t max_batch;        /* requests fused into one launch (>= 1)         */
    int32_t queue_capacity;   /* submitted, not yet admitted; beyond it
                                 submission returns XG_BUSY                    */
    int32_t plain;            /* 1: no fusion, each request runs the plain
                                 per-expression path (the reference sequence
                                 of assign / reduce launches) on its own      */
} xg_server_config;

typedef struct xg_completion {
    uint64_t tag;
    int32_t status;          /* xg_status of the request                      */
    int32_t batch_size;      /* requests in the launch that ran it            */
    int64_t arrival_ns;
    int64_t admit_ns;
    int64_t launch_ns;
    int64_t done_ns;         /* when the loop saw the launch complete         */
} xg_completion;

typedef struct xg_server_stats {
    int64_t submitted;
    int64_t completed;
    int64_t rejected;
    int64_t launches;        /* kernel-group launches (fused or plain)        */
    int64_t inflight_bytes;
    int64_t peak_inflight_bytes;
    int64_t queued;          /* submitted, not yet admitted                   */
    int64_t admitted;        /* admitted, not yet complete                    */
} xg_server_stats;

typedef struct xg_server_s* xg_server_t;

/* Monotonic clock of the server, nanoseconds. */
int64_t xg_server_now_ns(void);
/* Start a server (and its loop thread) on the current device. */
xg_status xg_server_create(const xg_server_config* cfg, xg_server_t* server);
/* Queue a request.  XG_BUSY: queue full (retry later); XG_OUT_OF_MEMORY:
 * the request alone exceeds the budget; XG_TYPE_ERROR / XG_INVALID_ARGUMENT:
 * malformed request.  Returns without waiting for the device. */
xg_status xg_server_submit(xg_server_t server, const xg_request* req);
/* Move up to `cap` completions into `out`; *n receives how many. */
xg_status xg_server_poll(xg_server_t server, xg_completion* out, int32_t cap, int32_t* n);
/* Wait until every submitted request has completed (timeout_ms < 0: no limit). */
xg_status xg_server_drain(xg_server_t server, int64_t timeout_ms);
xg_status xg_server_stats_get(xg_server_t server, xg_server_stats* stats);
/* hold = 1: keep admitting requests but start no launches (requests
 * accumulate into groups); hold = 0: resume. */
xg_status xg_server_hold(xg_server_t server, int32_t hold);
/* Drain, stop the loop thread and release the server's resources. */
void xg_server_destroy(xg_server_t server);

#ifdef __cplusplus
}
#endif

#endif /* XG_H */
