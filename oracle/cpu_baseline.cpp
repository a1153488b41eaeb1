// cpu_baseline.cpp — typed CPU restatements of the benchmark workloads, timed
// as the CPU baseline (test / benchmark infrastructure only; never part of
// the product).
//
// The reference cannot be built here (xtl / xsimd / TBB headers absent), so
// these restate what xtensor's CPU path executes for each workload, with the
// same C++ types and loop order:
//  * affine + clip into a strided view: xtensor's assigner falls back to the
//    scalar single-threaded stepper_assigner for this expression
//    (core/xassign.hpp:1051-1087, 1339-1343; core/xassign.hpp:644-695):
//    one element at a time, row-major over the view, float - float, then the
//    double divide / multiply / add, clip via select, narrowing store.
//  * mul_add_exp: the strided-loop assigner (core/xassign.hpp:1100-1328),
//    outer rows x inner contiguous run; threads over rows (OpenMP here in
//    place of TBB).
//  * row / column reductions: xreducer_stepper::aggregate_impl, a sequential
//    fold per output (reducers/xreducer.hpp:1801-1828), single-threaded; and
//    evaluation_strategy::immediate (reducers/xreducer.hpp:289-565): a
//    std::accumulate per contiguous run, or a std::transform that adds whole
//    rows into the output when the reduced axis is strided, single-threaded.
//  * accumulators: a copy, then the linear sweep res[k + s] = res[k] + res[k + s]
//    with s the axis stride (reducers/xaccumulator.hpp:215-297).
// Built with -O3 -march=native -ffp-contract=off.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>

#include <omp.h>

namespace {
inline double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
template <class T, class U> inline auto xsel(bool c, T a, U b) -> std::common_type_t<T, U> { return c ? a : b; }
}  // namespace

extern "C" {

// y[:, :, 0:h:2, :] = clip((x - mean) / sqrt(var + 1e-5) * gamma + beta, -3, 3)
// on the same view of x; returns seconds for `reps` passes.
double xb_affine_clip_f32(const float* x, const float* mean, const float* var, const float* gamma,
                          const float* beta, float* y, int n, int c, int h, int w, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r) {
        for (int in = 0; in < n; ++in)
            for (int ic = 0; ic < c; ++ic)
                for (int ih = 0; ih < h; ih += 2)
                    for (int iw = 0; iw < w; ++iw) {
                        size_t o = ((size_t(in) * c + ic) * h + ih) * w + iw;
                        float d = x[o] - mean[ic];
                        double v = double(d) / std::sqrt(double(var[ic]) + 1e-5) * double(gamma[ic]) + double(beta[ic]);
                        double lo = -3, hi = 3;
                        double res = xsel(lo < hi, xsel(v < lo, lo, xsel(hi < v, hi, v)), hi);
                        y[o] = static_cast<float>(res);
                    }
    }
    return now() - t0;
}

// int32 variant: y = (x * 3 + (x >> 2)) & 0xFFFF (linear assigner)
double xb_bitmix_i32(const int32_t* x, int32_t* y, size_t n, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r)
#pragma omp parallel for schedule(static)
        for (long i = 0; i < long(n); ++i) y[i] = (x[i] * 3 + (x[i] >> 2)) & 0xFFFF;
    return now() - t0;
}

// c = a * b + exp(-a * a), b broadcast over rows
double xb_mul_add_exp_f64(const double* a, const double* b, double* c, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r)
#pragma omp parallel for schedule(static)
        for (int i = 0; i < rows; ++i)
            for (int j = 0; j < cols; ++j) {
                double x = a[size_t(i) * cols + j];
                c[size_t(i) * cols + j] = x * b[j] + std::exp(-x * x);
            }
    return now() - t0;
}

// sum over the contiguous axis of a [rows, cols] float array, sequential float fold
double xb_row_sum_f32(const float* x, float* out, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r)
        for (int i = 0; i < rows; ++i) {
            float acc = 0.f;
            const float* p = x + size_t(i) * cols;
            for (int j = 0; j < cols; ++j) acc = acc + p[j];
            out[i] = acc;
        }
    return now() - t0;
}

// sum over the strided axis (columns of a row-major array), per output a fold down the rows
double xb_col_sum_f32(const float* x, float* out, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r)
        for (int j = 0; j < cols; ++j) {
            float acc = 0.f;
            for (int i = 0; i < rows; ++i) acc = acc + x[size_t(i) * cols + j];
            out[j] = acc;
        }
    return now() - t0;
}

// cumsum along the contiguous axis: copy then in-place sequential recurrence
double xb_cumsum_rows_f32(const float* x, float* out, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r) {
        std::memcpy(out, x, sizeof(float) * size_t(rows) * cols);
        for (int i = 0; i < rows; ++i) {
            float* p = out + size_t(i) * cols;
            for (int j = 1; j < cols; ++j) p[j] = p[j - 1] + p[j];
        }
    }
    return now() - t0;
}

// row softmax of one [rows, cols] instance: amax(keep_dims), exp(x - m), sum(keep_dims), divide
double xb_softmax_f32(const float* x, float* out, float* tmp, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r)
        for (int i = 0; i < rows; ++i) {
            const float* p = x + size_t(i) * cols;
            float* q = out + size_t(i) * cols;
            float m = -3.402823466e38f;
            for (int j = 0; j < cols; ++j) m = xsel(m > p[j], m, p[j]);
            float s = 0.f;
            for (int j = 0; j < cols; ++j) s = s + std::exp(p[j] - m);
            for (int j = 0; j < cols; ++j) q[j] = std::exp(p[j] - m) / s;
            tmp[i] = s;
        }
    return now() - t0;
}

// ---- reductions and scans of [rows, cols] arrays, single-threaded --------

// row sums of a float64 array (config 1's s = sum(c, {1}), lazy fold)
double xb_row_sum_f64(const double* x, double* out, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r)
        for (int i = 0; i < rows; ++i) {
            double acc = 0.0;
            const double* p = x + size_t(i) * cols;
            for (int j = 0; j < cols; ++j) acc = acc + p[j];
            out[i] = acc;
        }
    return now() - t0;
}

// column maxima of a float64 array (config 1's m = amax(c, {0})), immediate
// form: the first row copied, then every row folded in with maximum's select
double xb_col_amax_f64(const double* x, double* out, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r) {
        for (int j = 0; j < cols; ++j) out[j] = xsel(-1.7976931348623157e308 > x[j], -1.7976931348623157e308, x[j]);
        for (int i = 1; i < rows; ++i) {
            const double* p = x + size_t(i) * cols;
            for (int j = 0; j < cols; ++j) out[j] = xsel(out[j] > p[j], out[j], p[j]);
        }
    }
    return now() - t0;
}

// column sums of a float32 array, immediate form (rows added into the output)
double xb_col_sum_f32_immediate(const float* x, float* out, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r) {
        for (int j = 0; j < cols; ++j) out[j] = 0.f + x[j];
        for (int i = 1; i < rows; ++i) {
            const float* p = x + size_t(i) * cols;
            for (int j = 0; j < cols; ++j) out[j] = out[j] + p[j];
        }
    }
    return now() - t0;
}

// amax over every element (std::accumulate with maximum)
double xb_amax_all_f32(const float* x, float* out, size_t n, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r) {
        float acc = -3.402823466e38f;
        for (size_t k = 0; k < n; ++k) acc = xsel(acc > x[k], acc, x[k]);
        out[0] = acc;
    }
    return now() - t0;
}

// norm_l2 over every element: sqrt of the sum of squares in double
double xb_norm_l2_all_f32(const float* x, double* out, size_t n, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r) {
        double acc = 0.0;
        for (size_t k = 0; k < n; ++k) acc = acc + double(x[k]) * double(x[k]);
        out[0] = std::sqrt(acc);
    }
    return now() - t0;
}

// cumsum along the strided axis 0: copy, then the linear sweep
// out[k + cols] = out[k] + out[k + cols]
double xb_cumsum_cols_f32(const float* x, float* out, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r) {
        std::memcpy(out, x, sizeof(float) * size_t(rows) * cols);
        const size_t n = size_t(rows - 1) * cols;
        for (size_t k = 0; k < n; ++k) out[k + cols] = out[k] + out[k + cols];
    }
    return now() - t0;
}

// sum<int64_t> of an int32 array over axis 1 (lazy fold) and axis 0 (immediate rows)
double xb_row_sum_i32_i64(const int32_t* x, int64_t* out, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r)
        for (int i = 0; i < rows; ++i) {
            int64_t acc = 0;
            const int32_t* p = x + size_t(i) * cols;
            for (int j = 0; j < cols; ++j) acc = acc + p[j];
            out[i] = acc;
        }
    return now() - t0;
}

double xb_col_sum_i32_i64(const int32_t* x, int64_t* out, int rows, int cols, int reps) {
    double t0 = now();
    for (int r = 0; r < reps; ++r) {
        for (int j = 0; j < cols; ++j) out[j] = int64_t(0) + x[j];
        for (int i = 1; i < rows; ++i) {
            const int32_t* p = x + size_t(i) * cols;
            for (int j = 0; j < cols; ++j) out[j] = out[j] + p[j];
        }
    }
    return now() - t0;
}

// threads an OpenMP loop of this library runs on
int xb_threads(void) {
    int n = 1;
#pragma omp parallel
    {
#pragma omp single
        n = omp_get_num_threads();
    }
    return n;
}

}  // extern "C"
