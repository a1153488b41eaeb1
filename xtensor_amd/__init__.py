"""xtensor_amd — MI355X (gfx950) evaluator for xtensor's lazy expressions.

`import xtensor_amd as xt` gives xtensor's names over device arrays:
`xt.noalias(out).assign(a * b + xt.exp(-a * a))`, `xt.sum(e, [1])`,
`xt.cumsum(e, 1)`, `xt.view(y, xt.all_(), xt.range_(0, 56, 2))` ...
Every evaluation runs hand-written HIP kernels in `lib/libxg.so` through the C
ABI of `include/xg.h`; there is no CPU fallback.
"""

from .core import (  # noqa: F401
    BroadcastError, DeviceArray, DeviceError, Expr, Func, Reducer, Scalar, Scan, assign_checked, assign_reduce,
    Filtration, IndexView, abs, acos, all_, amax, amin, argwhere, asarray, asin, assign, atan, atan2, broadcast,
    cast, cbrt, ceil, clip, count_nonzero, filter, filtration, functor_view, imag, index_view, masked_view, real,
    cos, cosh, cumprod, cumsum, dtype_code, dtype_name, dump_npy, empty, equal, erf, eval, exp, exp2, expm1,
    expr_info, fabs, floor, fmod, from_numpy, greater, hypot, last_kernel, less, log, log1p, log2,
    load_npy, log10, logical_and, logical_not, logical_or, maximum, mean, minimum, minmax, newaxis, noalias,
    npy_info,
    norm_l0, norm_l1, norm_l2, norm_linf, norm_sq, not_equal, numpy_dtype, plan, Plan, pow, prod, range_,
    reduce_info, reshape_view, round, scalar, serialize, sin, sinh, sqrt, stddev, strided_view, sum, tan,
    tanh, transpose, trunc, variance, view, where, zeros,
)

from . import blockwise  # noqa: F401,E402
from .chunked import ChunkedArray, chunked_array, is_chunked  # noqa: F401,E402
from .batch import Batch  # noqa: F401,E402
from .graph import PlanGraph  # noqa: F401,E402
from .core import execute_many  # noqa: F401,E402
from .serve import Server  # noqa: F401,E402

__version__ = "0.1.0"
