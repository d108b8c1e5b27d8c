"""Every math function name the DuckDB extension's translator can emit, compiled by hipRTC and run through
GpuEngine::ExecutePlan (a PROJECTION over a scanned column), in strict IEEE mode and with fast math on, for
DOUBLE and FLOAT inputs.

The name tables restate extension/src/expression_translator.cpp:121-170 (CudaMathFunction for DOUBLE,
CudaMathFunctionFp32 for FLOAT, where fast math swaps in the __expf-style intrinsics); the bodies have the
translator's text form (`((T*)out)[idx] = f(((const T*)inputs[k])[row]);`), and radians/degrees go through
the __vgpu_radians/__vgpu_degrees helpers of the kernel prelude (cuda_engine/src/code_generator.cpp:279-280).

Expected values come from Python's math module (the host libm) in double precision, rounded to float for
FLOAT. Stated bounds:
  * sqrt, floor, ceil, radians, degrees: bit-exact (one correctly rounded operation), both precisions (with
    fast math sqrt is the hardware approximation: within 1 ulp);
  * DOUBLE transcendentals (exp, ln, log10, log2, sin, cos, tan, asin, acos, atan, pow, atan2): within 2 ulp
    of libm (each side is within about 1 ulp of the true value), 4 ulp with fast math;
  * FLOAT strict (expf, logf, ... powf, atan2f): within 4 ulp of the correctly rounded float;
  * pow / powf with fast math: within 32 ulp (the fast form skips the extra-precision logarithm);
  * FLOAT fast-math intrinsics (__expf, __logf, __log10f, __log2f, __sinf, __cosf, __tanf) on their
    documented ranges: relative error <= 2^-16 (exp, and the log family away from 1), absolute <= 2^-19
    (sin, cos, and the log family near 1), tan within 2^-18 x max(1, |tan|).
Special values (NaN for sqrt(-1) and log(-1), -inf for log(0), +inf for exp overflow) must match in strict
mode."""
import math
import zlib

import numpy as np
import pytest

from vectisdb_amd.engine import (GpuExpr, GpuOpType as Op, GpuPlanNode, GpuTableRef, GpuValueType as T,
                                 GpuEngine, make_column, set_gpu_fast_math_enabled)

pytestmark = pytest.mark.gpu

# DuckDB name -> (DOUBLE name, FLOAT strict name, FLOAT fast-math name); arity 1 unless listed in BINARY
NAMES = {
    "sqrt": ("sqrt", "sqrtf", "sqrtf"),
    "exp": ("exp", "expf", "__expf"),
    "ln": ("log", "logf", "__logf"),
    "log10": ("log10", "log10f", "__log10f"),
    "log": ("log10", "log10f", "__log10f"),  # DuckDB log is base 10
    "log2": ("log2", "log2f", "__log2f"),
    "sin": ("sin", "sinf", "__sinf"),
    "cos": ("cos", "cosf", "__cosf"),
    "tan": ("tan", "tanf", "__tanf"),
    "asin": ("asin", "asinf", "asinf"),
    "acos": ("acos", "acosf", "acosf"),
    "atan": ("atan", "atanf", "atanf"),
    "floor": ("floor", "floorf", "floorf"),
    "ceil": ("ceil", "ceilf", "ceilf"),
    "ceiling": ("ceil", "ceilf", "ceilf"),
    "radians": ("__vgpu_radians", "__vgpu_radians", "__vgpu_radians"),
    "degrees": ("__vgpu_degrees", "__vgpu_degrees", "__vgpu_degrees"),
    "pow": ("pow", "powf", "powf"),
    "power": ("pow", "powf", "powf"),
    "atan2": ("atan2", "atan2f", "atan2f"),
}
BINARY = {"pow", "power", "atan2"}

# host meaning of each DuckDB function, in double precision
HOST = {
    "sqrt": math.sqrt, "exp": math.exp, "ln": math.log, "log10": math.log10, "log": math.log10, "log2": math.log2,
    "sin": math.sin, "cos": math.cos, "tan": math.tan, "asin": math.asin, "acos": math.acos, "atan": math.atan,
    "floor": math.floor, "ceil": math.ceil, "ceiling": math.ceil,
    "radians": lambda d: d * 0.017453292519943295, "degrees": lambda r: r * 57.29577951308232,
    "pow": math.pow, "power": math.pow, "atan2": math.atan2,
}
EXACT = {"sqrt", "floor", "ceil", "ceiling", "radians", "degrees"}

# input ranges: the intrinsics' documented ranges for the fast-math FLOAT forms, wider otherwise
RANGES = {
    "sqrt": (0.0, 1e6), "exp": (-40.0, 40.0), "ln": (1e-3, 1e6), "log10": (1e-3, 1e6), "log": (1e-3, 1e6),
    "log2": (1e-3, 1e6), "sin": (-math.pi, math.pi), "cos": (-math.pi, math.pi), "tan": (-1.3, 1.3),
    "asin": (-1.0, 1.0), "acos": (-1.0, 1.0), "atan": (-1e3, 1e3), "floor": (-1e6, 1e6), "ceil": (-1e6, 1e6),
    "ceiling": (-1e6, 1e6), "radians": (-1e4, 1e4), "degrees": (-100.0, 100.0), "pow": (0.1, 10.0),
    "power": (0.1, 10.0), "atan2": (-100.0, 100.0),
}
N = 4096


def _ulps(got, want, dtype):
    """Distance in units in the last place of `dtype` (inf where exactly one side is not finite)."""
    got = np.asarray(got, dtype)
    want = np.asarray(want, dtype)
    itype = np.int32 if dtype == np.float32 else np.int64
    g = got.view(itype).astype(np.int64)
    w = want.view(itype).astype(np.int64)
    # map the sign-magnitude bit patterns onto a monotonic integer line
    bias = np.int64(np.iinfo(itype).min)
    g = np.where(g < 0, bias - g, g)
    w = np.where(w < 0, bias - w, w)
    d = np.abs(g - w).astype(np.float64)
    both = np.isfinite(got) & np.isfinite(want)
    same_special = (~np.isfinite(got)) & (~np.isfinite(want)) & ((got == want) | (np.isnan(got) & np.isnan(want)))
    return np.where(both, d, np.where(same_special, 0.0, np.inf))


def _inputs(name, dtype, seed):
    rng = np.random.default_rng(seed)
    lo, hi = RANGES[name]
    x = rng.uniform(lo, hi, N)
    if name in ("floor", "ceil", "ceiling"):
        x[:8] = [-2.5, -2.0, -0.5, -0.0, 0.0, 0.5, 2.0, 2.5]
    y = rng.uniform(-3.0, 3.0, N) if name in ("pow", "power") else rng.uniform(lo, hi, N)
    return x.astype(dtype), y.astype(dtype)


def _run(fname, dtype, x, y, binary):
    ctype, vt = ("double", T.FLOAT64) if dtype == np.float64 else ("float", T.FLOAT32)
    args = f"((const {ctype}*)inputs[0])[row]"
    cols = [make_column("x", vt, x)]
    inputs = ["x"]
    if binary:
        args += f", ((const {ctype}*)inputs[1])[row]"
        cols.append(make_column("y", vt, y))
        inputs.append("y")
    body = f"(({ctype}*)out)[idx] = {fname}({args});"
    scan = GpuPlanNode(Op.SCAN, table=GpuTableRef("memory", "main", "m"), column_names=inputs)
    plan = GpuPlanNode(Op.PROJECTION, children=[scan], expressions=[GpuExpr(body, inputs, vt)])
    r = GpuEngine.instance().execute_plan(plan, cols)
    assert r.success, f"{fname}: {r.error_message}"
    return np.asarray(r.columns[0].data, dtype)


def _host(name, x, y, dtype):
    f = HOST[name]
    xs = x.astype(np.float64)
    ys = y.astype(np.float64)
    out = np.array([f(a, b) if name in BINARY else f(a) for a, b in zip(xs.tolist(), ys.tolist())], np.float64)
    return out.astype(dtype)


@pytest.fixture
def fast_math(request):
    set_gpu_fast_math_enabled(request.param)
    yield request.param
    set_gpu_fast_math_enabled(False)


@pytest.mark.parametrize("fast_math", [False, True], indirect=True, ids=["strict", "fast_math"])
@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("name", sorted(NAMES))
def test_translator_math_name(dev, name, precision, fast_math):
    dtype = np.float64 if precision == "double" else np.float32
    double_name, float_name, fast_name = NAMES[name]
    fname = double_name if precision == "double" else (fast_name if fast_math else float_name)
    x, y = _inputs(name, dtype, seed=zlib.crc32(f"{name}-{precision}".encode()))
    got = _run(fname, dtype, x, y, name in BINARY)
    want = _host(name, x, y, dtype)
    if name in EXACT:
        if fast_math and name == "sqrt":  # -ffast-math drops the correctly rounded square root
            assert _ulps(got, want, dtype).max() <= 1, fname
        else:
            np.testing.assert_array_equal(got, want, err_msg=fname)
        return
    if fast_math and name in ("pow", "power"):  # fast pow: log2/exp2 without the extra-precision steps
        assert _ulps(got, want, dtype).max() <= 32, (fname, _ulps(got, want, dtype).max())
        return
    if precision == "double":
        bound = 4 if fast_math else 2
        assert _ulps(got, want, dtype).max() <= bound, (fname, _ulps(got, want, dtype).max())
        return
    if not fast_math or fname == float_name:
        assert _ulps(got, want, dtype).max() <= 4, (fname, _ulps(got, want, dtype).max())
        return
    # fast-math FLOAT intrinsics: hardware approximations with stated absolute/relative bounds
    g = got.astype(np.float64)
    w = _host(name, x, y, np.float64)
    if name in ("sin", "cos"):
        assert np.abs(g - w).max() <= 2.0 ** -19, fname
    elif name == "tan":
        assert (np.abs(g - w) / np.maximum(np.abs(w), 1.0)).max() <= 2.0 ** -18, fname
    elif name in ("ln", "log10", "log", "log2"):
        near_one = np.abs(x.astype(np.float64) - 1.0) < 0.5
        assert np.abs(g - w)[near_one].max(initial=0.0) <= 2.0 ** -19, fname
        assert (np.abs(g - w) / np.abs(w))[~near_one].max(initial=0.0) <= 2.0 ** -16, fname
    else:  # exp
        assert (np.abs(g - w) / np.abs(w)).max() <= 2.0 ** -16, fname


def test_special_values_strict(dev):
    set_gpu_fast_math_enabled(False)
    for precision, dtype in (("double", np.float64), ("float", np.float32)):
        x = np.array([-1.0, 0.0, 1000.0, -0.0], dtype)
        i = 0 if precision == "double" else 1
        s = _run(NAMES["sqrt"][i], dtype, x, x, False)
        assert np.isnan(s[0]) and s[1] == 0.0 and s[3] == 0.0 and np.signbit(s[3])
        lg = _run(NAMES["ln"][i], dtype, x, x, False)
        assert np.isnan(lg[0]) and lg[1] == -np.inf and lg[3] == -np.inf
        e = _run(NAMES["exp"][i], dtype, x, x, False)
        assert e[2] == np.inf and e[1] == 1.0
