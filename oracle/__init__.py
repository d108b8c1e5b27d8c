"""CPU oracle for the vectis engine. TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package, and only to
check results against (or as the timed CPU baseline). Nothing under vectisdb_amd/ imports it: the product path fails
loudly when its HIP library is missing instead of falling back to this code.

Pinning (what keeps this restatement honest), per function -- see tests/test_oracle.py:
  * exact_sum: math.fsum on random, cancelling and special inputs (same published algorithm), plus
    closed forms (sum of k copies of 0.1 vs exact rational rounding).
  * group_by: brute force per-key Python loops on small inputs (dict + fractions.Fraction for the exact
    double sum), invariants (counts sum to n, min <= avg <= max), and the reference's own known answers
    (cuda_engine/tests/test_gpu_executor.cpp GROUP BY cases, restated in tests/golden/).
  * filter_project_cfg0: numpy float64 elementwise (IEEE, no contraction) on the same rows.
  * join_sum_count / join_pairs: closed form for permutation build keys (COUNT = #fk < Nb, SUM via the
    inverse permutation) and nested-loop brute force on tiny inputs; the reference's join known answers.
  * cross_filter_*: nested-loop brute force on small shapes and the left-major order of the reference's
    cross-product test (cuda_engine/tests/test_gpu_executor.cpp:1171-1175).
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
SRC = HERE / "oracle.c"
LIB = HERE / "lib" / "liboracle.so"


def build() -> Path:
    LIB.parent.mkdir(exist_ok=True)
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC", "-pthread", "-o", str(LIB),
           str(SRC), "-lm"]
    subprocess.run(cmd, check=True)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists() or LIB.stat().st_mtime < SRC.stat().st_mtime:
            build()
        L = C.CDLL(str(LIB))
        P = C.c_void_p
        u64 = C.c_uint64
        L.vo_mix_counter.restype = u64
        L.vo_mix_counter.argtypes = [u64, u64, u64]
        L.vo_gen_uniform_i64.argtypes = [P, u64, C.c_int64, C.c_int64, u64, u64]
        L.vo_gen_uniform_i32.argtypes = [P, u64, C.c_int64, C.c_int64, u64, u64]
        L.vo_gen_uniform_f64.argtypes = [P, u64, u64, u64]
        L.vo_filter_project_cfg0.restype = u64
        L.vo_filter_project_cfg0.argtypes = [P, P, P, u64, C.c_int32, C.c_double, P, P]
        L.vo_exact_sum.restype = C.c_double
        L.vo_exact_sum.argtypes = [P, u64]
        L.vo_group_by_i64.restype = u64
        L.vo_group_by_i64.argtypes = [P, P, P, u64, P, P, P, P, P, P, P]
        L.vo_group_by_i64_mt.restype = u64
        L.vo_group_by_i64_mt.argtypes = [P, P, P, u64, C.c_int, P, P, P, P, P, P, P]
        L.vo_join_sum_count.restype = u64
        L.vo_join_sum_count.argtypes = [P, P, u64, P, u64, C.POINTER(u64), C.POINTER(C.c_int64)]
        L.vo_join_pairs.restype = u64
        L.vo_join_pairs.argtypes = [P, u64, P, u64, P, P, u64]
        L.vo_cross_filter_count.restype = u64
        L.vo_cross_filter_count.argtypes = [P, u64, P, u64, C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.vo_cross_filter_pairs.restype = u64
        L.vo_cross_filter_pairs.argtypes = [P, u64, P, u64, C.c_int64, P, P, u64]
        L.vo_filter_project_cfg0_mt.restype = u64
        L.vo_filter_project_cfg0_mt.argtypes = [P, P, P, u64, C.c_int32, C.c_double, P, P, C.c_int]
        L.vo_join_sum_count_mt.restype = u64
        L.vo_join_sum_count_mt.argtypes = [P, P, u64, P, u64, C.c_int, C.POINTER(u64), C.POINTER(C.c_int64)]
        L.vo_cross_filter_pairs_mt.restype = u64
        L.vo_cross_filter_pairs_mt.argtypes = [P, u64, P, u64, C.c_int64, P, P, u64, C.c_int]
        L.vo_batched_filter_group_by_mt.restype = u64
        L.vo_batched_filter_group_by_mt.argtypes = [P, P, P, P, P, P, u64, C.c_int32, C.c_double, C.c_int,
                                                    P, P, P, P, P, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else C.c_void_p(0)


def mix_counter(seed, stream, i):
    return lib().vo_mix_counter(seed, stream, i)


def gen_uniform_i64(n, lo, hi, seed, stream):
    out = np.empty(n, np.int64)
    lib().vo_gen_uniform_i64(_p(out), n, lo, hi, seed, stream)
    return out


def gen_uniform_i32(n, lo, hi, seed, stream):
    out = np.empty(n, np.int32)
    lib().vo_gen_uniform_i32(_p(out), n, lo, hi, seed, stream)
    return out


def gen_uniform_f64(n, seed, stream):
    out = np.empty(n, np.float64)
    lib().vo_gen_uniform_f64(_p(out), n, seed, stream)
    return out


def filter_project_cfg0(a, b, c, a_lt=500000, c_gt=0.25):
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int64)
    c = np.ascontiguousarray(c, np.float64)
    o1 = np.empty(len(a), np.int64)
    o2 = np.empty(len(a), np.float64)
    k = lib().vo_filter_project_cfg0(_p(a), _p(b), _p(c), len(a), a_lt, c_gt, _p(o1), _p(o2))
    return o1[:k], o2[:k]


def exact_sum(x):
    x = np.ascontiguousarray(x, np.float64)
    return lib().vo_exact_sum(_p(x), len(x))


def group_by_i64(key, v1=None, v2=None, threads=1, raw=False):
    """Returns dict of per-group arrays in ascending key order (sum_v1 as Python ints, or with raw=True as a
    (groups, 2) int64 array of HUGEINT {lower, upper} words). threads > 1 runs the multithreaded restatement
    (same results)."""
    key = np.ascontiguousarray(key, np.int64)
    n = len(key)
    v1 = None if v1 is None else np.ascontiguousarray(v1, np.int64)
    v2 = None if v2 is None else np.ascontiguousarray(v2, np.float64)
    cap = max(n, 1)
    out = {k: np.empty(cap, np.int64) for k in ("key", "count", "min_v1", "max_v1")}
    out["sum_v1"] = np.empty(2 * cap, np.int64)
    out["sum_v2"] = np.empty(cap, np.float64)
    out["avg_v2"] = np.empty(cap, np.float64)
    if threads > 1:
        g = lib().vo_group_by_i64_mt(_p(key), _p(v1), _p(v2), n, threads, _p(out["key"]), _p(out["count"]),
                                     _p(out["sum_v1"]), _p(out["min_v1"]), _p(out["max_v1"]), _p(out["sum_v2"]),
                                     _p(out["avg_v2"]))
    else:
        g = lib().vo_group_by_i64(_p(key), _p(v1), _p(v2), n, _p(out["key"]), _p(out["count"]), _p(out["sum_v1"]),
                                  _p(out["min_v1"]), _p(out["max_v1"]), _p(out["sum_v2"]), _p(out["avg_v2"]))
    res = {k: v[:g] for k, v in out.items() if k != "sum_v1"}
    pairs = out["sum_v1"][: 2 * g].reshape(-1, 2)
    res["sum_v1"] = pairs if raw else [int(lo) % (1 << 64) + (int(hi) << 64) for lo, hi in pairs]
    return res


def time_group_by_i64(key, v1, v2, threads=1):
    """Seconds taken by the C GROUP BY alone (outputs allocated beforehand, no Python conversion): the CPU
    baseline timing. Returns (seconds, groups)."""
    import time
    key = np.ascontiguousarray(key, np.int64)
    v1 = np.ascontiguousarray(v1, np.int64)
    v2 = np.ascontiguousarray(v2, np.float64)
    n = len(key)
    cap = max(n, 1)
    out = {k: np.empty(cap, np.int64) for k in ("key", "count", "min_v1", "max_v1")}
    out["sum_v1"] = np.empty(2 * cap, np.int64)
    out["sum_v2"] = np.empty(cap, np.float64)
    out["avg_v2"] = np.empty(cap, np.float64)
    t0 = time.perf_counter()
    g = lib().vo_group_by_i64_mt(_p(key), _p(v1), _p(v2), n, threads, _p(out["key"]), _p(out["count"]),
                                 _p(out["sum_v1"]), _p(out["min_v1"]), _p(out["max_v1"]), _p(out["sum_v2"]),
                                 _p(out["avg_v2"]))
    return time.perf_counter() - t0, g


def join_sum_count(build_key, payload, fk):
    bk = np.ascontiguousarray(build_key, np.int64)
    pl = np.ascontiguousarray(payload, np.int32)
    fk = np.ascontiguousarray(fk, np.int64)
    lo = C.c_uint64(0)
    hi = C.c_int64(0)
    cnt = lib().vo_join_sum_count(_p(bk), _p(pl), len(bk), _p(fk), len(fk), C.byref(lo), C.byref(hi))
    return (hi.value << 64) + lo.value, cnt


def join_pairs(left, right):
    left = np.ascontiguousarray(left, np.int64)
    right = np.ascontiguousarray(right, np.int64)
    n = lib().vo_join_pairs(_p(left), len(left), _p(right), len(right), C.c_void_p(0), C.c_void_p(0), 0)
    ol = np.empty(max(n, 1), np.uint64)
    orr = np.empty(max(n, 1), np.uint64)
    lib().vo_join_pairs(_p(left), len(left), _p(right), len(right), _p(ol), _p(orr), n)
    return ol[:n], orr[:n]


def cross_filter_count(x, y, c):
    x = np.ascontiguousarray(x, np.int32)
    y = np.ascontiguousarray(y, np.int32)
    sx = C.c_int64(0)
    sy = C.c_int64(0)
    n = lib().vo_cross_filter_count(_p(x), len(x), _p(y), len(y), c, C.byref(sx), C.byref(sy))
    return n, sx.value, sy.value


def cross_filter_pairs(x, y, c):
    x = np.ascontiguousarray(x, np.int32)
    y = np.ascontiguousarray(y, np.int32)
    n, _, _ = cross_filter_count(x, y, c)
    ox = np.empty(max(n, 1), np.int32)
    oy = np.empty(max(n, 1), np.int32)
    lib().vo_cross_filter_pairs(_p(x), len(x), _p(y), len(y), c, _p(ox), _p(oy), n)
    return ox[:n], oy[:n]


def _timed(fn, *args):
    import time
    t0 = time.perf_counter()
    r = fn(*args)
    return r, time.perf_counter() - t0


def filter_project_cfg0_mt(a, b, c, threads, a_lt=500000, c_gt=0.25):
    """(out1, out2, seconds of the C call): vo_filter_project_cfg0 on `threads` threads."""
    a = np.ascontiguousarray(a, np.int32)
    b = np.ascontiguousarray(b, np.int64)
    c = np.ascontiguousarray(c, np.float64)
    o1 = np.empty(max(len(a), 1), np.int64)
    o2 = np.empty(max(len(a), 1), np.float64)
    k, sec = _timed(lib().vo_filter_project_cfg0_mt, _p(a), _p(b), _p(c), len(a), a_lt, c_gt, _p(o1), _p(o2),
                    threads)
    return o1[:k], o2[:k], sec


def join_sum_count_mt(build_key, payload, fk, threads):
    """(SUM(payload), COUNT(*), seconds of the C call): the hash-join restatement on `threads` threads."""
    bk = np.ascontiguousarray(build_key, np.int64)
    pl = np.ascontiguousarray(payload, np.int32)
    fk = np.ascontiguousarray(fk, np.int64)
    lo = C.c_uint64(0)
    hi = C.c_int64(0)
    cnt, sec = _timed(lib().vo_join_sum_count_mt, _p(bk), _p(pl), len(bk), _p(fk), len(fk), threads, C.byref(lo),
                      C.byref(hi))
    return (hi.value << 64) + lo.value, cnt, sec


def cross_filter_pairs_mt(x, y, c, threads):
    """(xs, ys, seconds of the C call): the materialized left-major pairs on `threads` threads."""
    x = np.ascontiguousarray(x, np.int32)
    y = np.ascontiguousarray(y, np.int32)
    n, _, _ = cross_filter_count(x, y, c)
    ox = np.empty(max(n, 1), np.int32)
    oy = np.empty(max(n, 1), np.int32)
    k, sec = _timed(lib().vo_cross_filter_pairs_mt, _p(x), len(x), _p(y), len(y), c, _p(ox), _p(oy), n, threads)
    assert k == n
    return ox[:n], oy[:n], sec


def batched_filter_group_by_mt(a, c, key, v1, v2, offsets, threads, a_lt=500000, c_gt=0.25):
    """(dict of per-group arrays as group_by_i64(raw=True), group offsets per query, seconds of the C call)."""
    a = np.ascontiguousarray(a, np.int32)
    c = np.ascontiguousarray(c, np.float64)
    key = np.ascontiguousarray(key, np.int64)
    v1 = np.ascontiguousarray(v1, np.int64)
    v2 = np.ascontiguousarray(v2, np.float64)
    offsets = np.ascontiguousarray(offsets, np.int64)
    nq = len(offsets) - 1
    cap = max(len(key), 1)
    out = {k: np.empty(cap, np.int64) for k in ("key", "count", "min_v1", "max_v1")}
    out["sum_v1"] = np.empty(2 * cap, np.int64)
    out["sum_v2"] = np.empty(cap, np.float64)
    out["avg_v2"] = np.empty(cap, np.float64)
    goffs = np.empty(nq + 1, np.uint64)
    g, sec = _timed(lib().vo_batched_filter_group_by_mt, _p(a), _p(c), _p(key), _p(v1), _p(v2), _p(offsets), nq,
                    a_lt, c_gt, threads, _p(out["key"]), _p(out["count"]), _p(out["sum_v1"]), _p(out["min_v1"]),
                    _p(out["max_v1"]), _p(out["sum_v2"]), _p(out["avg_v2"]), _p(goffs))
    res = {k: v[:g] for k, v in out.items() if k != "sum_v1"}
    res["sum_v1"] = out["sum_v1"][: 2 * g].reshape(-1, 2)
    return res, goffs.astype(np.int64), sec
