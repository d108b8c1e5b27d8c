"""Pins the CPU oracle (oracle/oracle.c) before results are checked against it. CPU only.

Each oracle function is checked against something independent of its own formula: Python's math.fsum
(Shewchuk's exact summation, a library routine), exact rational arithmetic (fractions.Fraction), plain
Python brute force on small inputs, closed forms, and the reference's own known answers (tests/golden).
"""
import json
import math
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

import oracle

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_exact_sum_matches_fsum_and_fractions():
    rng = np.random.default_rng(0)
    cases = [rng.standard_normal(1000), rng.standard_normal(10) * 1e300, np.array([1e16, 1.0, -1e16]),
             np.array([0.1] * 10), np.array([1.0, 1e100, 1.0, -1e100]), np.array([]),
             np.array([2.0 ** -1074] * 3), rng.uniform(-1, 1, 5000) * 10.0 ** rng.integers(-20, 20, 5000)]
    for x in cases:
        got = oracle.exact_sum(x)
        assert got == math.fsum(x)
        if len(x) and np.all(np.isfinite(x)):
            exact = sum((Fraction(float(v)) for v in x), Fraction(0))
            assert got == float(exact)  # Fraction -> float is correctly rounded


def test_exact_sum_specials():
    assert math.isinf(oracle.exact_sum(np.array([1.0, np.inf, 2.0])))
    assert math.isnan(oracle.exact_sum(np.array([np.inf, -np.inf])))


def _brute_group_by(key, v1, v2):
    groups = {}
    for k, a, b in zip(key.tolist(), v1.tolist(), v2.tolist()):
        g = groups.setdefault(k, [0, 0, None, None, []])
        g[0] += 1
        g[1] += a
        g[2] = a if g[2] is None else min(g[2], a)
        g[3] = a if g[3] is None else max(g[3], a)
        g[4].append(Fraction(b))
    out = {}
    for k, (c, s, mn, mx, fr) in groups.items():
        exact = sum(fr, Fraction(0))
        out[k] = (c, s, mn, mx, float(exact), float(exact / c) if False else float(exact) / c)
    return out


@pytest.mark.parametrize("n,distinct", [(1, 1), (50, 3), (2000, 37), (5000, 5000)])
def test_group_by_matches_brute_force(n, distinct):
    rng = np.random.default_rng(n)
    key = rng.integers(-distinct, distinct, n).astype(np.int64)
    v1 = rng.integers(-10**6, 10**6, n).astype(np.int64)
    v2 = rng.standard_normal(n)
    res = oracle.group_by_i64(key, v1, v2)
    exp = _brute_group_by(key, v1, v2)
    assert list(res["key"]) == sorted(exp)
    for i, k in enumerate(res["key"].tolist()):
        c, s, mn, mx, s2, avg = exp[k]
        assert res["count"][i] == c and res["sum_v1"][i] == s
        assert res["min_v1"][i] == mn and res["max_v1"][i] == mx
        assert res["sum_v2"][i] == s2
        assert res["avg_v2"][i] == avg


def test_group_by_int128_sum_no_overflow():
    key = np.zeros(4, np.int64)
    v1 = np.array([2**62, 2**62, 2**62, 2**62], np.int64)
    res = oracle.group_by_i64(key, v1, np.zeros(4))
    assert res["sum_v1"][0] == 2**64


def test_group_by_wide_key_range_uses_sort_path():
    key = np.array([2**62, -2**62, 5, 2**62, 5], np.int64)
    res = oracle.group_by_i64(key, np.arange(5, dtype=np.int64), np.ones(5))
    assert list(res["key"]) == [-2**62, 5, 2**62]
    assert list(res["count"]) == [1, 2, 2]


def test_filter_project_cfg0_matches_numpy():
    rng = np.random.default_rng(1)
    n = 100_000
    a = rng.integers(0, 10**6, n).astype(np.int32)
    b = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    c = rng.random(n)
    o1, o2 = oracle.filter_project_cfg0(a, b, c)
    keep = (a < 500000) & (c > 0.25)
    np.testing.assert_array_equal(o1, (a[keep] * 2).astype(np.int64) + b[keep])
    np.testing.assert_array_equal(o2, c[keep] * 1.5 + a[keep].astype(np.float64))
    assert abs(keep.mean() - 0.375) < 0.01


def test_join_sum_count_closed_form_and_brute_force():
    rng = np.random.default_rng(2)
    nb = 1 << 12
    build = rng.permutation(nb).astype(np.int64)
    payload = rng.integers(0, 1000, nb).astype(np.int32)
    fk = np.where(rng.random(50_000) < 0.9, rng.integers(0, nb, 50_000), rng.integers(nb, 2 * nb, 50_000))
    fk = fk.astype(np.int64)
    s, c = oracle.join_sum_count(build, payload, fk)
    # closed form: keys are a permutation of [0, nb) -> inverse permutation gives each key's payload
    inv = np.empty(nb, np.int64)
    inv[build] = np.arange(nb)
    hit = fk < nb
    assert c == int(hit.sum())
    assert s == int(payload[inv[fk[hit]]].astype(np.int64).sum())
    # duplicates on the build side: brute force
    b2 = rng.integers(0, 50, 300).astype(np.int64)
    p2 = rng.integers(0, 1000, 300).astype(np.int32)
    f2 = rng.integers(0, 60, 400).astype(np.int64)
    s2, c2 = oracle.join_sum_count(b2, p2, f2)
    es = ec = 0
    for f in f2.tolist():
        m = b2 == f
        ec += int(m.sum())
        es += int(p2[m].astype(np.int64).sum())
    assert (s2, c2) == (es, ec)


def test_join_pairs_brute_force():
    rng = np.random.default_rng(3)
    left = rng.integers(0, 20, 200).astype(np.int64)
    right = rng.integers(0, 20, 150).astype(np.int64)
    l, r = oracle.join_pairs(left, right)
    exp = sorted((i, j) for i in range(len(left)) for j in range(len(right)) if left[i] == right[j])
    assert sorted(zip(l.tolist(), r.tolist())) == exp


def test_cross_filter_closed_form_vs_brute_force():
    rng = np.random.default_rng(4)
    x = rng.integers(0, 1 << 30, 300).astype(np.int32)
    y = rng.integers(0, 1 << 30, 700).astype(np.int32)
    n, sx, sy = oracle.cross_filter_count(x, y, 1 << 30)
    ox, oy = oracle.cross_filter_pairs(x, y, 1 << 30)
    m = (x[:, None].astype(np.int64) + y[None, :]) < (1 << 30)
    assert n == int(m.sum()) == len(ox)
    assert sx == int((x[:, None].astype(np.int64) * m).sum()) and sy == int((y[None, :].astype(np.int64) * m).sum())
    # left-major order
    ii, jj = np.nonzero(m)
    np.testing.assert_array_equal(ox, x[ii])
    np.testing.assert_array_equal(oy, y[jj])


def test_generator_is_counter_based_and_in_range():
    a = oracle.gen_uniform_i64(1000, -5, 5, 7, 1)
    assert a.min() >= -5 and a.max() <= 5
    # row i depends only on (seed, stream, i)
    b = oracle.gen_uniform_i64(10, -5, 5, 7, 1)
    np.testing.assert_array_equal(a[:10], b)
    u = oracle.gen_uniform_f64(100_000, 3, 3)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01
    # the mixer is splitmix64's finalizer over (seed ^ stream * C1) + (i + 1) * golden
    z = ((7 ^ (1 * 0xD1B54A32D192ED03)) + 1 * 0x9E3779B97F4A7C15) & (2**64 - 1)
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
    z ^= z >> 31
    assert oracle.mix_counter(7, 1, 0) == z


def test_reference_known_answers_group_by():
    """GROUP BY cases of the reference's executor tests (cuda_engine/tests/test_gpu_executor.cpp), restated
    as data in tests/golden/reference_groupby.json."""
    cases = json.loads((GOLDEN / "reference_groupby.json").read_text())
    for case in cases:
        key = np.array(case["key"], np.int64)
        v = np.array(case["value"], np.int64)
        res = oracle.group_by_i64(key, v, v.astype(np.float64))
        got = {int(k): (int(c), int(s)) for k, c, s in zip(res["key"], res["count"], res["sum_v1"])}
        exp = {int(k): tuple(x) for k, x in case["expected"].items()}
        assert got == exp, case["name"]


@pytest.mark.parametrize("threads,n,distinct", [(2, 10, 3), (8, 100_000, 1000), (16, 300_000, 70_000),
                                                (5, 50_001, 1), (8, 1000, 1000)])
def test_multithreaded_group_by_matches_single_threaded(threads, n, distinct):
    """The multithreaded restatement (the CPU baseline the bench times) returns exactly the single-threaded
    oracle's groups: same keys and order, counts, int128 sums, min/max and the same double bits."""
    rng = np.random.default_rng(n + threads)
    key = (rng.zipf(1.3, n) % distinct - distinct // 2).astype(np.int64)
    v1 = rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64)
    v2 = rng.standard_normal(n) * 10.0 ** rng.integers(-5, 5, n)
    a = oracle.group_by_i64(key, v1, v2)
    b = oracle.group_by_i64(key, v1, v2, threads=threads)
    for k in ("key", "count", "min_v1", "max_v1"):
        np.testing.assert_array_equal(a[k], b[k])
    assert a["sum_v1"] == b["sum_v1"]
    np.testing.assert_array_equal(a["sum_v2"].view(np.int64), b["sum_v2"].view(np.int64))
    np.testing.assert_array_equal(a["avg_v2"].view(np.int64), b["avg_v2"].view(np.int64))


def test_multithreaded_group_by_wide_range_falls_back():
    """Keys spanning more than 2^24 take the single-threaded path; the answer is the same."""
    rng = np.random.default_rng(3)
    key = rng.integers(-(1 << 40), 1 << 40, 5000).astype(np.int64)
    v1 = rng.integers(-100, 100, 5000).astype(np.int64)
    v2 = rng.random(5000)
    a = oracle.group_by_i64(key, v1, v2)
    b = oracle.group_by_i64(key, v1, v2, threads=4)
    np.testing.assert_array_equal(a["key"], b["key"])
    np.testing.assert_array_equal(a["sum_v2"], b["sum_v2"])


# ---------------------------------------------------------------- multithreaded CPU baselines
# Each _mt function must return exactly what its single-threaded restatement returns (itself pinned above).

@pytest.mark.parametrize("threads,n", [(1, 0), (3, 1), (8, 100_003), (16, 7)])
def test_filter_project_mt_matches_single_threaded(threads, n):
    rng = np.random.default_rng(n + threads)
    a = rng.integers(0, 10**6, n).astype(np.int32)
    b = rng.integers(-(1 << 40), 1 << 40, n).astype(np.int64)
    c = rng.random(n)
    o1, o2 = oracle.filter_project_cfg0(a, b, c)
    m1, m2, _ = oracle.filter_project_cfg0_mt(a, b, c, threads)
    np.testing.assert_array_equal(m1, o1)
    assert m2.tobytes() == o2.tobytes()


@pytest.mark.parametrize("threads", [1, 4, 8])
def test_join_hash_mt_matches_sorted_restatement(threads):
    rng = np.random.default_rng(threads)
    nb = 20_000
    build = rng.integers(-3000, 3000, nb).astype(np.int64)       # duplicate build keys, negative keys
    payload = rng.integers(-1000, 1000, nb).astype(np.int32)
    fk = rng.integers(-5000, 5000, 200_000).astype(np.int64)
    s, n = oracle.join_sum_count(build, payload, fk)
    s2, n2, _ = oracle.join_sum_count_mt(build, payload, fk, threads)
    assert (s2, n2) == (s, n)
    # a permutation build side: the closed form
    perm = rng.permutation(1 << 14).astype(np.int64)
    pay = rng.integers(0, 1000, 1 << 14).astype(np.int32)
    fk = rng.integers(0, 1 << 15, 100_000).astype(np.int64)
    inv = np.empty(1 << 14, np.int64)
    inv[perm] = np.arange(1 << 14)
    hit = fk < (1 << 14)
    s3, n3, _ = oracle.join_sum_count_mt(perm, pay, fk, threads)
    assert n3 == int(hit.sum()) and s3 == int(pay[inv[fk[hit]]].astype(np.int64).sum())
    # empty sides
    assert oracle.join_sum_count_mt(perm, pay, fk[:0], threads)[:2] == (0, 0)
    assert oracle.join_sum_count_mt(perm[:0], pay[:0], fk, threads)[:2] == (0, 0)


@pytest.mark.parametrize("threads,nl,nr", [(1, 5, 7), (4, 33, 1000), (16, 64, 4096), (8, 0, 10)])
def test_cross_pairs_mt_matches_single_threaded(threads, nl, nr):
    rng = np.random.default_rng(nl * 31 + nr)
    x = rng.integers(0, 1 << 30, nl).astype(np.int32)
    y = rng.integers(0, 1 << 30, nr).astype(np.int32)
    ex, ey = oracle.cross_filter_pairs(x, y, 1 << 30)
    gx, gy, _ = oracle.cross_filter_pairs_mt(x, y, 1 << 30, threads)
    np.testing.assert_array_equal(gx, ex)
    np.testing.assert_array_equal(gy, ey)


@pytest.mark.parametrize("threads", [1, 6])
def test_batched_group_by_mt_matches_per_query_restatement(threads):
    rng = np.random.default_rng(threads + 100)
    sizes = np.array([0, 1, 5000, 17, 40_000, 3])
    n = int(sizes.sum())
    a = rng.integers(0, 10**6, n).astype(np.int32)
    c = rng.random(n)
    key = rng.integers(0, 300, n).astype(np.int64)
    v1 = rng.integers(-10**6, 10**6, n).astype(np.int64)
    v2 = rng.standard_normal(n)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    res, goffs, _ = oracle.batched_filter_group_by_mt(a, c, key, v1, v2, offs, threads)
    assert goffs[0] == 0 and goffs[-1] == len(res["key"])
    for q in range(len(sizes)):
        r0, r1 = offs[q], offs[q + 1]
        keep = (a[r0:r1] < 500000) & (c[r0:r1] > 0.25)
        g0, g1 = goffs[q], goffs[q + 1]
        if not keep.any():
            assert g0 == g1
            continue
        ref = oracle.group_by_i64(key[r0:r1][keep], v1[r0:r1][keep], v2[r0:r1][keep], raw=True)
        for k in ("key", "count", "min_v1", "max_v1", "sum_v1"):
            np.testing.assert_array_equal(res[k][g0:g1], ref[k])
        assert res["sum_v2"][g0:g1].tobytes() == ref["sum_v2"].tobytes()
        assert res["avg_v2"][g0:g1].tobytes() == ref["avg_v2"].tobytes()
