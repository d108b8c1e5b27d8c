"""Fused FILTER + PROJECTION (hipRTC) and the fused cross-product filter, through the C ABI.

Bar: bit-exact for integer outputs, row sets and order; doubles bit-exact for +,-,*,/ and casts (IEEE
round-to-nearest, no contraction on either side); transcendentals within 2 ulp of the host libm (device
and host libm are different correctly-rounded-or-nearly implementations).
"""
import math

import numpy as np
import pytest

from vectisdb_amd import capi
from vectisdb_amd.workloads import cast, cfg0_exprs, col, copy_body, dlit, out

pytestmark = pytest.mark.gpu


def _dev(a, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev)


def cfg0_data(n, seed=0):
    import oracle
    a = oracle.gen_uniform_i32(n, 0, 999_999, seed, 1)
    b = oracle.gen_uniform_i64(n, -(1 << 40), (1 << 40) - 1, seed, 2)
    c = oracle.gen_uniform_f64(n, seed, 3)
    return a, b, c


@pytest.mark.parametrize("n", [1, 5, 2047, 2048, 2049, 100_003, 1_048_576, (1 << 22) - 1, (1 << 22) + 8_191])
def test_config0_filter_project_bit_exact(dev, n):
    """The last two sizes straddle the switch to the 8,192-row-tile variant of the generated kernel."""
    import oracle
    a, b, c = cfg0_data(n, seed=n)
    preds, projs = cfg0_exprs()
    outs, _, sel, k = capi.filter_project([_dev(a, dev), _dev(b, dev), _dev(c, dev)], preds, projs,
                                          emit_selection=True)
    e1, e2 = oracle.filter_project_cfg0(a, b, c)
    assert k == len(e1)
    np.testing.assert_array_equal(outs[0].cpu().numpy(), e1)
    np.testing.assert_array_equal(outs[1].cpu().numpy().view(np.uint64), e2.view(np.uint64))
    keep = np.nonzero((a < 500000) & (c > 0.25))[0]
    np.testing.assert_array_equal(sel.cpu().numpy().astype(np.int64), keep)


def test_device_generator_matches_host_generator(dev):
    """The device column generator and the oracle's host generator agree bit for bit (integers, uniform
    doubles), so GPU-generated tables can be checked against host-regenerated data."""
    import torch
    import oracle
    n = 300_001
    t = torch.empty(n, dtype=torch.int64, device=dev)
    capi.generate_column(t, capi.INT64, 0, lo=-(1 << 40), hi=(1 << 40) - 1, seed=5, stream_id=2)
    np.testing.assert_array_equal(t.cpu().numpy(), oracle.gen_uniform_i64(n, -(1 << 40), (1 << 40) - 1, 5, 2))
    u = torch.empty(n, dtype=torch.float64, device=dev)
    capi.generate_column(u, capi.FLOAT64, 1, seed=5, stream_id=3)
    np.testing.assert_array_equal(u.cpu().numpy(), oracle.gen_uniform_f64(n, 5, 3))


def test_projection_only_and_selection_order(dev):
    n = 70_000
    rng = np.random.default_rng(1)
    x = rng.standard_normal(n)
    y = rng.integers(-1000, 1000, n).astype(np.int32)
    body = out(f"(({col(0, capi.FLOAT64)} * 1.18) - {cast(col(1, capi.INT32), capi.FLOAT64)})", capi.FLOAT64)
    outs, _, _, k = capi.filter_project([_dev(x, dev), _dev(y, dev)], [], [(body, [0, 1], capi.FLOAT64)])
    assert k == n
    np.testing.assert_array_equal(outs[0].cpu().numpy(), x * 1.18 - y.astype(np.float64))


def test_filter_everything_out_and_empty_input(dev):
    import torch
    x = np.arange(10_000, dtype=np.int64)
    pred = (out(f"({col(0, capi.INT64)} > 99999999LL)", capi.BOOLEAN), [0], capi.BOOLEAN)
    proj = (copy_body(capi.INT64), [0], capi.INT64)
    outs, _, _, k = capi.filter_project([_dev(x, dev)], [pred], [proj])
    assert k == 0
    e = torch.empty(0, dtype=torch.int64, device=dev)
    outs, _, _, k = capi.filter_project([e], [pred], [proj])
    assert k == 0


def test_null_predicate_drops_row_and_projection_propagates_null(dev):
    import torch
    n = 5000
    rng = np.random.default_rng(2)
    x = rng.integers(0, 100, n).astype(np.int32)
    y = rng.standard_normal(n)
    xv = (rng.random(n) > 0.2).astype(np.uint8)
    yv = (rng.random(n) > 0.3).astype(np.uint8)
    pred = (out(f"({col(0, capi.INT32)} < 50)", capi.BOOLEAN), [0], capi.BOOLEAN)
    proj = (out(f"({cast(col(0, capi.INT32), capi.FLOAT64)} + {col(1, capi.FLOAT64)})", capi.FLOAT64), [0, 1],
            capi.FLOAT64)
    outs, ovalid, sel, k = capi.filter_project([_dev(x, dev), _dev(y, dev)], [pred], [proj],
                                               valid=[_dev(xv, dev), _dev(yv, dev)], out_valid=True,
                                               emit_selection=True)
    keep = np.nonzero((x < 50) & (xv == 1))[0]
    assert k == len(keep)
    np.testing.assert_array_equal(sel.cpu().numpy().astype(np.int64), keep)
    got_valid = ovalid[0].cpu().numpy().astype(bool)
    np.testing.assert_array_equal(got_valid, (xv[keep] == 1) & (yv[keep] == 1))
    vals = outs[0].cpu().numpy()
    exp = x[keep].astype(np.float64) + y[keep]
    np.testing.assert_array_equal(vals[got_valid], exp[got_valid])


def test_float32_division_bit_exact_with_denormal(dev):
    """FP32 division is correctly rounded on both sides, including a denormal result
    (cuda_engine/tests/test_code_generator.cpp:178-221)."""
    a = np.array([1.0, 3.0, 1e-38, 7.0, -2.5, 1.17549435e-38], np.float32)
    b = np.array([3.0, 7.0, 3.0, 0.1, 3.3, 4.0], np.float32)
    body = out(f"({col(0, capi.FLOAT32)} / {col(1, capi.FLOAT32)})", capi.FLOAT32)
    outs, _, _, k = capi.filter_project([_dev(a, dev), _dev(b, dev)], [], [(body, [0, 1], capi.FLOAT32)])
    np.testing.assert_array_equal(outs[0].cpu().numpy().view(np.uint32), (a / b).view(np.uint32))


def test_transcendentals_within_2ulp(dev):
    n = 20_000
    rng = np.random.default_rng(3)
    x = rng.uniform(0.1, 100.0, n)
    cases = [
        ("sin", np.sin), ("cos", np.cos), ("exp", None), ("log", np.log), ("log10", np.log10), ("sqrt", np.sqrt),
    ]
    for fn, ref in cases:
        arg = f"(double)({col(0, capi.FLOAT64)})" if fn != "exp" else f"(double)(({col(0, capi.FLOAT64)} / 10.0))"
        body = out(f"{fn}({arg})", capi.FLOAT64)
        outs, _, _, _ = capi.filter_project([_dev(x, dev)], [], [(body, [0], capi.FLOAT64)])
        got = outs[0].cpu().numpy()
        exp = np.exp(x / 10.0) if fn == "exp" else ref(x)
        ulp = np.abs(got - exp) / np.spacing(np.abs(exp))
        assert ulp.max() <= 2.0, (fn, ulp.max())
        if fn == "sqrt":
            np.testing.assert_array_equal(got, exp)  # correctly rounded on both sides


def test_radians_degrees_and_integer_modulo(dev):
    x = np.array([0.0, 90.0, 180.0, -45.5, 1e-3], np.float64)
    i = np.array([7, -7, 13, 100, -1], np.int64)
    b1 = out(f"__vgpu_radians((double)({col(0, capi.FLOAT64)}))", capi.FLOAT64)
    b2 = out(f"({col(0, capi.INT64)} % 3LL)", capi.INT64)
    outs, _, _, _ = capi.filter_project([_dev(x, dev), _dev(i, dev)], [], [(b1, [0], capi.FLOAT64),
                                                                         (b2, [1], capi.INT64)])
    np.testing.assert_array_equal(outs[0].cpu().numpy(), x * 0.017453292519943295)
    np.testing.assert_array_equal(outs[1].cpu().numpy(), np.fmod(i, 3))  # C remainder truncates toward 0


def test_cross_filter_small_brute_force(dev):
    import oracle
    rng = np.random.default_rng(4)
    for nl, nr in [(1, 1), (3, 5), (17, 1000), (64, 4099)]:
        x = rng.integers(0, 1 << 30, nl).astype(np.int32)
        y = rng.integers(0, 1 << 30, nr).astype(np.int32)
        pred = (out(f"(({col(0, capi.INT32)} + {col(1, capi.INT32)}) < 1073741824)", capi.BOOLEAN), [0, 1],
                capi.BOOLEAN)
        projs = [(copy_body(capi.INT32), [0], capi.INT32), (copy_body(capi.INT32), [1], capi.INT32)]
        outs, k = capi.cross_filter([_dev(x, dev)], [_dev(y, dev)], [pred], projs)
        ex, ey = oracle.cross_filter_pairs(x, y, 1 << 30)
        assert k == len(ex)
        np.testing.assert_array_equal(outs[0].cpu().numpy(), ex)  # left-major order preserved
        np.testing.assert_array_equal(outs[1].cpu().numpy(), ey)


@pytest.mark.parametrize("nl,nr", [(300, 16384), (700, 9001), (257, 65536 + 3)])
def test_cross_filter_large_exact(dev, nl, nr):
    """Cross products of 4.9M to 16.8M pairs (hundreds of 8,192-pair tiles): every kept pair in left-major
    order against the oracle, with the right side a multiple of the tile and not."""
    import oracle
    rng = np.random.default_rng(nl + nr)
    x = rng.integers(0, 1 << 30, nl).astype(np.int32)
    y = rng.integers(0, 1 << 30, nr).astype(np.int32)
    assert nl * nr >= 1 << 22
    pred = (out(f"(({col(0, capi.INT32)} + {col(1, capi.INT32)}) < 1073741824)", capi.BOOLEAN), [0, 1],
            capi.BOOLEAN)
    projs = [(copy_body(capi.INT32), [0], capi.INT32), (copy_body(capi.INT32), [1], capi.INT32)]
    outs, k = capi.cross_filter([_dev(x, dev)], [_dev(y, dev)], [pred], projs)
    ex, ey = oracle.cross_filter_pairs(x, y, 1 << 30)
    assert k == len(ex)
    np.testing.assert_array_equal(outs[0].cpu().numpy(), ex)
    np.testing.assert_array_equal(outs[1].cpu().numpy(), ey)


@pytest.mark.parametrize("nl,chunks", [(1, 1), (5, 1), (257, 2), (64, 3), (33, 8)])
def test_cross_whole_tiles_against_oracle(dev, nl, chunks):
    """Right sides of 1 to 8 whole 8,192-row chunks (the count, scan and write kernels over static tiles),
    with one and with odd left-row counts: every kept pair in left-major order against the oracle."""
    import oracle
    rng = np.random.default_rng(nl * 31 + chunks)
    nr = 8192 * chunks
    x = rng.integers(0, 1 << 30, nl).astype(np.int32)
    y = rng.integers(0, 1 << 30, nr).astype(np.int32)
    pred = [(out(f"(({col(0, capi.INT32)} + {col(1, capi.INT32)}) < 1073741824)", capi.BOOLEAN), [0, 1],
             capi.BOOLEAN)]
    projs = [(copy_body(capi.INT32), [0], capi.INT32), (copy_body(capi.INT32), [1], capi.INT32)]
    outs, k = capi.cross_filter([_dev(x, dev)], [_dev(y, dev)], pred, projs)
    ex, ey = oracle.cross_filter_pairs(x, y, 1 << 30)
    assert k == len(ex)
    np.testing.assert_array_equal(outs[0].cpu().numpy(), ex)
    np.testing.assert_array_equal(outs[1].cpu().numpy(), ey)


def test_cross_whole_tiles_wide_outputs_and_extreme_selectivity(dev):
    """8-byte outputs (the LDS stage at 8 bytes per pair), a projection over both sides, and predicates that
    keep every pair and none, over a right side of two whole tiles."""
    rng = np.random.default_rng(12)
    nl, nr = 37, 8192 * 2
    x = rng.integers(-(1 << 40), 1 << 40, nl).astype(np.int64)
    y = rng.random(nr)
    projs = [(copy_body(capi.INT64), [0], capi.INT64),
             (out(f"((double){col(0, capi.INT64)} * {col(1, capi.FLOAT64)})", capi.FLOAT64), [0, 1], capi.FLOAT64),
             (copy_body(capi.FLOAT64), [1], capi.FLOAT64)]
    for thr, expect in [("0.5", None), ("2.0", nl * nr), ("-1.0", 0)]:
        pred = [(out(f"({col(0, capi.FLOAT64)} < {thr})", capi.BOOLEAN), [1], capi.BOOLEAN)]
        o, k = capi.cross_filter([_dev(x, dev)], [_dev(y, dev)], pred, projs)
        keep = y < float(thr)
        assert k == nl * int(keep.sum())
        if expect is not None:
            assert k == expect
        ex = np.repeat(x, int(keep.sum()))
        ey = np.tile(y[keep], nl)
        np.testing.assert_array_equal(o[0].cpu().numpy(), ex)
        np.testing.assert_array_equal(o[2].cpu().numpy(), ey)
        np.testing.assert_array_equal(o[1].cpu().numpy(), ex.astype(np.float64) * ey)


def test_cross_whole_tiles_capacity_guard(dev):
    """Every pair kept over a whole tile, output capacity below the kept count: the kernels drop writes past
    the capacity and the call reports the overflow."""
    x = np.zeros(2, np.int32)
    y = np.zeros(8192, np.int32)
    pred = [(out(f"(({col(0, capi.INT32)} + {col(1, capi.INT32)}) < 5)", capi.BOOLEAN), [0, 1], capi.BOOLEAN)]
    with pytest.raises(capi.VgpuError, match="more than the output capacity"):
        capi.cross_filter([_dev(x, dev)], [_dev(y, dev)], pred,
                          [(copy_body(capi.INT32), [0], capi.INT32), (copy_body(capi.INT32), [1], capi.INT32)],
                          capacity=10000)


def test_cross_product_unfiltered_left_major(dev):
    x = np.array([10, 20, 30], np.int64)
    y = np.array([1.5, 2.5], np.float64)
    projs = [(copy_body(capi.INT64), [0], capi.INT64), (copy_body(capi.FLOAT64), [1], capi.FLOAT64)]
    outs, k = capi.cross_filter([_dev(x, dev)], [_dev(y, dev)], [], projs)
    assert k == 6
    np.testing.assert_array_equal(outs[0].cpu().numpy(), [10, 10, 20, 20, 30, 30])
    np.testing.assert_array_equal(outs[1].cpu().numpy(), [1.5, 2.5, 1.5, 2.5, 1.5, 2.5])


def test_cross_filter_config_shape_counts(dev):
    """4,096 x 65,536 pairs with x + y < 2^30: count and checksums against the closed form."""
    import oracle
    rng = np.random.default_rng(6)
    x = rng.integers(0, 1 << 30, 4096).astype(np.int32)
    y = rng.integers(0, 1 << 30, 65536).astype(np.int32)
    pred = (out(f"(({col(0, capi.INT32)} + {col(1, capi.INT32)}) < 1073741824)", capi.BOOLEAN), [0, 1],
            capi.BOOLEAN)
    projs = [(copy_body(capi.INT32), [0], capi.INT32), (copy_body(capi.INT32), [1], capi.INT32)]
    outs, k = capi.cross_filter([_dev(x, dev)], [_dev(y, dev)], [pred], projs)
    cnt, sx, sy = oracle.cross_filter_count(x, y, 1 << 30)
    assert k == cnt
    assert int(outs[0].to(dtype=__import__("torch").int64).sum().item()) == sx
    assert int(outs[1].to(dtype=__import__("torch").int64).sum().item()) == sy
    # left-major: the left value is non-decreasing in blocks of equal x order
    xs = outs[0][:100000].cpu().numpy()
    first = np.nonzero(np.diff(xs))[0]
    assert len(np.unique(xs[first])) == len(first)


def test_cross_filter_capacity_guard(dev):
    x = np.zeros(10, np.int32)
    y = np.zeros(10, np.int32)
    pred = (out(f"(({col(0, capi.INT32)} + {col(1, capi.INT32)}) < 5)", capi.BOOLEAN), [0, 1], capi.BOOLEAN)
    with pytest.raises(capi.VgpuError, match="more than the output capacity"):
        capi.cross_filter([_dev(x, dev)], [_dev(y, dev)], [pred], [(copy_body(capi.INT32), [0], capi.INT32)],
                          capacity=50)
