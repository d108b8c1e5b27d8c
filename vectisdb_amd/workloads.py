"""Synthetic workloads of the engine's measured configurations, generated on the device, and the expression
bodies they run.

Expression bodies use the exact text form the DuckDB extension emits (extension/src/expression_translator.cpp):
column reads `((const T*)inputs[k])[row]`, output `((U*)out)[idx] = ...;`, casts `((T)(...))`, BIGINT literals
with `LL`, doubles printed with 17 significant digits.

Tables (every column seeded, made by vgpu_generate_column on the device, so the same call gives the same data
on any MI355X):
  * filter table: a INT32 uniform [0, 1e6), b INT64 uniform [-2^40, 2^40), c DOUBLE uniform [0, 1);
  * GROUP BY table: key INT64 Zipf(1.1) over 1M values through a seeded permutation, v1 INT64 uniform
    [-1e6, 1e6], v2 DOUBLE normal(0, 1);
  * join: build 2^24 unique keys (a seeded permutation of [0, 2^24)) with payload INT32 uniform [0, 1000),
    probe 2^28 foreign keys, 90% uniform over the build keys and 10% uniform over [2^24, 2^25);
  * cross: 4,096 x 65,536 INT32 uniform [0, 2^30) with the filter x + y < 2^30;
  * batched: 512 row groups of log-uniform size in [2K, 4M] rows with per-group Zipf keys;
  * served: ~40 GB of the tables above, and query streams over row ranges of them.
"""
from __future__ import annotations

from . import capi

CTYPE = {capi.INT16: "short", capi.INT32: "int", capi.INT64: "long long", capi.FLOAT32: "float",
         capi.FLOAT64: "double", capi.BOOLEAN: "bool"}


def col(k, vtype):
    return f"((const {CTYPE[vtype]}*)inputs[{k}])[row]"


def cast(expr, vtype):
    return f"(({CTYPE[vtype]})({expr}))"


def out(expr, vtype):
    return f"(({CTYPE[vtype]}*)out)[idx] = {expr};"


def dlit(v: float) -> str:
    text = repr(float(v)) if abs(v) < 1e16 else f"{v:.17g}"
    return text if any(ch in text for ch in ".eE") else text + ".0"


# The two config-0 projections and predicates (SELECT a*2+b, c*1.5+a FROM t WHERE a < 500000 AND c > 0.25),
# with DuckDB's binding: a*2 is INTEGER, widened to BIGINT for + b; a is cast to DOUBLE for c*1.5 + a.
CFG0_PRED_A = out(f"({col(0, capi.INT32)} < 500000)", capi.BOOLEAN)
CFG0_PRED_C = out(f"({col(0, capi.FLOAT64)} > 0.25)", capi.BOOLEAN)
CFG0_PROJ_1 = out(f"({cast(f'({col(0, capi.INT32)} * 2)', capi.INT64)} + {col(1, capi.INT64)})", capi.INT64)
CFG0_PROJ_2 = out(f"(({col(0, capi.FLOAT64)} * 1.5) + {cast(col(1, capi.INT32), capi.FLOAT64)})", capi.FLOAT64)

CROSS_LIMIT = 1 << 30
CROSS_PRED = out(f"(({col(0, capi.INT32)} + {col(1, capi.INT32)}) < 1073741824)", capi.BOOLEAN)

# The 7-aggregate GROUP BY list (key is the group key): COUNT(*), SUM(v1), MIN(v1), MAX(v1), SUM(v2), AVG(v2).
AGGS7 = [(capi.COUNT_STAR, -1, capi.INT64), (capi.SUM, 0, capi.HUGEINT), (capi.MIN, 0, capi.INT64),
         (capi.MAX, 0, capi.INT64), (capi.SUM, 1, capi.FLOAT64), (capi.AVG, 1, capi.FLOAT64)]


def cfg0_exprs():
    """(preds, projs) over columns [a INT32, b INT64, c DOUBLE]."""
    preds = [(CFG0_PRED_A, [0], capi.BOOLEAN), (CFG0_PRED_C, [2], capi.BOOLEAN)]
    projs = [(CFG0_PROJ_1, [0, 1], capi.INT64), (CFG0_PROJ_2, [2, 0], capi.FLOAT64)]
    return preds, projs


def copy_body(vtype):
    return out(col(0, vtype), vtype)


def cross_exprs():
    """(preds, projs) of SELECT x, y FROM L, R WHERE x + y < 2^30 (left column 0, right column 1)."""
    return ([(CROSS_PRED, [0, 1], capi.BOOLEAN)],
            [(copy_body(capi.INT32), [0], capi.INT32), (copy_body(capi.INT32), [1], capi.INT32)])


def zipf_tables(torch, dev, distinct, s, seed):
    """Inverse-CDF table of Zipf(s) over ranks 1..distinct and a seeded permutation of [0, distinct)."""
    import numpy as np
    ranks = np.arange(1, distinct + 1, dtype=np.float64)
    w = ranks ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    cdf[-1] = 1.0 + 1e-12  # every u in [0,1) lands inside the table
    perm = np.random.RandomState(seed).permutation(distinct).astype(np.int64)
    return torch.from_numpy(cdf).to(dev), torch.from_numpy(perm).to(dev)


def make_filter_table(torch, dev, rows, seed):
    a = torch.empty(rows, dtype=torch.int32, device=dev)
    b = torch.empty(rows, dtype=torch.int64, device=dev)
    x = torch.empty(rows, dtype=torch.float64, device=dev)
    capi.generate_column(a, capi.INT32, 0, lo=0, hi=999_999, seed=seed, stream_id=1)
    capi.generate_column(b, capi.INT64, 0, lo=-(1 << 40), hi=(1 << 40) - 1, seed=seed, stream_id=2)
    capi.generate_column(x, capi.FLOAT64, 1, seed=seed, stream_id=3)
    return a, b, x


def make_groupby_table(torch, dev, rows, seed=1):
    cdf, perm = zipf_tables(torch, dev, 1_000_000, 1.1, seed)
    key = torch.empty(rows, dtype=torch.int64, device=dev)
    v1 = torch.empty(rows, dtype=torch.int64, device=dev)
    v2 = torch.empty(rows, dtype=torch.float64, device=dev)
    capi.generate_column(key, capi.INT64, 3, seed=seed, stream_id=1, cdf=cdf, perm=perm)
    capi.generate_column(v1, capi.INT64, 0, lo=-1_000_000, hi=1_000_000, seed=seed, stream_id=2)
    capi.generate_column(v2, capi.FLOAT64, 2, seed=seed, stream_id=3)
    torch.cuda.synchronize()
    return key, v1, v2


# fk = sel == 0 ? miss : hit  (sel uniform over 0..9: 10% misses)
_PICK_FK = ("((long long*)out)[idx] = (((const long long*)inputs[2])[row] == 0LL) ? "
            "((const long long*)inputs[1])[row] : ((const long long*)inputs[0])[row];")


def make_join_tables(torch, dev, nb=1 << 24, nprobe=1 << 28, seed=21):
    """(build keys, build payload, probe fk): build keys a seeded permutation of [0, nb), payload INT32 uniform
    [0, 1000); fk 90% uniform over [0, nb) and 10% uniform over [nb, 2 nb)."""
    import numpy as np
    build = torch.from_numpy(np.random.RandomState(seed).permutation(nb).astype(np.int64)).to(dev)
    payload = torch.empty(nb, dtype=torch.int32, device=dev)
    capi.generate_column(payload, capi.INT32, 0, lo=0, hi=999, seed=seed, stream_id=1)
    fk = torch.empty(nprobe, dtype=torch.int64, device=dev)
    chunk = min(nprobe, 1 << 27)
    hit = torch.empty(chunk, dtype=torch.int64, device=dev)
    miss = torch.empty(chunk, dtype=torch.int64, device=dev)
    sel = torch.empty(chunk, dtype=torch.int64, device=dev)
    for i, lo in enumerate(range(0, nprobe, chunk)):
        m = min(chunk, nprobe - lo)
        capi.generate_column(hit[:m], capi.INT64, 0, lo=0, hi=nb - 1, seed=seed + i, stream_id=2)
        capi.generate_column(miss[:m], capi.INT64, 0, lo=nb, hi=2 * nb - 1, seed=seed + i, stream_id=3)
        capi.generate_column(sel[:m], capi.INT64, 0, lo=0, hi=9, seed=seed + i, stream_id=4)
        outs, _, _, _ = capi.filter_project([hit[:m], miss[:m], sel[:m]], [], [(_PICK_FK, [0, 1, 2], capi.INT64)])
        capi.device_copy(fk[lo:lo + m], outs[0], m * 8)
        del outs
    del hit, miss, sel
    torch.cuda.synchronize()
    return build, payload, fk


def make_cross_tables(torch, dev, nl=4096, nr=65536, seed=11):
    x = torch.empty(nl, dtype=torch.int32, device=dev)
    y = torch.empty(nr, dtype=torch.int32, device=dev)
    capi.generate_column(x, capi.INT32, 0, lo=0, hi=CROSS_LIMIT - 1, seed=seed, stream_id=1)
    capi.generate_column(y, capi.INT32, 0, lo=0, hi=CROSS_LIMIT - 1, seed=seed, stream_id=2)
    torch.cuda.synchronize()
    return x, y


# WHERE a < 500000 AND c > 0.25 over the batched table's columns [a, c, key, v1, v2]
BATCHED_PREDS = [(out(f"({col(0, capi.INT32)}) < 500000", capi.BOOLEAN), [0], capi.BOOLEAN),
                 (out(f"({col(0, capi.FLOAT64)}) > 0.25", capi.BOOLEAN), [1], capi.BOOLEAN)]


def make_batched_table(torch, dev, nq=512, seed=7):
    """nq row groups of one table, sizes log-uniform in [2K, 4M] rows; a INT32 uniform [0,1e6), c DOUBLE
    uniform [0,1), v1 INT64 uniform [-1e6,1e6], v2 DOUBLE normal(0,1) as in the plain configs; the group
    count of query g is min(rows, log-uniform in [16, 1M]) and its keys are Zipf(1.1) ranks over that count
    mapped through a seeded permutation of [0, count) (inverse CDF on the device).
    Returns ([a, c, key, v1, v2], device int64 offsets (nq + 1), sizes, group counts)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(2048), np.log(4 << 20), nq)).astype(np.int64)
    cards = np.minimum(sizes, np.exp(rng.uniform(np.log(16), np.log(1_000_000), nq)).astype(np.int64))
    rows = int(sizes.sum())
    a = torch.empty(rows, dtype=torch.int32, device=dev)
    c = torch.empty(rows, dtype=torch.float64, device=dev)
    capi.generate_column(a, capi.INT32, 0, lo=0, hi=999_999, seed=seed, stream_id=1)
    capi.generate_column(c, capi.FLOAT64, 1, seed=seed, stream_id=2)
    _, v1, v2 = make_groupby_table(torch, dev, rows, seed=seed)
    key = torch.empty(rows, dtype=torch.int64, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    for q in range(nq):
        g = int(cards[q])
        w = torch.arange(1, g + 1, dtype=torch.float64, device=dev).pow_(-1.1)
        cdf = torch.cumsum(w, 0)
        cdf /= cdf[-1].clone()
        u = torch.rand(int(sizes[q]), dtype=torch.float64, device=dev, generator=gen)
        rank = torch.searchsorted(cdf, u).clamp_(max=g - 1)
        perm = torch.randperm(g, device=dev, generator=gen)
        key[offsets[q]:offsets[q + 1]] = perm[rank]
    offsets_t = torch.from_numpy(offsets).to(dev)
    torch.cuda.synchronize()
    return [a, c, key, v1, v2], offsets_t, sizes, cards


def make_served_tables(torch, dev, scale=1.0, seed=31):
    """Device-resident tables of the served workload (~40 GB at scale 1): the filter table, the GROUP BY table,
    the probe table (fk as in make_join_tables over 1.6G rows at scale 1) with the 2^24-row build side, and the
    4,096 / 65,536-row cross sides."""
    nf = int(600_000_000 * scale)
    ng = int(600_000_000 * scale)
    nprobe = int(1_600_000_000 * scale)
    a, b, c = make_filter_table(torch, dev, nf, seed)
    key, v1, v2 = make_groupby_table(torch, dev, ng, seed=seed)
    build, payload, fk = make_join_tables(torch, dev, 1 << 24, nprobe, seed)
    x, y = make_cross_tables(torch, dev, seed=seed)
    torch.cuda.synchronize()
    tables = {"filter": [a, b, c], "group": [key, v1, v2], "probe": fk, "build": build, "payload": payload,
              "left": [x], "right": [y]}
    total = sum(t.numel() * t.element_size() for t in [a, b, c, key, v1, v2, fk, build, payload, x, y])
    return tables, total


def served_queries(n, rows, seed, lo=10_000, hi=50_000_000, left_rows=4096):
    """n queries drawn uniformly from the four templates (filter/project, GROUP BY, join + aggregate, cross
    product + filter), input sizes log-uniform in [lo, hi] rows, each over a uniformly placed row range of its
    template's table. A cross query's input size counts pairs: it takes size / 4,096 rows of the right side
    against the whole 4,096-row left side."""
    import numpy as np
    rng = np.random.default_rng(seed)
    kinds = rng.integers(0, 4, n)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.uint64)
    sizes = np.where(kinds == 3, np.maximum(sizes // left_rows, 1), sizes).astype(np.uint64)
    starts = np.array([rng.integers(0, rows[k] - int(s) + 1) for k, s in zip(kinds, sizes)], dtype=np.uint64)
    return kinds, starts, sizes
