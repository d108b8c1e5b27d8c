"""Writes reference_groupby.json: the GROUP BY known-answer cases of the reference's executor tests
(cuda_engine/tests/test_gpu_executor.cpp:455-770), restated as data.

The reference generates its inputs with closed-form row formulas and checks against a std::map brute
force; this script applies the same formulas and a plain-Python brute force, and stores inputs and
expected per-group (count, sum) so the fixture is self-contained. Run: python tests/golden/make_reference_cases.py
"""
import json
from pathlib import Path

CASES = [
    # TestGroupBySumOnGpu (:455-492): k = i % 1000, v = i % 64, n = 100000 (shrunk to 5000 rows here)
    ("sum_by_mod1000", 5000, lambda i: i % 1000, lambda i: i % 64),
    # TestGroupByCountStarOnGpu (:494-530): k = (i * i) % 251, n = 60000 (uneven group sizes)
    ("count_by_square_mod251", 6000, lambda i: (i * i) % 251, lambda i: 1),
    # TestGroupByMinMaxOnGpu (:532-576): k = i % 97, v = (i * 7919) % 100000 - 50000, n = 50000
    ("minmax_by_mod97", 5000, lambda i: i % 97, lambda i: (i * 7919) % 100000 - 50000),
    # TestGroupByHighCardinalityOnGpu (:733-770): every row its own group (k = n - i), n = 20000
    ("high_cardinality", 3000, lambda i: 3000 - i, lambda i: i),
]


def main():
    out = []
    for name, n, kf, vf in CASES:
        key = [kf(i) for i in range(n)]
        val = [vf(i) for i in range(n)]
        exp = {}
        for k, v in zip(key, val):
            c, s = exp.get(k, (0, 0))
            exp[k] = (c + 1, s + v)
        out.append({"name": name, "key": key, "value": val,
                    "expected": {str(k): list(cs) for k, cs in sorted(exp.items())}})
    path = Path(__file__).resolve().parent / "reference_groupby.json"
    path.write_text(json.dumps(out, separators=(",", ":")))
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
