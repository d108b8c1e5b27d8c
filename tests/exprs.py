"""Expression bodies in the exact text form the DuckDB extension emits (extension/src/expression_translator.cpp:
column reads `((const T*)inputs[k])[row]`, output `((U*)out)[idx] = ...;`, casts `((T)(...))`, BIGINT
literals with `LL`, doubles printed with 17 significant digits)."""
from vectisdb_amd import capi

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


def cfg0_exprs():
    """(preds, projs) over columns [a INT32, b INT64, c DOUBLE]."""
    preds = [(CFG0_PRED_A, [0], capi.BOOLEAN), (CFG0_PRED_C, [2], capi.BOOLEAN)]
    projs = [(CFG0_PROJ_1, [0, 1], capi.INT64), (CFG0_PROJ_2, [2, 0], capi.FLOAT64)]
    return preds, projs


def copy_body(vtype):
    return out(col(0, vtype), vtype)
