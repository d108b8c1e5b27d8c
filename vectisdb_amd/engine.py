"""Python mirror of the engine API in include/gpu_engine.hpp (cuda_engine/include/gpu_engine.hpp).

Same type names, fields, enum values and functions as the C++ header, bound through the C ABI
(vgpu_execute_plan and the routing entry points of include/vgpu.h): a plan built here is serialized to
JSON, executed by GpuEngine::ExecutePlan inside libvgpu.so, and the host result columns come back as numpy
arrays. There is no Python or CPU execution path: a missing library raises.
"""
from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass, field
from enum import IntEnum
from typing import List, Optional

import numpy as np

from . import capi


class GpuOpType(IntEnum):
    SCAN = 0
    FILTER = 1
    PROJECTION = 2
    HASH_JOIN = 3
    GROUP_BY_AGGREGATE = 4
    CROSS_PRODUCT = 5


class ExecutionMode(IntEnum):
    PIPELINE_STREAM_CACHED = 0
    PIPELINE_STREAM_UNCACHED = 1
    BATCH_DISPATCH = 2


class GpuValueType(IntEnum):
    INT16 = 0
    INT32 = 1
    INT64 = 2
    FLOAT32 = 3
    FLOAT64 = 4
    BOOLEAN = 5
    DICTIONARY_STRING = 6
    HUGEINT = 7


class GpuWindowPrimitive(IntEnum):
    NONE = 0
    RUNNING_SUM = 1


class GpuAggregateKind(IntEnum):
    COUNT_STAR = 0
    COUNT = 1
    SUM = 2
    MIN = 3
    MAX = 4
    AVG = 5


class GpuWorkDensity(IntEnum):
    INSUFFICIENT = 0
    INCONCLUSIVE = 1
    SUFFICIENT = 2


kMinOpsPerScannedRow = 3.0
kOffloadOpsPerRow = 10.0

#: DuckDB's hugeint_t layout: low 64 bits unsigned, high 64 bits signed.
HUGEINT_DTYPE = np.dtype([("lower", "<u8"), ("upper", "<i8")])

NUMPY_DTYPE = {
    GpuValueType.INT16: np.dtype(np.int16),
    GpuValueType.INT32: np.dtype(np.int32),
    GpuValueType.INT64: np.dtype(np.int64),
    GpuValueType.FLOAT32: np.dtype(np.float32),
    GpuValueType.FLOAT64: np.dtype(np.float64),
    GpuValueType.BOOLEAN: np.dtype(np.bool_),
    GpuValueType.HUGEINT: HUGEINT_DTYPE,
}


def hugeint_to_int(values) -> List[int]:
    """Python ints from a HUGEINT column."""
    return [(int(v["upper"]) << 64) | int(v["lower"]) for v in values]


@dataclass
class GpuExpr:
    generated_cuda_source: str = ""
    input_columns: List[str] = field(default_factory=list)
    output_type: GpuValueType = GpuValueType.FLOAT64
    source_hash: str = ""
    window_primitive: GpuWindowPrimitive = GpuWindowPrimitive.NONE
    estimated_ops_per_row: float = 0.0

    def to_json(self):
        d = {"source": self.generated_cuda_source, "inputs": list(self.input_columns),
             "output_type": int(self.output_type), "ops": float(self.estimated_ops_per_row)}
        if self.source_hash:
            d["hash"] = self.source_hash
        if self.window_primitive == GpuWindowPrimitive.RUNNING_SUM:
            d["window"] = "RUNNING_SUM"
        return d


@dataclass
class GpuAggregate:
    kind: GpuAggregateKind = GpuAggregateKind.COUNT_STAR
    input_column: str = ""
    output_type: GpuValueType = GpuValueType.INT64

    def to_json(self):
        return {"kind": int(self.kind), "input": self.input_column, "output_type": int(self.output_type)}


@dataclass
class GpuTableRef:
    catalog: str = ""
    schema: str = ""
    table: str = ""

    def to_string(self) -> str:
        return f"{self.catalog}.{self.schema}.{self.table}"


@dataclass
class GpuPlanNode:
    op_type: GpuOpType = GpuOpType.SCAN
    execution_mode: ExecutionMode = ExecutionMode.PIPELINE_STREAM_CACHED
    children: List["GpuPlanNode"] = field(default_factory=list)
    estimated_rows: int = 0
    estimated_row_width_bytes: int = 0
    table: GpuTableRef = field(default_factory=GpuTableRef)
    column_names: List[str] = field(default_factory=list)
    expressions: List[GpuExpr] = field(default_factory=list)
    build_keys: List[str] = field(default_factory=list)
    probe_keys: List[str] = field(default_factory=list)
    output_columns: List[str] = field(default_factory=list)
    group_keys: List[str] = field(default_factory=list)
    aggregate_exprs: List[str] = field(default_factory=list)
    aggregates: List[GpuAggregate] = field(default_factory=list)

    def to_dict(self):
        return {
            "op": int(self.op_type), "mode": int(self.execution_mode),
            "estimated_rows": int(self.estimated_rows),
            "estimated_row_width_bytes": int(self.estimated_row_width_bytes),
            "table": {"catalog": self.table.catalog, "schema": self.table.schema, "table": self.table.table},
            "column_names": list(self.column_names),
            "expressions": [e.to_json() for e in self.expressions],
            "build_keys": list(self.build_keys), "probe_keys": list(self.probe_keys),
            "output_columns": list(self.output_columns), "group_keys": list(self.group_keys),
            "aggregate_exprs": list(self.aggregate_exprs),
            "aggregates": [a.to_json() for a in self.aggregates],
            "children": [c.to_dict() for c in self.children],
        }

    def to_json(self) -> str:
        return json.dumps(self.to_dict())


@dataclass
class GpuColumn:
    """Host column. `data`: contiguous numpy array of the type's dtype; `validity`: Arrow bitmap (uint8,
    bit r%8 of byte r/8, 1 = valid) or None when the column has no NULLs."""
    name: str
    type: GpuValueType
    data: Optional[np.ndarray] = None
    validity: Optional[np.ndarray] = None
    row_count: int = 0

    def valid_mask(self) -> np.ndarray:
        if self.validity is None:
            return np.ones(self.row_count, dtype=bool)
        return np.unpackbits(self.validity, bitorder="little")[: self.row_count].astype(bool)


def make_column(name: str, vtype: GpuValueType, values, validity=None) -> GpuColumn:
    """GpuColumn over `values` (converted to the type's dtype), with an optional bool NULL mask
    (True = valid) packed into the Arrow bitmap."""
    arr = np.ascontiguousarray(np.asarray(values, dtype=NUMPY_DTYPE[GpuValueType(vtype)]))
    packed = None
    if validity is not None:
        packed = np.packbits(np.asarray(validity, dtype=bool), bitorder="little")
    return GpuColumn(name, GpuValueType(vtype), arr, packed, int(arr.shape[0]))


@dataclass
class GpuExecutionResult:
    success: bool = False
    error_message: str = ""
    columns: List[GpuColumn] = field(default_factory=list)

    def column(self, name: str) -> GpuColumn:
        for c in self.columns:
            if c.name == name:
                return c
        raise KeyError(name)


@dataclass
class GpuExecuteOptions:
    stage_scan_to_cache: bool = False
    cache_chunk_index: int = -1
    serve_whole_scan_from_cache: bool = False


class GpuEngine:
    _instance = None

    @classmethod
    def instance(cls) -> "GpuEngine":
        if cls._instance is None:
            cls._instance = GpuEngine()
        return cls._instance

    def execute_plan(self, plan: GpuPlanNode, inputs: List[GpuColumn],
                     options: Optional[GpuExecuteOptions] = None) -> GpuExecutionResult:
        lib = capi.load()
        n = len(inputs)
        hosts = (capi.HostColumn * max(n, 1))()
        keep = []
        for i, col in enumerate(inputs):
            data = None if col.data is None else np.ascontiguousarray(col.data)
            validity = None if col.validity is None else np.ascontiguousarray(col.validity, dtype=np.uint8)
            keep.extend([data, validity])
            name = col.name.encode()
            keep.append(name)
            hosts[i].name = name
            hosts[i].type = int(col.type)
            hosts[i].data = None if data is None or data.size == 0 else data.ctypes.data
            hosts[i].validity = None if validity is None else validity.ctypes.data
            hosts[i].rows = int(col.row_count)
        opts = None
        if options is not None:
            opts = capi.ExecuteOptions(int(options.stage_scan_to_cache), int(options.cache_chunk_index),
                                       int(options.serve_whole_scan_from_cache))
        handle = C.c_void_p()
        capi.check(lib.vgpu_execute_plan(plan.to_json().encode(), hosts, n,
                                         C.byref(opts) if opts is not None else None, C.byref(handle)))
        try:
            result = GpuExecutionResult(bool(lib.vgpu_result_success(handle)),
                                        (lib.vgpu_result_error(handle) or b"").decode())
            for i in range(lib.vgpu_result_num_columns(handle)):
                hc = capi.HostColumn()
                capi.check(lib.vgpu_result_column(handle, i, C.byref(hc)))
                vtype = GpuValueType(hc.type)
                dtype = NUMPY_DTYPE[vtype]
                rows = int(hc.rows)
                data = np.empty(rows, dtype=dtype)
                if rows > 0:
                    C.memmove(data.ctypes.data, hc.data, rows * dtype.itemsize)
                validity = None
                if hc.validity:
                    validity = np.empty((rows + 7) // 8, dtype=np.uint8)
                    C.memmove(validity.ctypes.data, hc.validity, validity.size)
                result.columns.append(GpuColumn(hc.name.decode(), vtype, data, validity, rows))
            return result
        finally:
            lib.vgpu_result_free(handle)


# ---------------------------------------------------------------- routing (no device needed)

@dataclass
class GpuPlanShape:
    scanned_rows: int = 0
    scanned_bytes: int = 0
    max_rows: int = 0
    total_ops: float = 0.0
    ops_per_scanned_row: float = 0.0
    ops_per_scanned_byte: float = 0.0
    density: GpuWorkDensity = GpuWorkDensity.INSUFFICIENT
    amplification: float = 0.0


@dataclass
class GpuWorkingSetEstimate:
    total_bytes: int = 0
    chunk_capped: bool = False
    largest_operator: str = ""
    largest_bytes: int = 0


def analyze_plan_shape(plan: GpuPlanNode) -> GpuPlanShape:
    s = capi.PlanShape()
    capi.check(capi.load().vgpu_analyze_plan_shape(plan.to_json().encode(), C.byref(s)))
    return GpuPlanShape(s.scanned_rows, s.scanned_bytes, s.max_rows, s.total_ops, s.ops_per_scanned_row,
                        s.ops_per_scanned_byte, GpuWorkDensity(s.density), s.amplification)


def is_row_independent(plan: GpuPlanNode) -> bool:
    out = C.c_int32()
    capi.check(capi.load().vgpu_is_row_independent(plan.to_json().encode(), C.byref(out)))
    return bool(out.value)


def estimate_working_set(plan: GpuPlanNode, max_chunk_rows: int) -> GpuWorkingSetEstimate:
    total, largest = C.c_uint64(), C.c_uint64()
    capped = C.c_int32()
    name = C.create_string_buffer(64)
    capi.check(capi.load().vgpu_estimate_working_set(plan.to_json().encode(), max_chunk_rows, C.byref(total),
                                                     C.byref(capped), C.byref(largest), name, 64))
    return GpuWorkingSetEstimate(total.value, bool(capped.value), name.value.decode(), largest.value)


def estimate_columns_size(plan: GpuPlanNode) -> int:
    out = C.c_uint64()
    capi.check(capi.load().vgpu_estimate_columns_size(plan.to_json().encode(), C.byref(out)))
    return out.value


def determine_execution_mode(plan: GpuPlanNode, vram_budget: int, threshold: float = 0.75) -> ExecutionMode:
    out = C.c_int32()
    capi.check(capi.load().vgpu_determine_execution_mode(plan.to_json().encode(), vram_budget, threshold,
                                                         C.byref(out)))
    return ExecutionMode(out.value)


def hash_source(text: str) -> str:
    """KernelCache::HashSource: lowercase hex SHA-256."""
    raw = text.encode()
    buf = C.create_string_buffer(65)
    capi.check(capi.load().vgpu_hash_source(raw, len(raw), buf))
    return buf.value.decode()


def set_gpu_fast_math_enabled(on: bool) -> None:
    capi.load().vgpu_set_fast_math(1 if on else 0)


def is_gpu_fast_math_enabled() -> bool:
    return bool(capi.load().vgpu_fast_math_enabled())


def set_gpu_extended_semantics_enabled(on: bool) -> None:
    """SetGpuExtendedSemanticsEnabled (include/gpu_semantics.hpp): DuckDB's NULL, floating-key and composite
    join-key semantics instead of the reference operators' refusals."""
    capi.load().vgpu_set_extended_semantics(1 if on else 0)


def is_gpu_extended_semantics_enabled() -> bool:
    return bool(capi.load().vgpu_extended_semantics_enabled())


# ---------------------------------------------------------------- column cache / device budget

def _t(ref: GpuTableRef):
    return ref.catalog.encode(), ref.schema.encode(), ref.table.encode()


class GpuColumnCache:
    """The process's GpuColumnCache singleton, through the C ABI."""

    @staticmethod
    def begin_staging(ref: GpuTableRef):
        capi.check(capi.load().vgpu_cache_begin_staging(*_t(ref)))

    @staticmethod
    def commit_staging(ref: GpuTableRef, scanned_rows: int):
        capi.check(capi.load().vgpu_cache_commit_staging(*_t(ref), scanned_rows))

    @staticmethod
    def abort_staging(ref: GpuTableRef):
        capi.check(capi.load().vgpu_cache_abort_staging(*_t(ref)))

    @staticmethod
    def invalidate_table(ref: GpuTableRef):
        capi.check(capi.load().vgpu_cache_invalidate_table(*_t(ref)))

    @staticmethod
    def invalidate_all():
        capi.check(capi.load().vgpu_cache_invalidate_all())

    @staticmethod
    def current_bytes() -> int:
        out = C.c_uint64()
        capi.check(capi.load().vgpu_cache_current_bytes(C.byref(out)))
        return out.value

    @staticmethod
    def table_generation(ref: GpuTableRef) -> int:
        out = C.c_uint64()
        capi.check(capi.load().vgpu_cache_table_generation(*_t(ref), C.byref(out)))
        return out.value

    @staticmethod
    def replayable_chunk_rows(ref: GpuTableRef, columns: List[str]) -> List[int]:
        names = (C.c_char_p * max(len(columns), 1))(*[c.encode() for c in columns])
        rows = (C.c_uint64 * 4096)()
        n = C.c_size_t()
        capi.check(capi.load().vgpu_cache_replayable_chunk_rows(*_t(ref), names, len(columns), rows, 4096,
                                                                C.byref(n)))
        return [rows[i] for i in range(n.value)]
