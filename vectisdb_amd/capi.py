"""ctypes binding of libvgpu.so (include/vgpu.h).

This is the same stub a maintainer would add on the reference side to call the engine from Python; the
tests and bench.py drive the HIP path only through it. Loading fails loudly: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libvgpu.so"

INT16, INT32, INT64, FLOAT32, FLOAT64, BOOLEAN, DICTIONARY_STRING, HUGEINT = range(8)
COUNT_STAR, COUNT, SUM, MIN, MAX, AVG = range(6)
TYPE_WIDTH = {INT16: 2, INT32: 4, INT64: 8, FLOAT32: 4, FLOAT64: 8, BOOLEAN: 1, HUGEINT: 16}


class VgpuError(RuntimeError):
    pass


class Column(C.Structure):
    _fields_ = [("data", C.c_void_p), ("type", C.c_int32), ("rows", C.c_uint64)]


class Agg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("input", C.c_int32), ("output_type", C.c_int32)]


class JoinAgg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("side", C.c_int32), ("input", C.c_int32), ("output_type", C.c_int32)]


class Expr(C.Structure):
    _fields_ = [("body", C.c_char_p), ("inputs", C.POINTER(C.c_int32)), ("ninputs", C.c_int32),
                ("output_type", C.c_int32)]


class HostColumn(C.Structure):
    _fields_ = [("name", C.c_char_p), ("type", C.c_int32), ("data", C.c_void_p), ("validity", C.c_void_p),
                ("rows", C.c_uint64)]


class ExecuteOptions(C.Structure):
    _fields_ = [("stage_scan_to_cache", C.c_int32), ("cache_chunk_index", C.c_int64),
                ("serve_whole_scan_from_cache", C.c_int32)]


class PlanShape(C.Structure):
    _fields_ = [("scanned_rows", C.c_uint64), ("scanned_bytes", C.c_uint64), ("max_rows", C.c_uint64),
                ("total_ops", C.c_double), ("ops_per_scanned_row", C.c_double),
                ("ops_per_scanned_byte", C.c_double), ("density", C.c_int32), ("amplification", C.c_double)]


# (name, restype, argtypes); every function returns int status unless noted
class ServerOptions(C.Structure):
    _fields_ = [("slots", C.c_int32), ("batching", C.c_int32), ("max_batch", C.c_int32),
                ("memory_budget", C.c_uint64), ("fast_math", C.c_int32), ("max_batch_rows", C.c_uint64),
                ("resident_join", C.c_int32)]


class Template(C.Structure):
    _fields_ = [("kind", C.c_int32), ("cols", C.POINTER(Column)), ("ncols", C.c_int32),
                ("preds", C.POINTER(Expr)), ("npreds", C.c_int32), ("projs", C.POINTER(Expr)), ("nprojs", C.c_int32),
                ("key_col", C.c_int32), ("value_cols", C.POINTER(C.c_int32)), ("nvalues", C.c_int32),
                ("aggs", C.POINTER(Agg)), ("naggs", C.c_int32), ("build_key", Column), ("build_payload", Column),
                ("left_cols", C.POINTER(Column)), ("nleft", C.c_int32)]


class ServedInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("kind", C.c_int32), ("rows", C.c_uint64), ("ncols", C.c_int32),
                ("on_device", C.c_int32), ("batch_size", C.c_uint32), ("submit_ns", C.c_uint64),
                ("start_ns", C.c_uint64), ("done_ns", C.c_uint64)]


class ServerStats(C.Structure):
    _fields_ = [("queries", C.c_uint64), ("batches", C.c_uint64), ("max_batch", C.c_uint64),
                ("peak_reserved", C.c_uint64), ("budget", C.c_uint64), ("rejected", C.c_uint64)]


TEMPLATE_FILTER_PROJECT, TEMPLATE_GROUP_BY, TEMPLATE_JOIN_AGG, TEMPLATE_CROSS = 0, 1, 2, 3

_PROTOS = {
    "vgpu_sort_autotune": (C.c_int, [C.c_int32]),
    "vgpu_sort_autotune_report": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vgpu_group_by_fused_last_pass": (C.c_int, [C.c_int32]),
    "vgpu_server_create": (C.c_int, [C.POINTER(ServerOptions), C.POINTER(C.c_void_p)]),
    "vgpu_server_destroy": (C.c_int, [C.c_void_p]),
    "vgpu_server_add_template": (C.c_int, [C.c_void_p, C.POINTER(Template), C.POINTER(C.c_int32)]),
    "vgpu_server_submit": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]),
    "vgpu_server_wait": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "vgpu_server_stats_read": (C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    "vgpu_served_result_info": (C.c_int, [C.c_void_p, C.POINTER(ServedInfo)]),
    "vgpu_served_result_column": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_int32),
                                            C.POINTER(C.c_int32)]),
    "vgpu_served_result_copy": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "vgpu_served_result_error": (C.c_char_p, [C.c_void_p]),
    "vgpu_served_result_free": (None, [C.c_void_p]),
    "vgpu_server_run_open_loop": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.c_uint64,
                                            C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "vgpu_last_error": (C.c_char_p, []),
    "vgpu_abi_version": (C.c_int, []),
    "vgpu_device_init": (C.c_int, [C.c_int]),
    "vgpu_group_by_bound": (C.c_int, [C.POINTER(Column), C.c_int32, C.c_uint64, C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_group_by": (C.c_int, [C.POINTER(Column), C.c_int32, C.POINTER(Column), C.c_int32, C.POINTER(Agg), C.c_int32,
                                C.c_uint64, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_uint64,
                                C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_sort_column": (C.c_int, [C.POINTER(Column), C.c_void_p, C.c_void_p, C.c_void_p]),
    "vgpu_filter_gt_i32": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_filter_gt_i32_noalloc": (C.c_int, [C.c_void_p, C.c_uint64, C.c_int32, C.c_void_p, C.POINTER(C.c_uint64),
                                             C.c_void_p]),
    "vgpu_filter_gt_i32_noalloc_max_rows": (C.c_uint64, []),
    "vgpu_gather": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32, C.c_uint64, C.c_void_p, C.c_void_p]),
    "vgpu_generate_column": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint64, C.c_int32, C.c_int64, C.c_int64, C.c_uint64,
                                       C.c_uint64, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]),
    "vgpu_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "vgpu_filter_project": (C.c_int, [C.POINTER(Column), C.c_void_p, C.c_int32, C.POINTER(Expr), C.c_int32,
                                      C.POINTER(Expr), C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_cross_filter": (C.c_int, [C.POINTER(Column), C.c_int32, C.c_int32, C.c_uint64, C.c_uint64,
                                    C.POINTER(Expr), C.c_int32, C.POINTER(Expr), C.c_int32, C.c_void_p, C.c_uint64,
                                    C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_join_aggregate": (C.c_int, [C.POINTER(Column), C.POINTER(Column), C.c_int32, C.POINTER(Column),
                                      C.POINTER(Column), C.c_int32, C.POINTER(JoinAgg), C.c_int32, C.c_void_p,
                                      C.c_void_p]),
    "vgpu_join_pairs": (C.c_int, [C.POINTER(Column), C.POINTER(Column), C.c_void_p, C.c_void_p, C.c_uint64,
                                  C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_profile_enable": (C.c_int, [C.c_int32]),
    "vgpu_profile_only": (C.c_int, [C.c_char_p]),
    "vgpu_profile_reset": (C.c_int, []),
    "vgpu_profile_read": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "vgpu_pool_stats": (C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "vgpu_pool_trim": (C.c_int, []),
    "vgpu_pool_reset_peak": (C.c_int, []),
    "vgpu_batched_filter_group_by": (C.c_int, [C.POINTER(Column), C.c_int32, C.c_uint64, C.c_void_p, C.c_int32,
                                               C.POINTER(Expr), C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                               C.POINTER(Agg), C.c_int32, C.c_void_p, C.c_void_p,
                                               C.POINTER(C.c_void_p), C.c_uint64, C.POINTER(C.c_uint64),
                                               C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_per_query_filter_group_by": (C.c_int, [C.POINTER(Column), C.c_int32, C.c_uint64, C.POINTER(C.c_uint64),
                                                 C.c_int32, C.POINTER(Expr), C.c_int32, C.c_int32,
                                                 C.POINTER(C.c_int32), C.c_int32, C.POINTER(Agg), C.c_int32,
                                                 C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.c_uint64,
                                                 C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_batched_ranges_filter_group_by": (C.c_int, [C.POINTER(Column), C.c_int32, C.c_uint64,
                                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int32,
                                                      C.POINTER(Expr), C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                                      C.c_int32, C.POINTER(Agg), C.c_int32, C.c_void_p, C.c_void_p,
                                                      C.POINTER(C.c_void_p), C.c_uint64, C.POINTER(C.c_uint64),
                                                      C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_batched_plan_create": (C.c_int, [C.POINTER(Column), C.c_int32, C.c_uint64, C.c_int32, C.c_uint64,
                                           C.POINTER(Expr), C.c_int32, C.c_int32, C.POINTER(C.c_int32), C.c_int32,
                                           C.POINTER(Agg), C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p),
                                           C.c_uint64, C.POINTER(C.c_void_p)]),
    "vgpu_batched_plan_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int32,
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_void_p]),
    "vgpu_batched_plan_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "vgpu_batched_plan_destroy": (None, [C.c_void_p]),
    "vgpu_execute_plan": (C.c_int, [C.c_char_p, C.POINTER(HostColumn), C.c_int32, C.POINTER(ExecuteOptions),
                                    C.POINTER(C.c_void_p)]),
    "vgpu_result_success": (C.c_int32, [C.c_void_p]),
    "vgpu_result_error": (C.c_char_p, [C.c_void_p]),
    "vgpu_result_num_columns": (C.c_int32, [C.c_void_p]),
    "vgpu_result_column": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(HostColumn)]),
    "vgpu_result_free": (None, [C.c_void_p]),
    "vgpu_analyze_plan_shape": (C.c_int, [C.c_char_p, C.POINTER(PlanShape)]),
    "vgpu_is_row_independent": (C.c_int, [C.c_char_p, C.POINTER(C.c_int32)]),
    "vgpu_estimate_working_set": (C.c_int, [C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_int32),
                                            C.POINTER(C.c_uint64), C.c_char_p, C.c_size_t]),
    "vgpu_estimate_columns_size": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64)]),
    "vgpu_determine_execution_mode": (C.c_int, [C.c_char_p, C.c_uint64, C.c_double, C.POINTER(C.c_int32)]),
    "vgpu_hash_source": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p]),
    "vgpu_fast_math_enabled": (C.c_int32, []),
    "vgpu_set_fast_math": (None, [C.c_int32]),
    "vgpu_extended_semantics_enabled": (C.c_int32, []),
    "vgpu_set_extended_semantics": (None, [C.c_int32]),
    "vgpu_cache_begin_staging": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p]),
    "vgpu_cache_commit_staging": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint64]),
    "vgpu_cache_abort_staging": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p]),
    "vgpu_cache_invalidate_table": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p]),
    "vgpu_cache_invalidate_all": (C.c_int, []),
    "vgpu_cache_current_bytes": (C.c_int, [C.POINTER(C.c_uint64)]),
    "vgpu_cache_table_generation": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_uint64)]),
    "vgpu_cache_replayable_chunk_rows": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p),
                                                   C.c_int32, C.POINTER(C.c_uint64), C.c_size_t,
                                                   C.POINTER(C.c_size_t)]),
    "vgpu_vram_budget_init": (C.c_int, [C.c_double]),
    "vgpu_vram_has_headroom": (C.c_int, [C.c_uint64, C.POINTER(C.c_int32)]),
    "vgpu_vram_budget_bytes": (C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
}

_lib = None


def load():
    """Loads libvgpu.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: run `python -m vectisdb_amd.build` first")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return list(_PROTOS)


def check(status: int):
    if status != 0:
        raise VgpuError(load().vgpu_last_error().decode())


def _stream_ptr(stream):
    if stream is None:
        import torch
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return C.c_void_p(stream)
    return C.c_void_p(stream.cuda_stream)


def _tensor_type(t):
    import torch
    return {torch.int16: INT16, torch.int32: INT32, torch.int64: INT64, torch.float32: FLOAT32,
            torch.float64: FLOAT64, torch.bool: BOOLEAN, torch.uint8: BOOLEAN}[t.dtype]


def column(t, vtype=None) -> Column:
    """Column descriptor of a contiguous 1-D device tensor."""
    assert t.is_contiguous() and t.dim() == 1
    return Column(t.data_ptr(), _tensor_type(t) if vtype is None else vtype, t.numel())


def torch_dtype(vtype):
    import torch
    return {INT16: torch.int16, INT32: torch.int32, INT64: torch.int64, FLOAT32: torch.float32,
            FLOAT64: torch.float64, BOOLEAN: torch.uint8}[vtype]


def empty_column(vtype, rows, device="cuda"):
    """Device buffer for `rows` values of `vtype` (HUGEINT as int64 pairs)."""
    import torch
    if vtype == HUGEINT:
        return torch.empty(max(rows, 1) * 2, dtype=torch.int64, device=device)
    return torch.empty(max(rows, 1), dtype=torch_dtype(vtype), device=device)


def group_by_bound(keys, rows, stream=None) -> int:
    lib = load()
    kc = (Column * max(len(keys), 1))(*keys)
    out = C.c_uint64(0)
    check(lib.vgpu_group_by_bound(kc, len(keys), rows, C.byref(out), _stream_ptr(stream)))
    return out.value


def group_by_outputs(keys, aggs, capacity):
    """Output buffers for group_by(..., out=...): one per key column and aggregate, `capacity` rows each."""
    return ([empty_column(column(k).type, capacity) for k in keys], [empty_column(a[2], capacity) for a in aggs])


def group_by(keys, values, aggs, rows, capacity=None, stream=None, out=None):
    """GROUP BY over device columns.

    keys/values: lists of device tensors; aggs: list of (kind, input_index, output_type). out: optional
    (key_buffers, agg_buffers) from group_by_outputs with at least `capacity` rows, reused by the caller
    across calls (the C ABI writes into caller-owned outputs).
    Returns (key_tensors, agg_tensors, groups) with tensors trimmed to `groups` rows.
    """
    lib = load()
    key_cols = [column(k) for k in keys]
    val_cols = [column(v) for v in values]
    if capacity is None:
        capacity = group_by_bound(key_cols, rows, stream)
    if out is not None:
        key_out, agg_out = out
        if len(key_out) != len(key_cols) or len(agg_out) != len(aggs):
            raise ValueError("group_by: out holds one buffer per key column and per aggregate")
        for t, a in zip(agg_out, aggs):
            if t.numel() < (2 * capacity if a[2] == HUGEINT else capacity):
                raise ValueError("group_by: an output buffer is smaller than the capacity")
        for t in key_out:
            if t.numel() < capacity:
                raise ValueError("group_by: an output buffer is smaller than the capacity")
    else:
        key_out = [empty_column(kc.type, capacity) for kc in key_cols]
        agg_out = [empty_column(a[2], capacity) for a in aggs]
    kc = (Column * max(len(key_cols), 1))(*key_cols)
    vc = (Column * max(len(val_cols), 1))(*val_cols)
    ac = (Agg * max(len(aggs), 1))(*[Agg(*a) for a in aggs])
    kp = (C.c_void_p * max(len(key_out), 1))(*[t.data_ptr() for t in key_out])
    ap = (C.c_void_p * max(len(agg_out), 1))(*[t.data_ptr() for t in agg_out])
    groups = C.c_uint64(0)
    check(lib.vgpu_group_by(kc, len(key_cols), vc, len(val_cols), ac, len(aggs), rows, kp, ap, capacity,
                            C.byref(groups), _stream_ptr(stream)))
    g = groups.value
    trimmed_aggs = [t[: 2 * g] if a[2] == HUGEINT else t[:g] for t, a in zip(agg_out, aggs)]
    return [t[:g] for t in key_out], trimmed_aggs, g


class PreparedGroupBy:
    """vgpu_group_by over fixed columns, aggregates and caller-owned outputs, with its argument arrays built
    once: each call is one C call and returns the group count (the outputs hold the groups). For loops that
    run the same GROUP BY many times, where building the arrays per call would cost more host time than the
    C entry point itself."""

    def __init__(self, keys, values, aggs, rows, capacity, out, stream=None):
        self._lib = load()
        key_out, agg_out = out
        self._keep = (list(keys), list(values), list(key_out), list(agg_out))  # the tensors the pointers name
        key_cols = [column(k) for k in keys]
        val_cols = [column(v) for v in values]
        if len(key_out) != len(key_cols) or len(agg_out) != len(aggs):
            raise ValueError("PreparedGroupBy: out holds one buffer per key column and per aggregate")
        for t, a in zip(agg_out, aggs):
            if t.numel() < (2 * capacity if a[2] == HUGEINT else capacity):
                raise ValueError("PreparedGroupBy: an output buffer is smaller than the capacity")
        for t in key_out:
            if t.numel() < capacity:
                raise ValueError("PreparedGroupBy: an output buffer is smaller than the capacity")
        self._aggs = list(aggs)
        self._args = ((Column * max(len(key_cols), 1))(*key_cols), len(key_cols),
                      (Column * max(len(val_cols), 1))(*val_cols), len(val_cols),
                      (Agg * max(len(aggs), 1))(*[Agg(*a) for a in aggs]), len(aggs), rows,
                      (C.c_void_p * max(len(key_out), 1))(*[t.data_ptr() for t in key_out]),
                      (C.c_void_p * max(len(agg_out), 1))(*[t.data_ptr() for t in agg_out]), capacity)
        self._groups = C.c_uint64(0)
        self._groups_ref = C.byref(self._groups)
        self._stream = _stream_ptr(stream)

    def __call__(self):
        check(self._lib.vgpu_group_by(*self._args, self._groups_ref, self._stream))
        return self._groups.value

    def result(self):
        """(key_tensors, agg_tensors, groups) of the last call, trimmed to its groups."""
        g = self._groups.value
        _, _, key_out, agg_out = self._keep
        return ([t[:g] for t in key_out],
                [t[: 2 * g] if a[2] == HUGEINT else t[:g] for t, a in zip(agg_out, self._aggs)], g)


def sort_column(key, stream=None):
    import torch
    lib = load()
    kc = column(key)
    out = torch.empty_like(key)
    perm = torch.empty(key.numel(), dtype=torch.int32, device=key.device)
    check(lib.vgpu_sort_column(C.byref(kc), C.c_void_p(out.data_ptr()), C.c_void_p(perm.data_ptr()),
                               _stream_ptr(stream)))
    return out, perm


def filter_gt_i32(x, threshold, stream=None):
    import torch
    lib = load()
    out = torch.empty(max(x.numel(), 1), dtype=torch.int32, device=x.device)
    count = C.c_uint64(0)
    check(lib.vgpu_filter_gt_i32(C.c_void_p(x.data_ptr()), x.numel(), threshold, C.c_void_p(out.data_ptr()),
                                 C.byref(count), _stream_ptr(stream)))
    return out[: count.value]


def filter_gt_i32_noalloc(x, out, threshold, stream=None):
    """vgpu_filter_gt_i32_noalloc into the caller's device buffer `out` (int32, at least x.numel() slots);
    returns the kept count. The library allocates nothing for it."""
    if out.numel() < x.numel():
        raise ValueError("filter_gt_i32_noalloc: output buffer smaller than the input")
    count = C.c_uint64(0)
    check(load().vgpu_filter_gt_i32_noalloc(C.c_void_p(x.data_ptr()), x.numel(), threshold,
                                            C.c_void_p(out.data_ptr()), C.byref(count), _stream_ptr(stream)))
    return count.value


def generate_column(out, vtype, dist, lo=0, hi=0, seed=0, stream_id=0, cdf=None, perm=None, stream=None):
    lib = load()
    check(lib.vgpu_generate_column(C.c_void_p(out.data_ptr()), vtype, out.numel() if vtype != HUGEINT else out.numel() // 2,
                                   dist, lo, hi, seed, stream_id,
                                   C.c_void_p(cdf.data_ptr() if cdf is not None else 0),
                                   cdf.numel() if cdf is not None else 0,
                                   C.c_void_p(perm.data_ptr() if perm is not None else 0), _stream_ptr(stream)))


def device_copy(dst, src, nbytes=None, stream=None):
    """dst[:nbytes] = src[:nbytes] (device tensors, bytes) through the library's copy kernel."""
    n = nbytes if nbytes is not None else src.numel() * src.element_size()
    if n > dst.numel() * dst.element_size() or n > src.numel() * src.element_size():
        raise ValueError("device_copy: byte count exceeds a tensor")
    check(load().vgpu_device_copy(C.c_void_p(dst.data_ptr()), C.c_void_p(src.data_ptr()), n, _stream_ptr(stream)))


def batched_filter_group_by(cols, segment_offsets, preds, key_col, value_cols, aggs, capacity=None, stream=None,
                            per_query=False):
    """Batched form: cols are device tensors of one table, segment_offsets a device int64 tensor of
    nseg + 1 row offsets; preds as (body, [col indices], BOOLEAN); aggs as (kind, value index, out type).
    Returns (segment ids, keys, [aggregate columns], per-segment group offsets as a Python list).
    per_query=True runs the same queries one by one through the plain path (same outputs)."""
    import torch
    lib = load()
    rows = cols[0].numel()
    nseg = segment_offsets.numel() - 1
    cc = (Column * len(cols))(*[column(c) for c in cols])
    pa, keep = _exprs(preds)
    vc = (C.c_int32 * max(len(value_cols), 1))(*value_cols)
    ac = (Agg * max(len(aggs), 1))(*[Agg(*a) for a in aggs])
    if capacity is None:
        capacity = max(rows, 1)
    # an out-of-range key column is left for the library to reject with its own message
    key_type = _tensor_type(cols[key_col]) if 0 <= key_col < len(cols) else INT64
    seg_out = torch.empty(capacity, dtype=torch.int32, device=cols[0].device)
    key_out = empty_column(key_type, capacity)
    agg_out = [empty_column(a[2], capacity) for a in aggs]
    ap = (C.c_void_p * max(len(agg_out), 1))(*[t.data_ptr() for t in agg_out])
    groups = C.c_uint64(0)
    offs = (C.c_uint64 * (nseg + 1))()
    if per_query:
        host = segment_offsets.cpu().numpy().astype("uint64")
        hoffs = (C.c_uint64 * (nseg + 1))(*[int(x) for x in host])
        check(lib.vgpu_per_query_filter_group_by(cc, len(cols), rows, hoffs, nseg, pa, len(preds), key_col, vc,
                                                 len(value_cols), ac, len(aggs), C.c_void_p(seg_out.data_ptr()),
                                                 C.c_void_p(key_out.data_ptr()), ap, capacity, C.byref(groups), offs,
                                                 _stream_ptr(stream)))
    else:
        check(lib.vgpu_batched_filter_group_by(cc, len(cols), rows, C.c_void_p(segment_offsets.data_ptr()), nseg,
                                               pa, len(preds), key_col, vc, len(value_cols), ac, len(aggs),
                                               C.c_void_p(seg_out.data_ptr()), C.c_void_p(key_out.data_ptr()), ap,
                                               capacity, C.byref(groups), offs, _stream_ptr(stream)))
    g = groups.value
    trimmed = [t[: 2 * g] if a[2] == HUGEINT else t[:g] for t, a in zip(agg_out, aggs)]
    return seg_out[:g], key_out[:g], trimmed, list(offs)


def _u64s(values):
    values = [int(v) for v in values]
    return (C.c_uint64 * max(len(values), 1))(*values)


def batched_ranges_filter_group_by(cols, row_starts, row_lens, preds, key_col, value_cols, aggs, capacity=None,
                                   stream=None):
    """batched_filter_group_by over row ranges (query g = table rows [row_starts[g], + row_lens[g])), planned
    on the host per batch. Returns (segment ids, keys, [aggregate columns], per-range group offsets)."""
    import torch
    lib = load()
    rows = cols[0].numel()
    nq = len(row_starts)
    cc = (Column * len(cols))(*[column(c) for c in cols])
    pa, keep = _exprs(preds)
    vc = (C.c_int32 * max(len(value_cols), 1))(*value_cols)
    ac = (Agg * max(len(aggs), 1))(*[Agg(*a) for a in aggs])
    if capacity is None:
        capacity = max(int(sum(row_lens)), 1)
    key_type = _tensor_type(cols[key_col]) if 0 <= key_col < len(cols) else INT64
    seg_out = torch.empty(capacity, dtype=torch.int32, device=cols[0].device)
    key_out = empty_column(key_type, capacity)
    agg_out = [empty_column(a[2], capacity) for a in aggs]
    ap = (C.c_void_p * max(len(agg_out), 1))(*[t.data_ptr() for t in agg_out])
    groups = C.c_uint64(0)
    offs = (C.c_uint64 * (nq + 1))()
    check(lib.vgpu_batched_ranges_filter_group_by(cc, len(cols), rows, _u64s(row_starts), _u64s(row_lens), nq, pa,
                                                  len(preds), key_col, vc, len(value_cols), ac, len(aggs),
                                                  C.c_void_p(seg_out.data_ptr()), C.c_void_p(key_out.data_ptr()), ap,
                                                  capacity, C.byref(groups), offs, _stream_ptr(stream)))
    g = groups.value
    trimmed = [t[: 2 * g] if a[2] == HUGEINT else t[:g] for t, a in zip(agg_out, aggs)]
    return seg_out[:g], key_out[:g], trimmed, list(offs)


class BatchedPlan:
    """The batched form prepared once over a device-resident table and run as one HIP graph per batch.

    cols: device tensors of the table; at most max_ranges row ranges holding at most max_rows rows per
    batch; preds / key_col / value_cols / aggs as in batched_filter_group_by. The plan owns its output
    columns (capacity rows; default: the group bound). run() returns views of them, valid until the next
    run()."""

    def __init__(self, cols, max_ranges, max_rows, preds, key_col, value_cols, aggs, capacity=None):
        import torch
        lib = load()
        self._lib = lib
        self._handle = None
        self.cols = list(cols)  # the plan reads these buffers at every run
        rows = cols[0].numel()
        self.aggs = list(aggs)
        self.max_ranges = max_ranges
        if capacity is None:
            capacity = max(int(max_rows), 1)
        key_type = _tensor_type(cols[key_col]) if 0 <= key_col < len(cols) else INT64
        self.seg_out = torch.empty(capacity, dtype=torch.int32, device=cols[0].device)
        self.key_out = empty_column(key_type, capacity)
        self.agg_out = [empty_column(a[2], capacity) for a in aggs]
        cc = (Column * len(cols))(*[column(c) for c in cols])
        pa, keep = _exprs(preds)
        vc = (C.c_int32 * max(len(value_cols), 1))(*value_cols)
        ac = (Agg * max(len(aggs), 1))(*[Agg(*a) for a in aggs])
        ap = (C.c_void_p * max(len(self.agg_out), 1))(*[t.data_ptr() for t in self.agg_out])
        h = C.c_void_p()
        check(lib.vgpu_batched_plan_create(cc, len(cols), rows, max_ranges, max_rows, pa, len(preds), key_col, vc,
                                           len(value_cols), ac, len(aggs), C.c_void_p(self.seg_out.data_ptr()),
                                           C.c_void_p(self.key_out.data_ptr()), ap, capacity, C.byref(h)))
        self._handle = h

    def run(self, row_starts, row_lens, stream=None):
        nq = len(row_starts)
        if len(row_lens) != nq:
            raise ValueError("row_starts and row_lens differ in length")
        groups = C.c_uint64(0)
        offs = (C.c_uint64 * (nq + 1))()
        check(self._lib.vgpu_batched_plan_run(self._handle, _u64s(row_starts), _u64s(row_lens), nq, C.byref(groups),
                                              offs, _stream_ptr(stream)))
        g = groups.value
        trimmed = [t[: 2 * g] if a[2] == HUGEINT else t[:g] for t, a in zip(self.agg_out, self.aggs)]
        return self.seg_out[:g], self.key_out[:g], trimmed, list(offs)

    def info(self):
        vals = [C.c_uint64(0) for _ in range(5)]
        check(self._lib.vgpu_batched_plan_info(self._handle, *[C.byref(v) for v in vals]))
        return dict(zip(["kernel_nodes", "copy_nodes", "other_nodes", "footprint_bytes", "group_bound"],
                        [v.value for v in vals]))

    def close(self):
        if self._handle is not None:
            self._lib.vgpu_batched_plan_destroy(self._handle)
            self._handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _exprs(exprs):
    """exprs: list of (body, [input column indices], output_type) -> ctypes array (kept alive by caller)."""
    keep = []
    arr = (Expr * max(len(exprs), 1))()
    for i, (body, inputs, otype) in enumerate(exprs):
        ins = (C.c_int32 * max(len(inputs), 1))(*inputs)
        b = body.encode() if isinstance(body, str) else body
        keep += [ins, b]
        arr[i] = Expr(b, ins, len(inputs), otype)
    return arr, keep


def filter_project(cols, preds, projs, valid=None, out_valid=False, emit_selection=False, stream=None):
    """Fused FILTER + PROJECTION over device tensors.

    cols: list of 1-D device tensors; preds/projs: list of (body, [col indices], output_type);
    valid: optional list (per column) of uint8 device tensors or None. Returns (outputs, out_valid, sel, kept).
    """
    import torch
    lib = load()
    rows = cols[0].numel() if cols else 0
    cc = (Column * max(len(cols), 1))(*[column(c) for c in cols])
    vptr = None
    if valid is not None:
        vptr = (C.c_void_p * len(cols))(*[(v.data_ptr() if v is not None else 0) for v in valid])
    pa, k1 = _exprs(preds)
    ja, k2 = _exprs(projs)
    outs = [empty_column(p[2], rows) for p in projs]
    op = (C.c_void_p * max(len(outs), 1))(*[o.data_ptr() for o in outs])
    ov = None
    ovt = []
    if out_valid:
        ovt = [torch.empty(max(rows, 1), dtype=torch.uint8, device=cols[0].device) for _ in projs]
        ov = (C.c_void_p * max(len(ovt), 1))(*[o.data_ptr() for o in ovt])
    sel = torch.empty(max(rows, 1), dtype=torch.int32, device=cols[0].device) if emit_selection else None
    kept = C.c_uint64(0)
    check(lib.vgpu_filter_project(cc, vptr, len(cols), pa, len(preds), ja, len(projs), op, ov,
                                  C.c_void_p(sel.data_ptr() if sel is not None else 0), C.byref(kept),
                                  _stream_ptr(stream)))
    k = kept.value
    return ([o[:k] for o in outs], [o[:k] for o in ovt], sel[:k] if sel is not None else None, k)


def cross_filter(left, right, preds, projs, capacity=None, stream=None):
    """Left-major cross product of `left` x `right` column lists with a fused filter.

    Column index i < len(left) is left[i], else right[i - len(left)]. Returns (outputs, kept).
    """
    lib = load()
    cols = list(left) + list(right)
    nl = left[0].numel()
    nr = right[0].numel()
    if capacity is None:
        capacity = nl * nr
    cc = (Column * len(cols))(*[column(c) for c in cols])
    pa, k1 = _exprs(preds)
    ja, k2 = _exprs(projs)
    outs = [empty_column(p[2], capacity) for p in projs]
    op = (C.c_void_p * max(len(outs), 1))(*[o.data_ptr() for o in outs])
    kept = C.c_uint64(0)
    check(lib.vgpu_cross_filter(cc, len(cols), len(left), nl, nr, pa, len(preds), ja, len(projs), op, capacity,
                                C.byref(kept), _stream_ptr(stream)))
    k = kept.value
    return [o[:k] for o in outs], k


def join_aggregate(akey, avals, bkey, bvals, aggs, stream=None):
    """Global aggregates over akey = bkey. aggs: list of (kind, side, input, output_type).
    Returns Python values (int for COUNT/SUM(int)/MIN/MAX ints, float otherwise)."""
    import struct
    lib = load()
    ak = column(akey)
    bk = column(bkey)
    av = (Column * max(len(avals), 1))(*[column(v) for v in avals])
    bv = (Column * max(len(bvals), 1))(*[column(v) for v in bvals])
    ja = (JoinAgg * max(len(aggs), 1))(*[JoinAgg(*a) for a in aggs])
    res = C.create_string_buffer(16 * max(len(aggs), 1))
    check(lib.vgpu_join_aggregate(C.byref(ak), av, len(avals), C.byref(bk), bv, len(bvals), ja, len(aggs), res,
                                  _stream_ptr(stream)))
    out = []
    raw = res.raw
    for i, (kind, side, inp, otype) in enumerate(aggs):
        chunk = raw[16 * i: 16 * i + 16]
        if otype == HUGEINT:
            lo, hi = struct.unpack("<Qq", chunk)
            out.append((hi << 64) + lo)
        elif otype in (FLOAT64, FLOAT32):
            out.append(struct.unpack("<d", chunk[:8])[0])
        else:
            out.append(struct.unpack("<q", chunk[:8])[0])
    return out


def join_pairs(akey, bkey, capacity=None, stream=None):
    """(a_rows, b_rows) int32 device tensors of the INNER equi-join."""
    import torch
    lib = load()
    ak = column(akey)
    bk = column(bkey)
    count = C.c_uint64(0)
    if capacity is None:
        st = lib.vgpu_join_pairs(C.byref(ak), C.byref(bk), None, None, 0, C.byref(count), _stream_ptr(stream))
        if st == 0:
            capacity = 0
        else:
            capacity = count.value
    a = torch.empty(max(capacity, 1), dtype=torch.int32, device=akey.device)
    b = torch.empty(max(capacity, 1), dtype=torch.int32, device=akey.device)
    check(lib.vgpu_join_pairs(C.byref(ak), C.byref(bk), C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                              capacity, C.byref(count), _stream_ptr(stream)))
    n = count.value
    return a[:n], b[:n]


def sort_autotune(on=True):
    check(load().vgpu_sort_autotune(1 if on else 0))


def group_by_fused_last_pass(mode):
    """0 never, 1 when it moves fewer bytes (default), 2 whenever the aggregate mix is supported."""
    check(load().vgpu_group_by_fused_last_pass(int(mode)))


def sort_autotune_report():
    import json
    lib = load()
    need = C.c_size_t(0)
    buf = C.create_string_buffer(1 << 16)
    check(lib.vgpu_sort_autotune_report(buf, len(buf), C.byref(need)))
    return json.loads(buf.value.decode())


def profile_enable(on=True):
    check(load().vgpu_profile_enable(1 if on else 0))


def profile_only(name=None):
    """Times only the launches profiled under `name` (None: every launch) while profiling is enabled."""
    check(load().vgpu_profile_only(name.encode() if name else None))


def profile_reset():
    check(load().vgpu_profile_reset())


def profile_read():
    lib = load()
    needed = C.c_size_t(0)
    buf = C.create_string_buffer(1 << 16)
    st = lib.vgpu_profile_read(buf, len(buf), C.byref(needed))
    if st != 0:
        buf = C.create_string_buffer(needed.value)
        check(lib.vgpu_profile_read(buf, len(buf), C.byref(needed)))
    return json.loads(buf.value.decode())


def pool_reset_peak():
    check(load().vgpu_pool_reset_peak())


def pool_stats():
    a, b, c = C.c_uint64(), C.c_uint64(), C.c_uint64()
    check(load().vgpu_pool_stats(C.byref(a), C.byref(b), C.byref(c)))
    return {"in_use": a.value, "cached": b.value, "peak": c.value}


_NP_DTYPES = {INT16: "int16", INT32: "int32", INT64: "int64", FLOAT32: "float32", FLOAT64: "float64",
              BOOLEAN: "uint8", HUGEINT: "int64"}


class ServedResult:
    """A finished served query: info fields, and columns copied to host numpy arrays (HUGEINT as an (n, 2)
    int64 array of low/high words); JOIN_AGG scalars carry a NULL flag."""

    def __init__(self, info, columns, types, nulls, error):
        self.status = info.status
        self.kind = info.kind
        self.rows = info.rows
        self.batch_size = info.batch_size
        self.latency_s = (info.done_ns - info.submit_ns) * 1e-9
        self.columns = columns
        self.types = types
        self.nulls = nulls
        self.error = error


class Server:
    """The served form over device-resident tables (tensors must stay alive while the server lives)."""

    def __init__(self, slots=4, batching=True, max_batch=512, memory_budget=0, fast_math=False, max_batch_rows=0,
                 resident_join=False):
        lib = load()
        self._lib = lib
        self._keep = []
        opts = ServerOptions(slots, 1 if batching else 0, max_batch, memory_budget, 1 if fast_math else 0,
                             max_batch_rows, 1 if resident_join else 0)
        h = C.c_void_p()
        check(lib.vgpu_server_create(C.byref(opts), C.byref(h)))
        self._h = h

    def close(self):
        if self._h:
            check(self._lib.vgpu_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _add(self, t, keep):
        self._keep.append(keep)
        tid = C.c_int32(0)
        check(self._lib.vgpu_server_add_template(self._h, C.byref(t), C.byref(tid)))
        return tid.value

    def add_filter_project(self, cols, preds, projs):
        cc = (Column * len(cols))(*[column(c) for c in cols])
        pa, k1 = _exprs(preds)
        qa, k2 = _exprs(projs)
        t = Template(kind=TEMPLATE_FILTER_PROJECT, cols=cc, ncols=len(cols), preds=pa, npreds=len(preds), projs=qa,
                     nprojs=len(projs), key_col=-1)
        return self._add(t, (cc, pa, k1, qa, k2))

    def add_group_by(self, cols, preds, key_col, value_cols, aggs):
        cc = (Column * len(cols))(*[column(c) for c in cols])
        pa, k1 = _exprs(preds)
        vc = (C.c_int32 * max(len(value_cols), 1))(*value_cols)
        ac = (Agg * max(len(aggs), 1))(*[Agg(*a) for a in aggs])
        t = Template(kind=TEMPLATE_GROUP_BY, cols=cc, ncols=len(cols), preds=pa, npreds=len(preds), key_col=key_col,
                     value_cols=vc, nvalues=len(value_cols), aggs=ac, naggs=len(aggs))
        return self._add(t, (cc, pa, k1, vc, ac))

    def add_join_agg(self, probe_key, build_key, build_payload):
        cc = (Column * 1)(column(probe_key))
        t = Template(kind=TEMPLATE_JOIN_AGG, cols=cc, ncols=1, key_col=-1, build_key=column(build_key),
                     build_payload=column(build_payload))
        return self._add(t, (cc,))

    def add_cross(self, left_cols, right_cols, preds, projs):
        """Queries take a row range of the right side; expression slots number the left columns first."""
        lc = (Column * len(left_cols))(*[column(c) for c in left_cols])
        rc = (Column * len(right_cols))(*[column(c) for c in right_cols])
        pa, k1 = _exprs(preds)
        qa, k2 = _exprs(projs)
        t = Template(kind=TEMPLATE_CROSS, cols=rc, ncols=len(right_cols), preds=pa, npreds=len(preds), projs=qa,
                     nprojs=len(projs), key_col=-1, left_cols=lc, nleft=len(left_cols))
        return self._add(t, (lc, rc, pa, k1, qa, k2))

    def submit(self, template_id, row_start, row_count):
        tk = C.c_uint64(0)
        check(self._lib.vgpu_server_submit(self._h, template_id, row_start, row_count, C.byref(tk)))
        return tk.value

    def wait(self, ticket):
        import numpy as np
        lib = self._lib
        r = C.c_void_p()
        check(lib.vgpu_server_wait(self._h, ticket, C.byref(r)))
        try:
            info = ServedInfo()
            check(lib.vgpu_served_result_info(r, C.byref(info)))
            cols, types, nulls = [], [], []
            if info.status == 0:
                for i in range(info.ncols):
                    ptr, ty, nl = C.c_void_p(), C.c_int32(0), C.c_int32(0)
                    check(lib.vgpu_served_result_column(r, i, C.byref(ptr), C.byref(ty), C.byref(nl)))
                    n = info.rows * (2 if ty.value == HUGEINT else 1)
                    a = np.empty(n, dtype=_NP_DTYPES[ty.value])
                    if n:
                        check(lib.vgpu_served_result_copy(r, i, C.c_void_p(a.ctypes.data)))
                    cols.append(a.reshape(-1, 2) if ty.value == HUGEINT else a)
                    types.append(ty.value)
                    nulls.append(bool(nl.value))
            err = lib.vgpu_served_result_error(r).decode()
            return ServedResult(info, cols, types, nulls, err)
        finally:
            lib.vgpu_served_result_free(r)

    def stats(self):
        s = ServerStats()
        check(self._lib.vgpu_server_stats_read(self._h, C.byref(s)))
        return {f: getattr(s, f) for f, _ in ServerStats._fields_}

    def run_open_loop(self, tmpl, starts, counts, arrivals):
        """Submits query i at arrivals[i] seconds (open loop); returns (latencies, batch sizes, elapsed s)."""
        import numpy as np
        n = len(tmpl)
        t = np.ascontiguousarray(tmpl, dtype=np.int32)
        s = np.ascontiguousarray(starts, dtype=np.uint64)
        c = np.ascontiguousarray(counts, dtype=np.uint64)
        a = np.ascontiguousarray(arrivals, dtype=np.float64)
        lat = np.zeros(n, dtype=np.float64)
        bs = np.zeros(n, dtype=np.uint32)
        el = C.c_double(0)
        check(self._lib.vgpu_server_run_open_loop(
            self._h, t.ctypes.data_as(C.POINTER(C.c_int32)), s.ctypes.data_as(C.POINTER(C.c_uint64)),
            c.ctypes.data_as(C.POINTER(C.c_uint64)), a.ctypes.data_as(C.POINTER(C.c_double)), n,
            lat.ctypes.data_as(C.POINTER(C.c_double)), bs.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(el)))
        return lat, bs, el.value
