/* vgpu.h -- the extern "C" boundary of the vectis MI355X engine (libvgpu.so).
 *
 * Plain pointers and sizes only. Every function returns VGPU_OK (0) on success and VGPU_ERROR (1) on
 * failure, with the reason in vgpu_last_error() (thread-local). Unless a comment says otherwise, pointers
 * named *_dev / columns are DEVICE pointers, and `stream` is a hipStream_t (NULL = the null stream); calls
 * are ordered on that stream and return once their work is queued, except where they say they
 * synchronize (calls that must read a size back).
 *
 * Each entry point names the reference interface it replaces (paths relative to the reference tree).
 * The C++ drop-in headers (gpu_engine.hpp and companions, same declarations as cuda_engine/include) sit on
 * top of the same library; this C layer is what ctypes / cgo / JNI callers and the parity tests bind.
 */
#ifndef VGPU_H
#define VGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VGPU_OK 0
#define VGPU_ERROR 1

/* Value types: numbered exactly like vector_gpu::GpuValueType (cuda_engine/include/gpu_engine.hpp:29-38). */
enum vgpu_type {
	VGPU_INT16 = 0,
	VGPU_INT32 = 1,
	VGPU_INT64 = 2,
	VGPU_FLOAT32 = 3,
	VGPU_FLOAT64 = 4,
	VGPU_BOOLEAN = 5,
	VGPU_DICTIONARY_STRING = 6,
	VGPU_HUGEINT = 7
};

/* Aggregate kinds: numbered like vector_gpu::GpuAggregateKind (cuda_engine/include/gpu_engine.hpp:85-92). */
enum vgpu_agg_kind { VGPU_COUNT_STAR = 0, VGPU_COUNT = 1, VGPU_SUM = 2, VGPU_MIN = 3, VGPU_MAX = 4, VGPU_AVG = 5 };

typedef struct vgpu_column {
	const void *data; /* device pointer, tightly packed */
	int32_t type;     /* enum vgpu_type */
	uint64_t rows;
} vgpu_column;

typedef struct vgpu_agg {
	int32_t kind;        /* enum vgpu_agg_kind */
	int32_t input;       /* index into the value columns, -1 for COUNT(*) */
	int32_t output_type; /* enum vgpu_type: COUNT -> INT64, SUM(int) -> HUGEINT, SUM(float) -> FLOAT64,
	                        MIN/MAX -> input type, AVG -> FLOAT64 */
} vgpu_agg;

/* ---------------------------------------------------------------- library */

/* Reason for the last VGPU_ERROR returned on this thread (empty string if none). */
const char *vgpu_last_error(void);
/* ABI version of this header (bumped on any signature change). */
int vgpu_abi_version(void);
/* Makes `device` current for this thread and initializes the runtime; replaces the context bring-up in
 * RmmPool::EnsureInitialized (cuda_engine/src/rmm_pool.cpp:90-150). */
int vgpu_device_init(int device);

/* ---------------------------------------------------------------- GROUP BY (radix sort + segmented reduce)
 * Replaces ExecuteGroupByNode (cuda_engine/src/operators/gpu_groupby.cu:303-592) for device-resident
 * columns. Keys: 0-4 integer/boolean columns (0 = one global group). Outputs: one device buffer per key
 * (key type) and per aggregate (output_type), each with room for `capacity` rows; the group count is
 * written to *groups (host). Synchronizes. Fails (without writing) when `capacity` is below the bound that
 * vgpu_group_by_bound returns. */
int vgpu_group_by_bound(const vgpu_column *keys, int32_t nkeys, uint64_t rows, uint64_t *bound, void *stream);
int vgpu_group_by(const vgpu_column *keys, int32_t nkeys, const vgpu_column *values, int32_t nvalues,
                  const vgpu_agg *aggs, int32_t naggs, uint64_t rows, void *const *key_out, void *const *agg_out,
                  uint64_t capacity, uint64_t *groups, void *stream);

/* Stable sort of one integer key column: keys_out receives the sorted keys, perm_out (uint32, may be NULL)
 * the source row of each output row. Replaces thrust::sequence + stable_sort_by_key
 * (cuda_engine/src/operators/gpu_groupby.cu:352-376, cuda_engine/src/operators/gpu_hash_join.cu:205-206).
 * Requires rows < 2^30. */
int vgpu_sort_column(const vgpu_column *key, void *keys_out, uint32_t *perm_out, void *stream);

/* ---------------------------------------------------------------- fused filter / projection (hipRTC)
 * Expression bodies are GpuExpr::generated_cuda_source text (cuda_engine/include/gpu_engine.hpp:56-61):
 * they read ((const T*)inputs[k])[row] and write ((U*)out)[idx]. inputs[k] of an expression is column
 * `inputs[k]` of the column array. */
typedef struct vgpu_expr {
	const char *body;
	const int32_t *inputs;
	int32_t ninputs;
	int32_t output_type; /* enum vgpu_type; BOOLEAN for predicates */
} vgpu_expr;

/* FILTER + PROJECTION in one kernel. Replaces ExecuteFilter + ExecuteProjection
 * (cuda_engine/src/gpu_executor.cu:645-746) and the per-expression NVRTC kernels they launch. Rows where
 * every predicate is true (a NULL predicate counts as false, as in SQL WHERE) are written in row order
 * through every projection into outs[e] (capacity: rows). col_valid[c]: dense byte validity (1 = valid)
 * or NULL (no NULLs); may itself be NULL. out_valid[e]: NULL or a byte buffer receiving output validity
 * (AND of the expression's inputs). sel_out (optional) receives the kept row ids. *kept receives the
 * output row count (synchronizes). With npreds == 0 every row is projected. */
int vgpu_filter_project(const vgpu_column *cols, const uint8_t *const *col_valid, int32_t ncols,
                        const vgpu_expr *preds, int32_t npreds, const vgpu_expr *projs, int32_t nprojs,
                        void *const *outs, uint8_t *const *out_valid, uint32_t *sel_out, uint64_t *kept,
                        void *stream);

/* CROSS_PRODUCT with an optional fused FILTER. Replaces ExecuteCrossProductNode
 * (cuda_engine/src/operators/gpu_cross_join.cu:111-192) plus a FILTER above it. Columns 0..nleft-1 are
 * the left side (left_rows rows), nleft..ncols-1 the right side (right_rows rows); pairs are enumerated
 * left-major (pair f = (f / right_rows, f % right_rows)), so output order matches the reference's. Kept
 * pairs are written through the projections (outs capacity: left_rows * right_rows when npreds == 0,
 * else the caller's bound `capacity`). Fails when more than `capacity` pairs are kept. Synchronizes. */
int vgpu_cross_filter(const vgpu_column *cols, int32_t ncols, int32_t nleft, uint64_t left_rows, uint64_t right_rows,
                      const vgpu_expr *preds, int32_t npreds, const vgpu_expr *projs, int32_t nprojs,
                      void *const *outs, uint64_t capacity, uint64_t *kept, void *stream);

/* ---------------------------------------------------------------- equi-join
 * Side A = children[0] (build_keys), side B = children[1] (probe_keys), one integer key of the same type on
 * both sides, INNER semantics (cuda_engine/src/operators/gpu_hash_join.cu:25-38). */
typedef struct vgpu_join_agg {
	int32_t kind;        /* enum vgpu_agg_kind */
	int32_t side;        /* 0 = A, 1 = B; ignored for COUNT(*) */
	int32_t input;       /* index into that side's value columns */
	int32_t output_type; /* COUNT -> INT64, SUM(int) -> HUGEINT, SUM(float) -> FLOAT64, AVG -> FLOAT64,
	                        MIN/MAX -> input type (BIGINT/DOUBLE) */
} vgpu_join_agg;

/* Global aggregates over A JOIN B ON akey = bkey without materializing the join. When every aggregate is a
 * COUNT or an integer SUM/AVG of one column of the smaller side and/or one column of the larger side, the
 * smaller side becomes a direct-address table (per key: row count, payload sum) and the larger side is read
 * once against it (partitioned by key range when the table outgrows L2 or the larger side's payload is
 * summed); otherwise both sides are radix-sorted, reduced to distinct keys and merge-path matched. Both
 * forms give the same results. Replaces a GROUP_BY_AGGREGATE without keys over
 * a HASH_JOIN (cuda_engine/src/operators/gpu_hash_join.cu:138-267 followed by gpu_groupby.cu). results:
 * naggs x 16 host bytes, each the aggregate in its output_type (HUGEINT = {lower, upper}). Synchronizes. */
int vgpu_join_aggregate(const vgpu_column *akey, const vgpu_column *avals, int32_t navals, const vgpu_column *bkey,
                        const vgpu_column *bvals, int32_t nbvals, const vgpu_join_agg *aggs, int32_t naggs,
                        void *results, void *stream);

/* Materialized join as row-id pairs (uint32), ordered by key, then A row, then B row. *count receives the
 * pair count; fails without writing when it exceeds `capacity` (call again with enough room). Replaces
 * the lower/upper_bound + scan + ExpandMatchesKernel steps of gpu_hash_join.cu:196-254. Synchronizes. */
int vgpu_join_pairs(const vgpu_column *akey, const vgpu_column *bkey, uint32_t *a_rows, uint32_t *b_rows,
                    uint64_t capacity, uint64_t *count, void *stream);

/* ---------------------------------------------------------------- batched form
 * Many `filter + GROUP BY` queries, each over its own row group of one device-resident table, in one
 * pipeline: row group g is rows [segment_offsets[g], segment_offsets[g+1]) (device, nsegments + 1 entries).
 * Every query applies the same predicates (inputs index `cols`) and groups by column `key_col` with the
 * aggregates `aggs` (inputs index `value_cols`). Outputs (device, `capacity` rows each): segment_out (INT32
 * segment of each group), key_out, agg_out[i]; groups are ordered by (segment, key), so query g's groups are
 * rows [segment_group_offsets[g], segment_group_offsets[g+1]) (host array, nsegments + 1 entries). Each
 * query's rows equal what the per-query path (vgpu_filter_project then vgpu_group_by) returns for it.
 * No single reference counterpart: it batches many ExecuteGroupByNode calls
 * (cuda_engine/src/operators/gpu_groupby.cu:303-592) over FILTER outputs (gpu_executor.cu:645-730).
 * Synchronizes. */
int vgpu_batched_filter_group_by(const vgpu_column *cols, int32_t ncols, uint64_t rows,
                                 const uint64_t *segment_offsets, int32_t nsegments, const vgpu_expr *preds,
                                 int32_t npreds, int32_t key_col, const int32_t *value_cols, int32_t nvalues,
                                 const vgpu_agg *aggs, int32_t naggs, void *segment_out, void *key_out,
                                 void *const *agg_out, uint64_t capacity, uint64_t *groups,
                                 uint64_t *segment_group_offsets, void *stream);

/* The same queries through the plain per-query path, one after another (per query: filter kernel, kept
 * count read back, key range, GROUP BY, group count read back), with the same arguments and output layout
 * as vgpu_batched_filter_group_by; segment_offsets_host is the host copy of the row offsets. The baseline
 * the batched form is compared with. Synchronizes. */
int vgpu_per_query_filter_group_by(const vgpu_column *cols, int32_t ncols, uint64_t rows,
                                   const uint64_t *segment_offsets_host, int32_t nsegments, const vgpu_expr *preds,
                                   int32_t npreds, int32_t key_col, const int32_t *value_cols, int32_t nvalues,
                                   const vgpu_agg *aggs, int32_t naggs, void *segment_out, void *key_out,
                                   void *const *agg_out, uint64_t capacity, uint64_t *groups,
                                   uint64_t *segment_group_offsets, void *stream);

/* vgpu_batched_filter_group_by over row ranges instead of contiguous row groups: query g reads table rows
 * [row_starts[g], row_starts[g] + row_lens[g]) (host arrays of nranges entries; ranges may overlap or
 * repeat). Same outputs, ordered by (range, key). Plans the batch on the host (reads the kept-row count
 * and the kept keys' ranges back). No single reference counterpart, as vgpu_batched_filter_group_by.
 * Synchronizes. */
int vgpu_batched_ranges_filter_group_by(const vgpu_column *cols, int32_t ncols, uint64_t rows,
                                        const uint64_t *row_starts, const uint64_t *row_lens, int32_t nranges,
                                        const vgpu_expr *preds, int32_t npreds, int32_t key_col,
                                        const int32_t *value_cols, int32_t nvalues, const vgpu_agg *aggs,
                                        int32_t naggs, void *segment_out, void *key_out, void *const *agg_out,
                                        uint64_t capacity, uint64_t *groups, uint64_t *segment_group_offsets,
                                        void *stream);

/* The batched form prepared once and replayed as ONE HIP graph per batch. create fixes the table, the
 * predicates, the GROUP BY and the bounds: at most max_ranges row ranges holding at most max_rows rows per
 * batch (max_rows + 3 x max_ranges < 2^30). The group key is packed from the range count and the key
 * column's table-wide value range (read once here), every buffer is sized for the bounds (capacity must
 * cover min(max_rows, max_ranges x key range), else VGPU_ERROR), and the pipeline -- descriptor upload,
 * device-side work-queue layout, the persistent filter kernel that sweeps the queue's tiles, radix passes
 * and segmented reduction reading the kept-row count on the device, per-range group offsets, result
 * download -- is captured into a graph. The output columns (capacity rows, as vgpu_batched_filter_group_by)
 * belong to the caller and must outlive the plan. No reference counterpart: the reference runs one
 * GROUP BY per query (cuda_engine/src/operators/gpu_groupby.cu:303-592) with one chunk in flight
 * (gpu_executor_internal.hpp:100-108).
 * run: one batch of nranges ranges, checked against the table and the bounds before any GPU work; one
 * hipGraphLaunch on `stream`, then waits. *groups and segment_group_offsets (nranges + 1 entries) as in
 * vgpu_batched_filter_group_by; results equal vgpu_batched_ranges_filter_group_by's for the same ranges.
 * Runs of one plan must not overlap.
 * info: the captured graph's kernel / copy / other node counts, the device bytes the plan holds, and the
 * group bound. */
typedef struct vgpu_batched_plan vgpu_batched_plan;
int vgpu_batched_plan_create(const vgpu_column *cols, int32_t ncols, uint64_t rows, int32_t max_ranges,
                             uint64_t max_rows, const vgpu_expr *preds, int32_t npreds, int32_t key_col,
                             const int32_t *value_cols, int32_t nvalues, const vgpu_agg *aggs, int32_t naggs,
                             void *segment_out, void *key_out, void *const *agg_out, uint64_t capacity,
                             vgpu_batched_plan **plan);
int vgpu_batched_plan_run(vgpu_batched_plan *plan, const uint64_t *row_starts, const uint64_t *row_lens,
                          int32_t nranges, uint64_t *groups, uint64_t *segment_group_offsets, void *stream);
int vgpu_batched_plan_info(const vgpu_batched_plan *plan, uint64_t *kernel_nodes, uint64_t *copy_nodes,
                           uint64_t *other_nodes, uint64_t *footprint_bytes, uint64_t *group_bound);
void vgpu_batched_plan_destroy(vgpu_batched_plan *plan);

/* ---------------------------------------------------------------- primitives */

/* Order-preserving copy of in[i] > threshold into out; *count_host receives the kept count (synchronizes).
 * Replaces FilterGreaterThanInt32 (cuda_engine/include/gpu_filter_ops.hpp:32,
 * cuda_engine/src/operators/gpu_filter.cu:36-85). */
int vgpu_filter_gt_i32(const int32_t *in, uint64_t n, int32_t threshold, int32_t *out, uint64_t *count_host,
                       void *stream);

/* vgpu_filter_gt_i32 without any allocation. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The kernel's look-back words, tile ticket and kept count live in a fixed array of 1 MiB per
 * GPU, held until the process ends; n is limited to vgpu_filter_gt_i32_noalloc_max_rows() (2^31) and larger
 * inputs fail with VGPU_ERROR before any GPU work. Same output and *count_host as vgpu_filter_gt_i32
 * (synchronizes). Every call on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * threads, other streams) race and must be serialized by the caller. */
int vgpu_filter_gt_i32_noalloc(const int32_t *in, uint64_t n, int32_t threshold, int32_t *out, uint64_t *count_host,
                               void *stream);
uint64_t vgpu_filter_gt_i32_noalloc_max_rows(void);

/* out[i] = src[idx[i]] for elements of `width` bytes (1/2/4/8/16); idx is uint32 (idx64 = 0) or uint64.
 * Replaces GatherColumn / MaterializeColumn (cuda_engine/src/gpu_executor.cu:144-252). */
int vgpu_gather(const void *src, int32_t width, const void *idx, int32_t idx64, uint64_t n, void *out,
                void *stream);

/* Seeded synthetic column: dist 0 uniform int [lo, hi], 1 uniform double [0,1), 2 normal(0,1), 3 Zipf
 * (cdf/perm device tables, cdf_len entries), 4 sequence lo + i. No reference counterpart (test/bench data). */
int vgpu_generate_column(void *out, int32_t type, uint64_t n, int32_t dist, int64_t lo, int64_t hi,
                         uint64_t seed, uint64_t stream_id, const double *cdf, uint64_t cdf_len,
                         const int64_t *perm, void *stream);

/* dst[0, bytes) = src[0, bytes) on the device (a block-chunked copy kernel; hipMemcpyAsync when a pointer is
 * not 16-byte aligned). The engine's own device-to-device copies use it; the bench times it as the
 * achievable stream-copy bandwidth. No reference counterpart. */
int vgpu_device_copy(void *dst, const void *src, uint64_t bytes, void *stream);

/* ---------------------------------------------------------------- plans
 * A plan is a GpuPlanNode tree (cuda_engine/include/gpu_engine.hpp:135-185) written as JSON:
 *   {"op": "SCAN"|"FILTER"|"PROJECTION"|"HASH_JOIN"|"GROUP_BY_AGGREGATE"|"CROSS_PRODUCT",
 *    "estimated_rows", "estimated_row_width_bytes", "table": {"catalog","schema","table"}, "column_names",
 *    "expressions": [{"source", "inputs", "output_type", "ops"}], "build_keys", "probe_keys",
 *    "output_columns", "group_keys", "aggregate_exprs", "aggregates": [{"kind", "input", "output_type"}],
 *    "children": [...]}
 * with types and kinds by name ("INT64", "SUM") or number. Host columns are HOST memory here. */
typedef struct vgpu_host_column {
	const char *name;
	int32_t type;             /* enum vgpu_type */
	const void *data;         /* host, tightly packed */
	const uint8_t *validity;  /* Arrow bitmap (bit r%8 of byte r/8, 1 = valid) or NULL */
	uint64_t rows;
} vgpu_host_column;

/* GpuExecuteOptions (cuda_engine/include/gpu_engine.hpp:357-385). */
typedef struct vgpu_execute_options {
	int32_t stage_scan_to_cache;
	int64_t cache_chunk_index; /* -1 = off */
	int32_t serve_whole_scan_from_cache;
} vgpu_execute_options;

typedef struct vgpu_result vgpu_result; /* GpuExecutionResult, owned by the library until vgpu_result_free */

/* GpuEngine::Instance().ExecutePlan (cuda_engine/include/gpu_engine.hpp:389-402,
 * cuda_engine/src/gpu_engine.cpp:222-259). Returns VGPU_ERROR only when the plan text cannot be parsed;
 * execution failures come back in the result (vgpu_result_success == 0 plus vgpu_result_error), exactly as
 * ExecutePlan reports them. `options` may be NULL. Synchronizes. */
int vgpu_execute_plan(const char *plan_json, const vgpu_host_column *inputs, int32_t ninputs,
                      const vgpu_execute_options *options, vgpu_result **result);
int32_t vgpu_result_success(const vgpu_result *result);
const char *vgpu_result_error(const vgpu_result *result);
int32_t vgpu_result_num_columns(const vgpu_result *result);
/* Column `index` of the result; its pointers stay valid until vgpu_result_free. */
int vgpu_result_column(const vgpu_result *result, int32_t index, vgpu_host_column *out);
void vgpu_result_free(vgpu_result *result);

/* ---------------------------------------------------------------- routing (no device needed)
 * AnalyzePlanShape, IsRowIndependent, EstimateWorkingSet, EstimateColumnsSize, DetermineExecutionMode
 * (cuda_engine/include/gpu_engine.hpp:248, :329-348; cuda_engine/src/gpu_cost_model.cpp). */
typedef struct vgpu_plan_shape {
	uint64_t scanned_rows;
	uint64_t scanned_bytes;
	uint64_t max_rows;
	double total_ops;
	double ops_per_scanned_row;
	double ops_per_scanned_byte;
	int32_t density; /* 0 INSUFFICIENT, 1 INCONCLUSIVE, 2 SUFFICIENT */
	double amplification;
} vgpu_plan_shape;
int vgpu_analyze_plan_shape(const char *plan_json, vgpu_plan_shape *out);
int vgpu_is_row_independent(const char *plan_json, int32_t *out);
/* largest_operator receives the operator name (NUL-terminated, truncated to `cap`). */
int vgpu_estimate_working_set(const char *plan_json, uint64_t max_chunk_rows, uint64_t *total_bytes,
                              int32_t *chunk_capped, uint64_t *largest_bytes, char *largest_operator, size_t cap);
int vgpu_estimate_columns_size(const char *plan_json, uint64_t *out);
/* *mode: 0 PIPELINE_STREAM_CACHED, 1 PIPELINE_STREAM_UNCACHED. */
int vgpu_determine_execution_mode(const char *plan_json, uint64_t vram_budget, double threshold, int32_t *mode);

/* KernelCache::HashSource (cuda_engine/include/kernel_cache.hpp:56): lowercase hex SHA-256 into hex[65]. */
int vgpu_hash_source(const char *text, size_t len, char *hex);
/* IsGpuFastMathEnabled / SetGpuFastMathEnabled (cuda_engine/include/code_generator.hpp:14-16). */
int32_t vgpu_fast_math_enabled(void);
void vgpu_set_fast_math(int32_t on);
/* IsGpuExtendedSemanticsEnabled / SetGpuExtendedSemanticsEnabled (include/gpu_semantics.hpp): the opt-in
 * DuckDB semantics the reference operators refuse (NULL-aware GROUP BY and join, FLOAT/DOUBLE keys,
 * composite and mixed-width join keys). No reference counterpart. Off by default. */
int32_t vgpu_extended_semantics_enabled(void);
void vgpu_set_extended_semantics(int32_t on);

/* ---------------------------------------------------------------- column cache and device budget
 * GpuColumnCache (cuda_engine/include/gpu_column_cache.hpp:130-175) and RmmPool
 * (cuda_engine/include/rmm_pool.hpp:66-95)
 * operations the extension calls, for plans staged with stage_scan_to_cache. */
int vgpu_cache_begin_staging(const char *catalog, const char *schema, const char *table);
int vgpu_cache_commit_staging(const char *catalog, const char *schema, const char *table, uint64_t scanned_rows);
int vgpu_cache_abort_staging(const char *catalog, const char *schema, const char *table);
int vgpu_cache_invalidate_table(const char *catalog, const char *schema, const char *table);
int vgpu_cache_invalidate_all(void);
int vgpu_cache_current_bytes(uint64_t *out);
int vgpu_cache_table_generation(const char *catalog, const char *schema, const char *table, uint64_t *out);
/* Chunk row counts when every named column is resident (*nchunks = 0 otherwise); fails when more than
 * `cap` chunks. */
int vgpu_cache_replayable_chunk_rows(const char *catalog, const char *schema, const char *table,
                                     const char *const *columns, int32_t ncolumns, uint64_t *rows, size_t cap,
                                     size_t *nchunks);
int vgpu_vram_budget_init(double fraction_of_vram);
int vgpu_vram_has_headroom(uint64_t bytes, int32_t *out);
int vgpu_vram_budget_bytes(uint64_t *total, uint64_t *allocated);

/* ---------------------------------------------------------------- served form
 * A request loop over device-resident tables. Queries are instances of prepared templates applied to a
 * row range of the template's table; `slots` executor threads (each with its own HIP stream) admit the
 * oldest waiting query whose memory footprint fits the unreserved budget and, with batching on, every later
 * waiting query of the same template that still fits (continuous batching: queries arriving while others
 * run join the next batch an executor forms). Batches of different executors overlap on the device.
 * No reference counterpart: the reference serves one plan at a time under a mutex
 * (cuda_engine/src/gpu_engine.cpp:222-259); this is the serving loop around the same operators. */
typedef struct vgpu_server vgpu_server;
typedef struct vgpu_served_result vgpu_served_result;

typedef struct vgpu_server_options {
	int32_t slots;          /* executor threads / HIP streams, 1..16 (0 = 4) */
	int32_t batching;       /* 1: batch same-template queries; 0: every query alone on the plain path */
	int32_t max_batch;      /* queries per batch (0 = 512) */
	uint64_t memory_budget; /* device bytes admitted work may reserve (0 = 3/4 of free memory) */
	int32_t fast_math;
	uint64_t max_batch_rows; /* table rows per batch (0 = no limit but the GROUP BY sort's 2^30) */
	int32_t resident_join;   /* 1: a JOIN_AGG template reduces its build side once and every query probes
	                            it, with batching off too (batching on always does) */
} vgpu_server_options;

enum vgpu_template_kind {
	VGPU_TEMPLATE_FILTER_PROJECT = 0, /* preds + projs over cols: one output column per projection */
	VGPU_TEMPLATE_GROUP_BY = 1,       /* preds, GROUP BY cols[key_col], aggs over value_cols: key, then aggs */
	VGPU_TEMPLATE_JOIN_AGG = 2,       /* SUM(build_payload), COUNT(*) of cols[0] = build_key: HUGEINT, INT64 */
	VGPU_TEMPLATE_CROSS = 3           /* left_cols x a row range of cols, preds + projs over (left..., right...) */
};

typedef struct vgpu_template {
	int32_t kind;               /* enum vgpu_template_kind */
	const vgpu_column *cols;    /* the table (device-resident while the server lives) */
	int32_t ncols;
	const vgpu_expr *preds;     /* inputs index cols */
	int32_t npreds;
	const vgpu_expr *projs;     /* FILTER_PROJECT */
	int32_t nprojs;
	int32_t key_col;            /* GROUP_BY */
	const int32_t *value_cols;  /* GROUP_BY; aggregate inputs index this list */
	int32_t nvalues;
	const vgpu_agg *aggs;       /* GROUP_BY */
	int32_t naggs;
	vgpu_column build_key;      /* JOIN_AGG: integer key without NULLs, same type as cols[0] */
	vgpu_column build_payload;  /* JOIN_AGG: integer payload */
	const vgpu_column *left_cols; /* CROSS: the whole left side; expression slots 0..nleft-1 are left columns */
	int32_t nleft;
} vgpu_template;

typedef struct vgpu_served_info {
	int32_t status;      /* 0 ok; else vgpu_served_result_error holds the message */
	int32_t kind;
	uint64_t rows;       /* kept rows, groups, or 1 */
	int32_t ncols;
	int32_t on_device;   /* column data are device pointers (0: host scalars) */
	uint32_t batch_size; /* queries in the batch that computed this result */
	uint64_t submit_ns, start_ns, done_ns; /* steady clock */
} vgpu_served_info;

typedef struct vgpu_server_stats {
	uint64_t queries, batches, max_batch, peak_reserved, budget, rejected;
} vgpu_server_stats;

int vgpu_server_create(const vgpu_server_options *options, vgpu_server **server);
/* Runs every queued query, then stops the executors. Results already collected stay valid. */
int vgpu_server_destroy(vgpu_server *server);
/* Prepares a template (a JOIN_AGG template with batching on reduces its build side once, here). */
int vgpu_server_add_template(vgpu_server *server, const vgpu_template *tmpl, int32_t *template_id);
int vgpu_server_submit(vgpu_server *server, int32_t template_id, uint64_t row_start, uint64_t row_count,
                       uint64_t *ticket);
/* Blocks until the query is done; the result belongs to the caller until vgpu_served_result_free. */
int vgpu_server_wait(vgpu_server *server, uint64_t ticket, vgpu_served_result **result);
int vgpu_server_stats_read(vgpu_server *server, vgpu_server_stats *stats);
int vgpu_served_result_info(const vgpu_served_result *result, vgpu_served_info *info);
/* Column i: data (device pointer, or host for JOIN_AGG scalars), type, and for scalars a NULL flag (SUM
 * over no matched rows). The memory stays valid until the result is freed; finish device reads first. */
int vgpu_served_result_column(const vgpu_served_result *result, int32_t i, const void **data, int32_t *type,
                              int32_t *is_null);
/* Copies column i (rows x type width bytes) to host memory `dst` (synchronous). */
int vgpu_served_result_copy(const vgpu_served_result *result, int32_t i, void *dst);
const char *vgpu_served_result_error(const vgpu_served_result *result);
void vgpu_served_result_free(vgpu_served_result *result);
/* Open-loop driver: query i (template tmpl[i], rows [starts[i], starts[i] + counts[i])) is submitted at
 * arrival_s[i] seconds after the start whatever the server is doing; latency_s[i] = completion - arrival.
 * *elapsed_s = start to last completion; batch_sizes (optional) = size of the batch that ran query i.
 * Fails if any query fails. */
int vgpu_server_run_open_loop(vgpu_server *server, const int32_t *tmpl, const uint64_t *starts,
                              const uint64_t *counts, const double *arrival_s, uint64_t n, double *latency_s,
                              uint32_t *batch_sizes, double *elapsed_s);

/* ---------------------------------------------------------------- measurement */

/* When on, every kernel launch is bracketed by HIP events on its own stream. */
int vgpu_profile_enable(int32_t on);
/* Restricts the bracketing to launches whose profile name is `name` (NULL or "": every launch), so one
 * kernel can be timed inside a timed loop without events around the others. No reference counterpart (a
 * measurement control of this library). */
int vgpu_profile_only(const char *name);
int vgpu_profile_reset(void);
/* JSON array [{"name":..,"ms":..,"launches":..,"bytes":..}, ...] into buf (synchronizes). Returns the
 * needed size in *needed when buf is too small. */
int vgpu_profile_read(char *buf, size_t cap, size_t *needed);

/* Sort autotuner: when on, the first sort of each shape times every radix-pass variant (digit width, items
 * per thread, payload prefetch) on its own input and caches the fastest for the shape. No reference
 * counterpart (thrust picks its own sort). vgpu_sort_autotune_report writes a JSON array of the tuned shapes. */
int vgpu_sort_autotune(int32_t on);
int vgpu_sort_autotune_report(char *buf, size_t cap, size_t *needed);

/* GROUP BY with the sort's last digit pass fused with the reduction (each tile of that pass reduces its runs
 * of equal keys to partial states, then the partials are combined): 0 never, 1 when it moves fewer bytes
 * (the default), 2 whenever the aggregate mix is supported. Results are the same in every mode. No reference
 * counterpart (the reference sorts fully, then reduces). */
int vgpu_group_by_fused_last_pass(int32_t mode);

/* Device pool statistics (bytes): in use, cached for reuse, and the peak in use since the last reset. The pool
 * is the engine's allocator for every plan temporary (replaces the per-plan DeviceArena,
 * cuda_engine/src/gpu_executor.cu:57-104); the peak is what EstimateWorkingSet predicts. */
int vgpu_pool_stats(uint64_t *in_use, uint64_t *cached, uint64_t *peak);
int vgpu_pool_trim(void);
int vgpu_pool_reset_peak(void);

#ifdef __cplusplus
}
#endif

#endif /* VGPU_H */
