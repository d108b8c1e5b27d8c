/* oracle.h -- CPU restatement of the engine's hot-path semantics. TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library, and only to
 * check results against (or as the timed CPU baseline). The product path (libvgpu.so, vectisdb_amd/) never calls it.
 *
 * What each function restates, with the reference site it follows:
 *   vo_mix_counter / vo_gen_*    the seeded column generator (same mixer as the device generator)
 *   vo_filter_project_cfg0       SELECT a*2+b, c*1.5+a FROM t WHERE a < 500000 AND c > 0.25 with DuckDB's
 *                                 operator typing (INTEGER*INTEGER -> INTEGER, then BIGINT +; DOUBLE * 1.5 +
 *                                 CAST(a AS DOUBLE)), filter order = row order
 *                                 (cuda_engine/src/gpu_executor.cu:645-729, extension/src/expression_translator.cpp:300-314)
 *   vo_group_by_i64              GROUP BY key: COUNT, SUM(int)->int128, MIN, MAX, SUM(double) correctly
 *                                 rounded (Shewchuk exact partials), AVG = SUM/COUNT; groups in ascending
 *                                 signed key order (cuda_engine/src/operators/gpu_groupby.cu:303-592)
 *   vo_join_sum_count            SELECT SUM(payload), COUNT(*) FROM probe JOIN build ON fk = key, INNER,
 *                                 one integer key (cuda_engine/src/operators/gpu_hash_join.cu:138-267)
 *   vo_join_pairs                materialized INNER equi-join as (left row, right row) pairs, sorted
 *   vo_cross_filter_count        COUNT/SUM(x)/SUM(y) of L x R WHERE x + y < c, and the left-major pair list
 *                                 (cuda_engine/src/operators/gpu_cross_join.cu:66-71)
 *   vo_exact_sum                 correctly rounded double sum (the reference for DOUBLE SUM/AVG tolerance)
 *   vo_*_mt                      the same results on all host threads: the timed CPU baselines
 */
#ifndef VECTIS_ORACLE_H
#define VECTIS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

uint64_t vo_mix_counter(uint64_t seed, uint64_t stream, uint64_t i);
void vo_gen_uniform_i64(int64_t *out, uint64_t n, int64_t lo, int64_t hi, uint64_t seed, uint64_t stream);
void vo_gen_uniform_i32(int32_t *out, uint64_t n, int64_t lo, int64_t hi, uint64_t seed, uint64_t stream);
void vo_gen_uniform_f64(double *out, uint64_t n, uint64_t seed, uint64_t stream);

uint64_t vo_filter_project_cfg0(const int32_t *a, const int64_t *b, const double *c, uint64_t n, int32_t a_lt,
                                double c_gt, int64_t *out1, double *out2);

double vo_exact_sum(const double *x, uint64_t n);

/* Per-group outputs, groups in ascending key order; returns the group count. sum_v1 is int128 as
 * (low, high) pairs. Any output pointer may be NULL. */
uint64_t vo_group_by_i64(const int64_t *key, const int64_t *v1, const double *v2, uint64_t n, int64_t *out_key,
                         int64_t *out_count, int64_t *out_sum_v1, int64_t *out_min_v1, int64_t *out_max_v1,
                         double *out_sum_v2, double *out_avg_v2);
/* The same GROUP BY on `threads` threads (stable counting sort with per-thread histograms, then every group
 * reduced by one thread with the same arithmetic): identical outputs. */
uint64_t vo_group_by_i64_mt(const int64_t *key, const int64_t *v1, const double *v2, uint64_t n, int threads,
                            int64_t *out_key, int64_t *out_count, int64_t *out_sum_v1, int64_t *out_min_v1,
                            int64_t *out_max_v1, double *out_sum_v2, double *out_avg_v2);

/* Returns COUNT(*); *sum_lo/*sum_hi receive SUM(payload) as int128. */
uint64_t vo_join_sum_count(const int64_t *build_key, const int32_t *payload, uint64_t nb, const int64_t *fk,
                           uint64_t np, uint64_t *sum_lo, int64_t *sum_hi);

/* Pairs (left_row, right_row) with left[l] == right[r], sorted by (l, r). Returns the count; writes at most
 * `cap` pairs. */
uint64_t vo_join_pairs(const int64_t *left, uint64_t nl, const int64_t *right, uint64_t nr, uint64_t *out_l,
                       uint64_t *out_r, uint64_t cap);

/* L x R WHERE x + y < c (x, y INT32, sum in INT32 as DuckDB types it; the inputs keep it in range). */
uint64_t vo_cross_filter_count(const int32_t *x, uint64_t nl, const int32_t *y, uint64_t nr, int64_t c,
                               int64_t *sum_x, int64_t *sum_y);
uint64_t vo_cross_filter_pairs(const int32_t *x, uint64_t nl, const int32_t *y, uint64_t nr, int64_t c,
                               int32_t *out_x, int32_t *out_y, uint64_t cap);

/* Multithreaded CPU baselines (all host threads): the same results as the single-threaded functions above.
 *   vo_filter_project_cfg0_mt       vo_filter_project_cfg0 split over row ranges
 *   vo_join_sum_count_mt            vo_join_sum_count as a hash join (shared open-addressing build table)
 *   vo_cross_filter_pairs_mt        vo_cross_filter_pairs split over left rows (left-major order kept);
 *                                    writes nothing unless the count fits `cap`
 *   vo_batched_filter_group_by_mt   per query: filter a < a_lt AND c > c_gt, then vo_group_by_i64_mt */
uint64_t vo_filter_project_cfg0_mt(const int32_t *a, const int64_t *b, const double *c, uint64_t n, int32_t a_lt,
                                   double c_gt, int64_t *out1, double *out2, int threads);
uint64_t vo_join_sum_count_mt(const int64_t *build_key, const int32_t *payload, uint64_t nb, const int64_t *fk,
                              uint64_t np, int threads, uint64_t *sum_lo, int64_t *sum_hi);
uint64_t vo_cross_filter_pairs_mt(const int32_t *x, uint64_t nl, const int32_t *y, uint64_t nr, int64_t c,
                                  int32_t *out_x, int32_t *out_y, uint64_t cap, int threads);
uint64_t vo_batched_filter_group_by_mt(const int32_t *a, const double *c, const int64_t *key, const int64_t *v1,
                                       const double *v2, const int64_t *offsets, uint64_t nq, int32_t a_lt,
                                       double c_gt, int threads, int64_t *out_key, int64_t *out_count,
                                       int64_t *out_sum_v1, int64_t *out_min_v1, int64_t *out_max_v1,
                                       double *out_sum_v2, double *out_avg_v2, uint64_t *group_offsets);

#ifdef __cplusplus
}
#endif

#endif
