/* oracle.c -- CPU restatement of the engine's hot path. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Built with gcc -O2 -ffp-contract=off so every double operation is one IEEE rounding, as DuckDB's CPU
 * path computes it. Single-threaded on purpose: when bench.py times it as the CPU baseline it reports
 * cores = 1.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;
typedef __int128 i128;

/* ------------------------------------------------------------------ generator
 * Same counter mixer as the device generator (vectisdb_amd/csrc/hip/datagen.hip): splitmix64's finalizer
 * over (seed ^ stream * C1) + (i + 1) * golden ratio. */
uint64_t vo_mix_counter(uint64_t seed, uint64_t stream, uint64_t i) {
	uint64_t z = (seed ^ (stream * 0xD1B54A32D192ED03ull)) + (i + 1) * 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

static int64_t uniform_int(uint64_t r, int64_t lo, int64_t hi) {
	uint64_t range = (uint64_t)(hi - lo) + 1ull;
	uint64_t off = range == 0 ? r : (uint64_t)(((u128)r * range) >> 64);
	return (int64_t)((uint64_t)lo + off);
}

void vo_gen_uniform_i64(int64_t *out, uint64_t n, int64_t lo, int64_t hi, uint64_t seed, uint64_t stream) {
	for (uint64_t i = 0; i < n; i++) {
		out[i] = uniform_int(vo_mix_counter(seed, stream, i), lo, hi);
	}
}

void vo_gen_uniform_i32(int32_t *out, uint64_t n, int64_t lo, int64_t hi, uint64_t seed, uint64_t stream) {
	for (uint64_t i = 0; i < n; i++) {
		out[i] = (int32_t)uniform_int(vo_mix_counter(seed, stream, i), lo, hi);
	}
}

void vo_gen_uniform_f64(double *out, uint64_t n, uint64_t seed, uint64_t stream) {
	for (uint64_t i = 0; i < n; i++) {
		out[i] = (double)(vo_mix_counter(seed, stream, i) >> 11) * 0x1.0p-53;
	}
}

/* ------------------------------------------------------------------ config 0: fused filter/project
 * DuckDB types `a*2` as INTEGER (both operands INTEGER), then widens to BIGINT for `+ b`; `c*1.5 + a`
 * is DOUBLE * DOUBLE + CAST(a AS DOUBLE). A row passes when both conjuncts are true. */
uint64_t vo_filter_project_cfg0(const int32_t *a, const int64_t *b, const double *c, uint64_t n, int32_t a_lt,
                                double c_gt, int64_t *out1, double *out2) {
	uint64_t k = 0;
	for (uint64_t i = 0; i < n; i++) {
		if (a[i] < a_lt && c[i] > c_gt) {
			int32_t twice = (int32_t)(a[i] * 2);
			out1[k] = (int64_t)twice + b[i];
			double prod = c[i] * 1.5;
			out2[k] = prod + (double)a[i];
			k++;
		}
	}
	return k;
}

/* ------------------------------------------------------------------ exact double summation
 * Shewchuk's non-overlapping partials (the algorithm behind Python's math.fsum): the partials sum exactly
 * to the running total, and the final pass rounds it once, correctly (half-even). */
typedef struct {
	double *p;
	size_t n, cap;
	double special; /* inf/nan accumulation */
	int has_special;
} msum_t;

static void msum_init(msum_t *m) {
	m->n = 0;
	m->has_special = 0;
	m->special = 0.0;
	if (m->cap == 0) {
		m->cap = 32;
		m->p = (double *)malloc(m->cap * sizeof(double));
	}
}

static void msum_add(msum_t *m, double x) {
	if (!isfinite(x)) {
		m->special = m->has_special ? m->special + x : x;
		m->has_special = 1;
		return;
	}
	size_t i = 0;
	for (size_t j = 0; j < m->n; j++) {
		double y = m->p[j];
		if (fabs(x) < fabs(y)) {
			double t = x;
			x = y;
			y = t;
		}
		double hi = x + y;
		double lo = y - (hi - x);
		if (lo != 0.0) {
			m->p[i++] = lo;
		}
		x = hi;
	}
	if (i + 1 > m->cap) {
		m->cap *= 2;
		m->p = (double *)realloc(m->p, m->cap * sizeof(double));
	}
	m->p[i++] = x;
	m->n = i;
}

static double msum_value(const msum_t *m) {
	if (m->has_special) {
		return m->special;
	}
	if (m->n == 0) {
		return 0.0;
	}
	size_t n = m->n;
	double hi = m->p[--n];
	double lo = 0.0;
	while (n > 0) {
		double x = hi;
		double y = m->p[--n];
		hi = x + y;
		double yr = hi - x;
		lo = y - yr;
		if (lo != 0.0) {
			break;
		}
	}
	/* half-even correction across the next partial */
	if (n > 0 && ((lo < 0.0 && m->p[n - 1] < 0.0) || (lo > 0.0 && m->p[n - 1] > 0.0))) {
		double y = lo * 2.0;
		double x = hi + y;
		double yr = x - hi;
		if (y == yr) {
			hi = x;
		}
	}
	return hi;
}

double vo_exact_sum(const double *x, uint64_t n) {
	msum_t m = {0};
	msum_init(&m);
	for (uint64_t i = 0; i < n; i++) {
		msum_add(&m, x[i]);
	}
	double v = msum_value(&m);
	free(m.p);
	return v;
}

/* ------------------------------------------------------------------ GROUP BY (config 1)
 * Rows are ordered by key with a counting sort when the key range is small, else by qsort on (key, row);
 * each run of equal keys is one group. */
typedef struct {
	int64_t key;
	uint64_t row;
} kr_t;

static int cmp_kr(const void *pa, const void *pb) {
	const kr_t *a = (const kr_t *)pa, *b = (const kr_t *)pb;
	if (a->key != b->key) {
		return a->key < b->key ? -1 : 1;
	}
	return a->row < b->row ? -1 : (a->row > b->row);
}

static uint64_t *order_by_key(const int64_t *key, uint64_t n) {
	uint64_t *order = (uint64_t *)malloc((n ? n : 1) * sizeof(uint64_t));
	if (n == 0) {
		return order;
	}
	int64_t mn = key[0], mx = key[0];
	for (uint64_t i = 1; i < n; i++) {
		if (key[i] < mn) {
			mn = key[i];
		}
		if (key[i] > mx) {
			mx = key[i];
		}
	}
	uint64_t range = (uint64_t)mx - (uint64_t)mn;
	if (range < (1ull << 26)) {
		uint64_t *count = (uint64_t *)calloc(range + 2, sizeof(uint64_t));
		for (uint64_t i = 0; i < n; i++) {
			count[(uint64_t)key[i] - (uint64_t)mn + 1]++;
		}
		for (uint64_t k = 1; k <= range + 1; k++) {
			count[k] += count[k - 1];
		}
		for (uint64_t i = 0; i < n; i++) {
			order[count[(uint64_t)key[i] - (uint64_t)mn]++] = i;
		}
		free(count);
	} else {
		kr_t *kr = (kr_t *)malloc(n * sizeof(kr_t));
		for (uint64_t i = 0; i < n; i++) {
			kr[i].key = key[i];
			kr[i].row = i;
		}
		qsort(kr, n, sizeof(kr_t), cmp_kr);
		for (uint64_t i = 0; i < n; i++) {
			order[i] = kr[i].row;
		}
		free(kr);
	}
	return order;
}

uint64_t vo_group_by_i64(const int64_t *key, const int64_t *v1, const double *v2, uint64_t n, int64_t *out_key,
                         int64_t *out_count, int64_t *out_sum_v1, int64_t *out_min_v1, int64_t *out_max_v1,
                         double *out_sum_v2, double *out_avg_v2) {
	uint64_t *order = order_by_key(key, n);
	msum_t m = {0};
	uint64_t g = 0;
	uint64_t i = 0;
	while (i < n) {
		int64_t k = key[order[i]];
		uint64_t j = i;
		i128 s1 = 0;
		int64_t mn = INT64_MAX, mx = INT64_MIN;
		msum_init(&m);
		while (j < n && key[order[j]] == k) {
			uint64_t r = order[j];
			if (v1) {
				s1 += (i128)v1[r];
				if (v1[r] < mn) {
					mn = v1[r];
				}
				if (v1[r] > mx) {
					mx = v1[r];
				}
			}
			if (v2) {
				msum_add(&m, v2[r]);
			}
			j++;
		}
		uint64_t cnt = j - i;
		if (out_key) {
			out_key[g] = k;
		}
		if (out_count) {
			out_count[g] = (int64_t)cnt;
		}
		if (out_sum_v1) {
			out_sum_v1[2 * g] = (int64_t)(uint64_t)(u128)s1;
			out_sum_v1[2 * g + 1] = (int64_t)(s1 >> 64);
		}
		if (out_min_v1) {
			out_min_v1[g] = mn;
		}
		if (out_max_v1) {
			out_max_v1[g] = mx;
		}
		double s2 = v2 ? msum_value(&m) : 0.0;
		if (out_sum_v2) {
			out_sum_v2[g] = s2;
		}
		if (out_avg_v2) {
			out_avg_v2[g] = s2 / (double)cnt;
		}
		g++;
		i = j;
	}
	free(m.p);
	free(order);
	return g;
}

/* ------------------------------------------------------------------ join (config 2a) */
static int cmp_i64(const void *pa, const void *pb) {
	int64_t a = *(const int64_t *)pa, b = *(const int64_t *)pb;
	return a < b ? -1 : (a > b);
}

uint64_t vo_join_sum_count(const int64_t *build_key, const int32_t *payload, uint64_t nb, const int64_t *fk,
                           uint64_t np, uint64_t *sum_lo, int64_t *sum_hi) {
	/* Sort the build side as (key, payload) pairs; per probe row, binary-search the key's run. */
	uint64_t *order = order_by_key(build_key, nb);
	int64_t *keys = (int64_t *)malloc((nb ? nb : 1) * sizeof(int64_t));
	int64_t *run_sum = (int64_t *)malloc((nb ? nb : 1) * sizeof(int64_t)); /* prefix sums of payload */
	int64_t acc = 0;
	for (uint64_t i = 0; i < nb; i++) {
		keys[i] = build_key[order[i]];
		acc += payload[order[i]];
		run_sum[i] = acc;
	}
	i128 sum = 0;
	uint64_t count = 0;
	for (uint64_t p = 0; p < np; p++) {
		int64_t k = fk[p];
		uint64_t lo = 0, hi = nb;
		while (lo < hi) {
			uint64_t mid = (lo + hi) >> 1;
			if (keys[mid] < k) {
				lo = mid + 1;
			} else {
				hi = mid;
			}
		}
		uint64_t a = lo;
		hi = nb;
		while (lo < hi) {
			uint64_t mid = (lo + hi) >> 1;
			if (keys[mid] <= k) {
				lo = mid + 1;
			} else {
				hi = mid;
			}
		}
		uint64_t b = lo;
		if (b > a) {
			count += b - a;
			sum += (i128)(run_sum[b - 1] - (a ? run_sum[a - 1] : 0));
		}
	}
	free(order);
	free(keys);
	free(run_sum);
	*sum_lo = (uint64_t)(u128)sum;
	*sum_hi = (int64_t)(sum >> 64);
	return count;
}

uint64_t vo_join_pairs(const int64_t *left, uint64_t nl, const int64_t *right, uint64_t nr, uint64_t *out_l,
                       uint64_t *out_r, uint64_t cap) {
	uint64_t *order = order_by_key(right, nr);
	int64_t *keys = (int64_t *)malloc((nr ? nr : 1) * sizeof(int64_t));
	for (uint64_t i = 0; i < nr; i++) {
		keys[i] = right[order[i]];
	}
	uint64_t k = 0;
	for (uint64_t l = 0; l < nl; l++) {
		uint64_t lo = 0, hi = nr;
		while (lo < hi) {
			uint64_t mid = (lo + hi) >> 1;
			if (keys[mid] < left[l]) {
				lo = mid + 1;
			} else {
				hi = mid;
			}
		}
		/* counting sort is stable, so equal keys appear in right-row order */
		for (uint64_t j = lo; j < nr && keys[j] == left[l]; j++) {
			if (k < cap) {
				out_l[k] = l;
				out_r[k] = order[j];
			}
			k++;
		}
	}
	free(order);
	free(keys);
	return k;
}

/* ------------------------------------------------------------------ cross product + filter (config 2b) */
uint64_t vo_cross_filter_count(const int32_t *x, uint64_t nl, const int32_t *y, uint64_t nr, int64_t c,
                               int64_t *sum_x, int64_t *sum_y) {
	int64_t *ys = (int64_t *)malloc((nr ? nr : 1) * sizeof(int64_t));
	int64_t *prefix = (int64_t *)malloc((nr + 1) * sizeof(int64_t));
	for (uint64_t j = 0; j < nr; j++) {
		ys[j] = y[j];
	}
	qsort(ys, nr, sizeof(int64_t), cmp_i64);
	prefix[0] = 0;
	for (uint64_t j = 0; j < nr; j++) {
		prefix[j + 1] = prefix[j] + ys[j];
	}
	uint64_t count = 0;
	int64_t sx = 0, sy = 0;
	for (uint64_t i = 0; i < nl; i++) {
		int64_t bound = c - (int64_t)x[i]; /* y < bound */
		uint64_t lo = 0, hi = nr;
		while (lo < hi) {
			uint64_t mid = (lo + hi) >> 1;
			if (ys[mid] < bound) {
				lo = mid + 1;
			} else {
				hi = mid;
			}
		}
		count += lo;
		sx += (int64_t)lo * x[i];
		sy += prefix[lo];
	}
	free(ys);
	free(prefix);
	if (sum_x) {
		*sum_x = sx;
	}
	if (sum_y) {
		*sum_y = sy;
	}
	return count;
}

uint64_t vo_cross_filter_pairs(const int32_t *x, uint64_t nl, const int32_t *y, uint64_t nr, int64_t c,
                               int32_t *out_x, int32_t *out_y, uint64_t cap) {
	uint64_t k = 0;
	for (uint64_t i = 0; i < nl; i++) {
		for (uint64_t j = 0; j < nr; j++) {
			if ((int64_t)x[i] + (int64_t)y[j] < c) {
				if (k < cap) {
					out_x[k] = x[i];
					out_y[k] = y[j];
				}
				k++;
			}
		}
	}
	return k;
}
