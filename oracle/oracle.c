/* oracle.c -- CPU restatement of the engine's hot path. TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Built with gcc -O2 -ffp-contract=off so every double operation is one IEEE rounding, as DuckDB's CPU
 * path computes it. Every function is single-threaded except vo_group_by_i64_mt, the multithreaded GROUP
 * BY the benchmark times as its CPU baseline (same per-group arithmetic, so the same results).
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
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

/* ------------------------------------------------------------------ multithreaded GROUP BY
 * Stable counting sort by key with one histogram per thread (thread t owns rows [t*n/T, (t+1)*n/T), so
 * bucket b's rows land thread after thread, in row order), then every group reduced exactly as in
 * vo_group_by_i64 by the thread that owns its bucket block. Falls back to vo_group_by_i64 when the key
 * range is wider than 2^24. */
typedef struct {
	const int64_t *key, *v1;
	const double *v2;
	uint64_t n, range;
	int64_t mn;
	int threads;
	uint32_t *hist;  /* threads x range */
	uint64_t *order;
	uint64_t *gfirst; /* range + 1: group index of each bucket (prefix of non-empty buckets) */
	uint64_t *bstart; /* range + 1: first position of each bucket in order */
	int64_t *out_key, *out_count, *out_sum_v1, *out_min_v1, *out_max_v1;
	double *out_sum_v2, *out_avg_v2;
	int phase;
	/* phase 3 work list: a task is a run of whole buckets [b0, b1), or rows [r0, r1) of one large bucket
	 * whose partial state goes to slot `slot` (merged afterwards) */
	uint64_t ntasks;
	uint64_t *task_b0, *task_b1, *task_r0, *task_r1;
	int64_t *task_slot;
	uint64_t next_task; /* atomic */
	i128 *part_s1;
	int64_t *part_mn, *part_mx;
	msum_t *part_m;
} mt_ctx;

typedef struct {
	mt_ctx *c;
	int t;
	int64_t mn, mx;
} mt_arg;

static void *mt_worker(void *p) {
	mt_arg *a = (mt_arg *)p;
	mt_ctx *c = a->c;
	const uint64_t lo = c->n * (uint64_t)a->t / (uint64_t)c->threads;
	const uint64_t hi = c->n * (uint64_t)(a->t + 1) / (uint64_t)c->threads;
	if (c->phase == 0) { /* min / max of the chunk */
		int64_t mn = INT64_MAX, mx = INT64_MIN;
		for (uint64_t i = lo; i < hi; i++) {
			mn = c->key[i] < mn ? c->key[i] : mn;
			mx = c->key[i] > mx ? c->key[i] : mx;
		}
		a->mn = mn;
		a->mx = mx;
	} else if (c->phase == 1) { /* histogram */
		uint32_t *h = c->hist + (uint64_t)a->t * c->range;
		for (uint64_t i = lo; i < hi; i++) {
			h[(uint64_t)c->key[i] - (uint64_t)c->mn]++;
		}
	} else if (c->phase == 2) { /* scatter: hist now holds this thread's first position per bucket */
		uint32_t *h = c->hist + (uint64_t)a->t * c->range;
		for (uint64_t i = lo; i < hi; i++) {
			c->order[h[(uint64_t)c->key[i] - (uint64_t)c->mn]++] = i;
		}
	} else { /* reduce: take tasks until none is left */
		msum_t m = {0};
		for (;;) {
			const uint64_t k = __atomic_fetch_add(&c->next_task, 1, __ATOMIC_RELAXED);
			if (k >= c->ntasks) {
				break;
			}
			if (c->task_slot[k] >= 0) { /* part of a large group */
				const int64_t sl = c->task_slot[k];
				i128 s1 = 0;
				int64_t mn = INT64_MAX, mx = INT64_MIN;
				msum_t *pm = &c->part_m[sl];
				msum_init(pm);
				for (uint64_t j = c->task_r0[k]; j < c->task_r1[k]; j++) {
					const uint64_t r = c->order[j];
					if (c->v1) {
						s1 += (i128)c->v1[r];
						mn = c->v1[r] < mn ? c->v1[r] : mn;
						mx = c->v1[r] > mx ? c->v1[r] : mx;
					}
					if (c->v2) {
						msum_add(pm, c->v2[r]);
					}
				}
				c->part_s1[sl] = s1;
				c->part_mn[sl] = mn;
				c->part_mx[sl] = mx;
				continue;
			}
		for (uint64_t b = c->task_b0[k]; b < c->task_b1[k]; b++) {
			const uint64_t i0 = c->bstart[b], i1 = c->bstart[b + 1];
			if (i0 == i1) {
				continue;
			}
			const uint64_t g = c->gfirst[b];
			i128 s1 = 0;
			int64_t mn = INT64_MAX, mx = INT64_MIN;
			msum_init(&m);
			for (uint64_t j = i0; j < i1; j++) {
				const uint64_t r = c->order[j];
				if (c->v1) {
					s1 += (i128)c->v1[r];
					mn = c->v1[r] < mn ? c->v1[r] : mn;
					mx = c->v1[r] > mx ? c->v1[r] : mx;
				}
				if (c->v2) {
					msum_add(&m, c->v2[r]);
				}
			}
			const uint64_t cnt = i1 - i0;
			if (c->out_key) {
				c->out_key[g] = (int64_t)((uint64_t)c->mn + b);
			}
			if (c->out_count) {
				c->out_count[g] = (int64_t)cnt;
			}
			if (c->out_sum_v1) {
				c->out_sum_v1[2 * g] = (int64_t)(uint64_t)(u128)s1;
				c->out_sum_v1[2 * g + 1] = (int64_t)(s1 >> 64);
			}
			if (c->out_min_v1) {
				c->out_min_v1[g] = mn;
			}
			if (c->out_max_v1) {
				c->out_max_v1[g] = mx;
			}
			const double s2 = c->v2 ? msum_value(&m) : 0.0;
			if (c->out_sum_v2) {
				c->out_sum_v2[g] = s2;
			}
			if (c->out_avg_v2) {
				c->out_avg_v2[g] = s2 / (double)cnt;
			}
		}
		}
		free(m.p);
	}
	return NULL;
}

static uint64_t mn_key_base(int64_t mn) {
	return (uint64_t)mn;
}

static void mt_run(mt_ctx *c, mt_arg *args, pthread_t *tid, int phase) {
	c->phase = phase;
	for (int t = 0; t < c->threads; t++) {
		pthread_create(&tid[t], NULL, mt_worker, &args[t]);
	}
	for (int t = 0; t < c->threads; t++) {
		pthread_join(tid[t], NULL);
	}
}

uint64_t vo_group_by_i64_mt(const int64_t *key, const int64_t *v1, const double *v2, uint64_t n, int threads,
                            int64_t *out_key, int64_t *out_count, int64_t *out_sum_v1, int64_t *out_min_v1,
                            int64_t *out_max_v1, double *out_sum_v2, double *out_avg_v2) {
	if (threads < 1) {
		threads = 1;
	}
	if (threads > 256) {
		threads = 256;
	}
	if (n == 0 || threads == 1 || n >= (1ull << 32)) {
		return vo_group_by_i64(key, v1, v2, n, out_key, out_count, out_sum_v1, out_min_v1, out_max_v1, out_sum_v2,
		                       out_avg_v2);
	}
	mt_ctx c;
	memset(&c, 0, sizeof(c));
	c.key = key;
	c.v1 = v1;
	c.v2 = v2;
	c.n = n;
	c.threads = threads;
	mt_arg args[256];
	pthread_t tid[256];
	for (int t = 0; t < threads; t++) {
		args[t].c = &c;
		args[t].t = t;
	}
	mt_run(&c, args, tid, 0);
	int64_t mn = INT64_MAX, mx = INT64_MIN;
	for (int t = 0; t < threads; t++) {
		mn = args[t].mn < mn ? args[t].mn : mn;
		mx = args[t].mx > mx ? args[t].mx : mx;
	}
	const uint64_t range = (uint64_t)mx - (uint64_t)mn + 1;
	if (range > (1ull << 24)) {
		return vo_group_by_i64(key, v1, v2, n, out_key, out_count, out_sum_v1, out_min_v1, out_max_v1, out_sum_v2,
		                       out_avg_v2);
	}
	c.mn = mn;
	c.range = range;
	c.hist = (uint32_t *)calloc((size_t)threads * range, sizeof(uint32_t));
	c.order = (uint64_t *)malloc(n * sizeof(uint64_t));
	c.gfirst = (uint64_t *)malloc((range + 1) * sizeof(uint64_t));
	c.bstart = (uint64_t *)malloc((range + 1) * sizeof(uint64_t));
	mt_run(&c, args, tid, 1);
	/* bucket-major, thread-minor exclusive prefix: each thread's first slot per bucket */
	uint64_t pos = 0, groups = 0;
	for (uint64_t b = 0; b < range; b++) {
		c.bstart[b] = pos;
		c.gfirst[b] = groups;
		const uint64_t before = pos;
		for (int t = 0; t < threads; t++) {
			uint32_t *h = c.hist + (uint64_t)t * range + b;
			const uint32_t cnt = *h;
			*h = (uint32_t)pos; /* positions stay below n < 2^32 */
			pos += cnt;
		}
		groups += pos > before ? 1 : 0;
	}
	c.bstart[range] = pos;
	c.gfirst[range] = groups;
	mt_run(&c, args, tid, 2);
	c.out_key = out_key;
	c.out_count = out_count;
	c.out_sum_v1 = out_sum_v1;
	c.out_min_v1 = out_min_v1;
	c.out_max_v1 = out_max_v1;
	c.out_sum_v2 = out_sum_v2;
	c.out_avg_v2 = out_avg_v2;
	/* Tasks of about `chunk` rows: runs of small buckets, and large buckets cut into chunk-sized parts. */
	const uint64_t chunk = n / ((uint64_t)threads * 16) + 1;
	const uint64_t cap = range + n / chunk + 2;
	c.task_b0 = (uint64_t *)malloc(cap * sizeof(uint64_t));
	c.task_b1 = (uint64_t *)malloc(cap * sizeof(uint64_t));
	c.task_r0 = (uint64_t *)malloc(cap * sizeof(uint64_t));
	c.task_r1 = (uint64_t *)malloc(cap * sizeof(uint64_t));
	c.task_slot = (int64_t *)malloc(cap * sizeof(int64_t));
	uint64_t nt = 0, nslots = 0, run_b0 = 0, run_rows = 0;
	for (uint64_t b = 0; b <= range; b++) {
		const uint64_t cnt = b < range ? c.bstart[b + 1] - c.bstart[b] : 0;
		if (b == range || cnt > chunk || run_rows >= chunk) {
			if (b > run_b0) { /* close the run of small buckets */
				c.task_b0[nt] = run_b0;
				c.task_b1[nt] = b;
				c.task_slot[nt++] = -1;
			}
			run_b0 = b;
			run_rows = 0;
		}
		if (b == range) {
			break;
		}
		if (cnt > chunk) {
			for (uint64_t r = c.bstart[b]; r < c.bstart[b + 1]; r += chunk) {
				c.task_b0[nt] = b;
				c.task_b1[nt] = b + 1;
				c.task_r0[nt] = r;
				c.task_r1[nt] = r + chunk < c.bstart[b + 1] ? r + chunk : c.bstart[b + 1];
				c.task_slot[nt++] = (int64_t)nslots++;
			}
			run_b0 = b + 1;
		} else {
			run_rows += cnt;
		}
	}
	c.ntasks = nt;
	c.next_task = 0;
	c.part_s1 = (i128 *)malloc((nslots + 1) * sizeof(i128));
	c.part_mn = (int64_t *)malloc((nslots + 1) * sizeof(int64_t));
	c.part_mx = (int64_t *)malloc((nslots + 1) * sizeof(int64_t));
	c.part_m = (msum_t *)calloc(nslots + 1, sizeof(msum_t));
	mt_run(&c, args, tid, 3);
	/* Merge the parts of each large group (exact: int128 adds, min/max, expansion components re-added). */
	for (uint64_t k = 0; k < nt;) {
		if (c.task_slot[k] < 0) {
			k++;
			continue;
		}
		const uint64_t b = c.task_b0[k];
		i128 s1 = 0;
		int64_t mn = INT64_MAX, mx = INT64_MIN;
		msum_t m = {0};
		msum_init(&m);
		for (; k < nt && c.task_slot[k] >= 0 && c.task_b0[k] == b; k++) {
			const int64_t sl = c.task_slot[k];
			s1 += c.part_s1[sl];
			mn = c.part_mn[sl] < mn ? c.part_mn[sl] : mn;
			mx = c.part_mx[sl] > mx ? c.part_mx[sl] : mx;
			const msum_t *pm = &c.part_m[sl];
			for (size_t j = 0; j < pm->n; j++) {
				msum_add(&m, pm->p[j]);
			}
			if (pm->has_special) {
				msum_add(&m, pm->special);
			}
			free(pm->p);
		}
		const uint64_t g = c.gfirst[b];
		const uint64_t cnt = c.bstart[b + 1] - c.bstart[b];
		if (out_key) {
			out_key[g] = (int64_t)((uint64_t)mn_key_base(c.mn) + b);
		}
		if (out_count) {
			out_count[g] = (int64_t)cnt;
		}
		if (out_sum_v1) {
			out_sum_v1[2 * g] = (int64_t)(uint64_t)(u128)s1;
			out_sum_v1[2 * g + 1] = (int64_t)(s1 >> 64);
		}
		if (out_min_v1) {
			out_min_v1[g] = v1 ? mn : INT64_MAX;
		}
		if (out_max_v1) {
			out_max_v1[g] = v1 ? mx : INT64_MIN;
		}
		const double s2 = v2 ? msum_value(&m) : 0.0;
		if (out_sum_v2) {
			out_sum_v2[g] = s2;
		}
		if (out_avg_v2) {
			out_avg_v2[g] = s2 / (double)cnt;
		}
		free(m.p);
	}
	free(c.task_b0);
	free(c.task_b1);
	free(c.task_r0);
	free(c.task_r1);
	free(c.task_slot);
	free(c.part_s1);
	free(c.part_mn);
	free(c.part_mx);
	free(c.part_m);
	free(c.hist);
	free(c.order);
	free(c.gfirst);
	free(c.bstart);
	return groups;
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

/* ------------------------------------------------------------------ multithreaded CPU baselines
 * The remaining configurations restated for all host threads, so each benchmark line can report a CPU
 * figure with a thread count. Each returns exactly what its single-threaded counterpart above returns
 * (tests/test_oracle.py checks them against each other); only the work is split. */

typedef struct {
	void (*fn)(void *ctx, int t, int threads);
	void *ctx;
	int t, threads;
} par_arg;

static void *par_trampoline(void *p) {
	par_arg *a = (par_arg *)p;
	a->fn(a->ctx, a->t, a->threads);
	return NULL;
}

/* Runs fn(ctx, t, threads) for t = 0..threads-1 on threads of their own, and waits for all of them. */
static void par_for(int threads, void (*fn)(void *, int, int), void *ctx) {
	pthread_t tid[256];
	par_arg args[256];
	if (threads < 1) {
		threads = 1;
	}
	if (threads > 256) {
		threads = 256;
	}
	for (int t = 0; t < threads; t++) {
		args[t].fn = fn;
		args[t].ctx = ctx;
		args[t].t = t;
		args[t].threads = threads;
		pthread_create(&tid[t], NULL, par_trampoline, &args[t]);
	}
	for (int t = 0; t < threads; t++) {
		pthread_join(tid[t], NULL);
	}
}

static int clamp_threads(int threads) {
	return threads < 1 ? 1 : (threads > 256 ? 256 : threads);
}

/* --- config 0 on all threads: each thread counts its row range, a prefix over the counts gives every range its
 * output offset, and each thread then writes its kept rows in row order. */
typedef struct {
	const int32_t *a;
	const int64_t *b;
	const double *c;
	uint64_t n;
	int32_t a_lt;
	double c_gt;
	int64_t *out1;
	double *out2;
	uint64_t count[256];
	uint64_t offset[256];
	int phase;
} fp_ctx;

static void fp_task(void *p, int t, int threads) {
	fp_ctx *x = (fp_ctx *)p;
	const uint64_t lo = x->n * (uint64_t)t / (uint64_t)threads, hi = x->n * (uint64_t)(t + 1) / (uint64_t)threads;
	if (x->phase == 0) {
		uint64_t k = 0;
		for (uint64_t i = lo; i < hi; i++) {
			k += (x->a[i] < x->a_lt) & (x->c[i] > x->c_gt);
		}
		x->count[t] = k;
	} else {
		uint64_t k = x->offset[t];
		for (uint64_t i = lo; i < hi; i++) {
			if (x->a[i] < x->a_lt && x->c[i] > x->c_gt) {
				const int32_t twice = (int32_t)(x->a[i] * 2);
				x->out1[k] = (int64_t)twice + x->b[i];
				x->out2[k] = x->c[i] * 1.5 + (double)x->a[i];
				k++;
			}
		}
	}
}

uint64_t vo_filter_project_cfg0_mt(const int32_t *a, const int64_t *b, const double *c, uint64_t n, int32_t a_lt,
                                   double c_gt, int64_t *out1, double *out2, int threads) {
	fp_ctx *x = (fp_ctx *)calloc(1, sizeof(fp_ctx));
	threads = clamp_threads(threads);
	x->a = a;
	x->b = b;
	x->c = c;
	x->n = n;
	x->a_lt = a_lt;
	x->c_gt = c_gt;
	x->out1 = out1;
	x->out2 = out2;
	x->phase = 0;
	par_for(threads, fp_task, x);
	uint64_t total = 0;
	for (int t = 0; t < threads; t++) {
		x->offset[t] = total;
		total += x->count[t];
	}
	x->phase = 1;
	par_for(threads, fp_task, x);
	free(x);
	return total;
}

/* --- config 2a on all threads: a hash join. The build side goes into an open-addressing table (linear probing,
 * multiplicative hash, twice the build rows rounded up to a power of two) that all threads fill at once: a
 * slot is claimed by a compare-and-swap of its state word, and the key's row count and payload sum are
 * added with atomics. Each thread then probes its range of foreign keys and sums its matches exactly. */
typedef struct {
	int64_t key;
	uint64_t state; /* 0 empty, 1 being claimed, 2 holds key */
	uint64_t count;
	int64_t sum;
} hj_slot;

typedef struct {
	const int64_t *bk;
	const int32_t *pay;
	uint64_t nb;
	const int64_t *fk;
	uint64_t np;
	hj_slot *tab;
	uint64_t mask;
	int shift;
	int phase;
	uint64_t count[256];
	i128 sum[256];
} hj_ctx;

static uint64_t hj_hash(int64_t k, int shift) {
	return ((uint64_t)k * 0x9E3779B97F4A7C15ull) >> shift;
}

static void hj_task(void *p, int t, int threads) {
	hj_ctx *x = (hj_ctx *)p;
	if (x->phase == 0) {
		const uint64_t lo = x->nb * (uint64_t)t / (uint64_t)threads, hi = x->nb * (uint64_t)(t + 1) / (uint64_t)threads;
		for (uint64_t i = lo; i < hi; i++) {
			const int64_t k = x->bk[i];
			for (uint64_t h = hj_hash(k, x->shift);; h = (h + 1) & x->mask) {
				hj_slot *s = &x->tab[h];
				uint64_t st = __atomic_load_n(&s->state, __ATOMIC_ACQUIRE);
				if (st == 0) {
					uint64_t zero = 0;
					if (__atomic_compare_exchange_n(&s->state, &zero, 1, 0, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE)) {
						s->key = k;
						__atomic_store_n(&s->state, 2, __ATOMIC_RELEASE);
						st = 2;
					} else {
						st = zero;
					}
				}
				while (st == 1) {
					st = __atomic_load_n(&s->state, __ATOMIC_ACQUIRE);
				}
				if (s->key == k) {
					__atomic_fetch_add(&s->count, 1, __ATOMIC_RELAXED);
					__atomic_fetch_add(&s->sum, (int64_t)x->pay[i], __ATOMIC_RELAXED);
					break;
				}
			}
		}
	} else {
		const uint64_t lo = x->np * (uint64_t)t / (uint64_t)threads, hi = x->np * (uint64_t)(t + 1) / (uint64_t)threads;
		uint64_t cnt = 0;
		i128 sum = 0;
		for (uint64_t i = lo; i < hi; i++) {
			const int64_t k = x->fk[i];
			for (uint64_t h = hj_hash(k, x->shift);; h = (h + 1) & x->mask) {
				const hj_slot *s = &x->tab[h];
				if (s->state == 0) {
					break;
				}
				if (s->key == k) {
					cnt += s->count;
					sum += (i128)s->sum;
					break;
				}
			}
		}
		x->count[t] = cnt;
		x->sum[t] = sum;
	}
}

uint64_t vo_join_sum_count_mt(const int64_t *build_key, const int32_t *payload, uint64_t nb, const int64_t *fk,
                              uint64_t np, int threads, uint64_t *sum_lo, int64_t *sum_hi) {
	hj_ctx *x = (hj_ctx *)calloc(1, sizeof(hj_ctx));
	threads = clamp_threads(threads);
	int bits = 1;
	while ((1ull << bits) < 2 * (nb ? nb : 1)) {
		bits++;
	}
	x->tab = (hj_slot *)calloc(1ull << bits, sizeof(hj_slot));
	x->mask = (1ull << bits) - 1;
	x->shift = 64 - bits;
	x->bk = build_key;
	x->pay = payload;
	x->nb = nb;
	x->fk = fk;
	x->np = np;
	x->phase = 0;
	par_for(threads, hj_task, x);
	x->phase = 1;
	par_for(threads, hj_task, x);
	uint64_t count = 0;
	i128 sum = 0;
	for (int t = 0; t < threads; t++) {
		count += x->count[t];
		sum += x->sum[t];
	}
	free(x->tab);
	free(x);
	*sum_lo = (uint64_t)(u128)sum;
	*sum_hi = (int64_t)(sum >> 64);
	return count;
}

/* --- config 2b on all threads: the materialized cross product + filter in left-major order. Each left row's
 * output count comes from a binary search over the sorted right side; a prefix over those counts gives each
 * left row its output offset, and each thread writes its left rows' pairs, scanning the right side in row
 * order. */
typedef struct {
	const int32_t *x, *y;
	uint64_t nl, nr;
	int64_t c;
	const int64_t *ys;
	uint64_t *offset;
	int32_t *ox, *oy;
	int phase;
} cf_ctx;

static void cf_task(void *p, int t, int threads) {
	cf_ctx *q = (cf_ctx *)p;
	const uint64_t lo = q->nl * (uint64_t)t / (uint64_t)threads, hi = q->nl * (uint64_t)(t + 1) / (uint64_t)threads;
	for (uint64_t i = lo; i < hi; i++) {
		const int64_t bound = q->c - (int64_t)q->x[i];
		if (q->phase == 0) {
			uint64_t a = 0, b = q->nr;
			while (a < b) {
				const uint64_t mid = (a + b) >> 1;
				if (q->ys[mid] < bound) {
					a = mid + 1;
				} else {
					b = mid;
				}
			}
			q->offset[i + 1] = a;
		} else {
			uint64_t k = q->offset[i];
			const int32_t xi = q->x[i];
			for (uint64_t j = 0; j < q->nr; j++) {
				if ((int64_t)q->y[j] < bound) {
					q->ox[k] = xi;
					q->oy[k] = q->y[j];
					k++;
				}
			}
		}
	}
}

uint64_t vo_cross_filter_pairs_mt(const int32_t *x, uint64_t nl, const int32_t *y, uint64_t nr, int64_t c,
                                  int32_t *out_x, int32_t *out_y, uint64_t cap, int threads) {
	cf_ctx q;
	memset(&q, 0, sizeof(q));
	threads = clamp_threads(threads);
	int64_t *ys = (int64_t *)malloc((nr ? nr : 1) * sizeof(int64_t));
	for (uint64_t j = 0; j < nr; j++) {
		ys[j] = y[j];
	}
	qsort(ys, nr, sizeof(int64_t), cmp_i64);
	q.x = x;
	q.y = y;
	q.nl = nl;
	q.nr = nr;
	q.c = c;
	q.ys = ys;
	q.offset = (uint64_t *)calloc(nl + 1, sizeof(uint64_t));
	q.ox = out_x;
	q.oy = out_y;
	q.phase = 0;
	par_for(threads, cf_task, &q);
	for (uint64_t i = 0; i < nl; i++) {
		q.offset[i + 1] += q.offset[i];
	}
	const uint64_t total = q.offset[nl];
	if (total <= cap) {
		q.phase = 1;
		par_for(threads, cf_task, &q);
	}
	free(ys);
	free(q.offset);
	return total;
}

/* --- config 3 on all threads: the batched queries one after another, as a CPU engine runs them: for query q,
 * rows [offsets[q], offsets[q+1]) are filtered (a < a_lt AND c > c_gt) into compact columns on all threads,
 * then grouped by vo_group_by_i64_mt. Query q's groups land at out rows [group_offsets[q],
 * group_offsets[q+1]) (nq + 1 entries written). Returns the total group count. */
typedef struct {
	const int32_t *a;
	const double *c;
	const int64_t *key, *v1;
	const double *v2;
	uint64_t r0, n;
	int32_t a_lt;
	double c_gt;
	int64_t *ok, *o1;
	double *o2;
	uint64_t count[256];
	uint64_t offset[256];
	int phase;
} bq_ctx;

static void bq_task(void *p, int t, int threads) {
	bq_ctx *x = (bq_ctx *)p;
	const uint64_t lo = x->r0 + x->n * (uint64_t)t / (uint64_t)threads;
	const uint64_t hi = x->r0 + x->n * (uint64_t)(t + 1) / (uint64_t)threads;
	if (x->phase == 0) {
		uint64_t k = 0;
		for (uint64_t i = lo; i < hi; i++) {
			k += (x->a[i] < x->a_lt) & (x->c[i] > x->c_gt);
		}
		x->count[t] = k;
	} else {
		uint64_t k = x->offset[t];
		for (uint64_t i = lo; i < hi; i++) {
			if (x->a[i] < x->a_lt && x->c[i] > x->c_gt) {
				x->ok[k] = x->key[i];
				x->o1[k] = x->v1[i];
				x->o2[k] = x->v2[i];
				k++;
			}
		}
	}
}

uint64_t vo_batched_filter_group_by_mt(const int32_t *a, const double *c, const int64_t *key, const int64_t *v1,
                                       const double *v2, const int64_t *offsets, uint64_t nq, int32_t a_lt,
                                       double c_gt, int threads, int64_t *out_key, int64_t *out_count,
                                       int64_t *out_sum_v1, int64_t *out_min_v1, int64_t *out_max_v1,
                                       double *out_sum_v2, double *out_avg_v2, uint64_t *group_offsets) {
	bq_ctx *x = (bq_ctx *)calloc(1, sizeof(bq_ctx));
	threads = clamp_threads(threads);
	uint64_t maxrows = 1;
	for (uint64_t q = 0; q < nq; q++) {
		const uint64_t r = (uint64_t)(offsets[q + 1] - offsets[q]);
		maxrows = r > maxrows ? r : maxrows;
	}
	x->ok = (int64_t *)malloc(maxrows * sizeof(int64_t));
	x->o1 = (int64_t *)malloc(maxrows * sizeof(int64_t));
	x->o2 = (double *)malloc(maxrows * sizeof(double));
	x->a = a;
	x->c = c;
	x->key = key;
	x->v1 = v1;
	x->v2 = v2;
	x->a_lt = a_lt;
	x->c_gt = c_gt;
	uint64_t g = 0;
	for (uint64_t q = 0; q < nq; q++) {
		group_offsets[q] = g;
		x->r0 = (uint64_t)offsets[q];
		x->n = (uint64_t)(offsets[q + 1] - offsets[q]);
		x->phase = 0;
		par_for(threads, bq_task, x);
		uint64_t kept = 0;
		for (int t = 0; t < threads; t++) {
			x->offset[t] = kept;
			kept += x->count[t];
		}
		x->phase = 1;
		par_for(threads, bq_task, x);
		g += vo_group_by_i64_mt(x->ok, x->o1, x->o2, kept, threads, out_key + g, out_count + g, out_sum_v1 + 2 * g,
		                        out_min_v1 + g, out_max_v1 + g, out_sum_v2 + g, out_avg_v2 + g);
	}
	group_offsets[nq] = g;
	free(x->ok);
	free(x->o1);
	free(x->o2);
	free(x);
	return g;
}
