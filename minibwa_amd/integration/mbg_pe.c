/*
 * mbg_pe.c -- the insert-size model of a mini-batch (mb_pestat, pe.c:68-124) with its per-fragment
 * pass split over threads.
 *
 * mb_pestat looks at every read pair: when both mates have exactly one primary hit with MAPQ >= 10
 * (mb_select_unique_se, pe.c:32-39) on the same contig, the pair's orientation and distance
 * (mb_insert_dir, pe.c:10-17) add the distance to that orientation's list, if below max_pe_ins.  Each
 * list is then sorted, and its quartiles, the mean and standard deviation inside the outlier bounds,
 * and the bounds for proper pairs follow.  The pass over the pairs touches every mate's hit array
 * once, scattered in memory, and takes most of the time; here it runs on several threads, each
 * filling its own lists.  The lists are only ever used sorted, and a sorted list does not depend on
 * the order its values came in, so every statistic is the reference's: the arithmetic after the sort
 * is the reference's, step for step, in the same types.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mbpriv.h"
#include "mbg_pe.h"

#define PE_MAX_THREADS 16

typedef struct {
	const mb_opt_t *opt;
	const int32_t *seg_off, *seg_cnt, *n_hit;
	mb_hit_t *const *hit;
	int32_t lo, hi;
	int64_t n[4], m[4];
	uint64_t *a[4];
} pe_part_t;

/* the one primary hit of a read with MAPQ >= 10, or null (mb_select_unique_se, pe.c:32-39) */
static const mb_hit_t *unique_hit(int32_t n_hit, const mb_hit_t *hit)
{
	int32_t j, n_pri = 0, mapq = 0, k = -1;
	for (j = 0; j < n_hit; ++j)
		if (hit[j].id == hit[j].parent) ++n_pri, mapq = hit[j].mapq, k = j;
	return n_pri == 1 && mapq >= 10 ? &hit[k] : 0;
}

/* orientation class 0-3 and distance of two hits (mb_insert_dir, pe.c:10-17) */
static int pair_dir(const mb_hit_t *h0, const mb_hit_t *h1, int64_t *dist)
{
	int64_t p0 = h0->rev ? h0->te : h0->ts, p1 = h1->rev ? h1->te : h1->ts;
	*dist = p0 > p1 ? p0 - p1 : p1 - p0;
	return ((int32_t)h0->rev << 1 | (int32_t)h1->rev) ^ (p0 < p1 ? 0 : 3);
}

static void *part_main(void *arg)
{
	pe_part_t *p = (pe_part_t*)arg;
	int32_t i;
	for (i = p->lo; i < p->hi; ++i) {
		const mb_hit_t *r0, *r1;
		int32_t off, dir;
		int64_t dist;
		if (p->seg_cnt[i] != 2) continue;
		off = p->seg_off[i];
		r0 = unique_hit(p->n_hit[off], p->hit[off]);
		r1 = unique_hit(p->n_hit[off + 1], p->hit[off + 1]);
		if (r0 == 0 || r1 == 0 || r0->tid != r1->tid) continue;
		dir = pair_dir(r0, r1, &dist);
		if (dist < p->opt->max_pe_ins) {
			if (p->n[dir] == p->m[dir]) {
				p->m[dir] = p->m[dir] ? p->m[dir] * 2 : 1024;
				p->a[dir] = (uint64_t*)realloc(p->a[dir], p->m[dir] * sizeof(uint64_t));
				if (p->a[dir] == 0) abort();
			}
			p->a[dir][p->n[dir]++] = dist;
		}
	}
	return 0;
}

void mbg_pestat(const mb_opt_t *opt, int32_t n_frag, const int32_t *seg_off, const int32_t *seg_cnt, const int32_t *n_hit,
                mb_hit_t *const *hit, mb_pestat_t pes[4], int n_threads)
{
	const int MIN_DIR_CNT = 20;
	const double MIN_DIR_RATIO = 0.05, OUTLIER_BOUND = 2.0, MAPPING_BOUND = 3.0, MAX_STDDEV = 4.0;
	pe_part_t part[PE_MAX_THREADS];
	pthread_t th[PE_MAX_THREADS];
	int started[PE_MAX_THREADS];
	int32_t n[4], i, max;
	uint64_t *a[4];
	int d, k, nt = n_threads < 1 ? 1 : n_threads > PE_MAX_THREADS ? PE_MAX_THREADS : n_threads;
	if (n_frag < 4096) nt = 1;
	memset(part, 0, sizeof(part));
	memset(pes, 0, sizeof(pes[0]) * 4);
	for (k = 0; k < nt; ++k) {
		part[k].opt = opt, part[k].seg_off = seg_off, part[k].seg_cnt = seg_cnt, part[k].n_hit = n_hit, part[k].hit = hit;
		part[k].lo = (int32_t)((int64_t)n_frag * k / nt), part[k].hi = (int32_t)((int64_t)n_frag * (k + 1) / nt);
	}
	for (k = 1; k < nt; ++k) started[k] = pthread_create(&th[k], 0, part_main, &part[k]) == 0;
	part_main(&part[0]);
	for (k = 1; k < nt; ++k) {
		if (started[k]) pthread_join(th[k], 0);
		else part_main(&part[k]);
	}
	for (d = 0; d < 4; ++d) {              /* each orientation's distances, in one array */
		int64_t tot = 0, o = 0;
		for (k = 0; k < nt; ++k) tot += part[k].n[d];
		a[d] = (uint64_t*)malloc((tot + 1) * sizeof(uint64_t));
		if (a[d] == 0) abort();
		for (k = 0; k < nt; ++k) {
			if (part[k].n[d]) memcpy(a[d] + o, part[k].a[d], part[k].n[d] * sizeof(uint64_t));
			o += part[k].n[d];
			free(part[k].a[d]);
		}
		n[d] = (int32_t)tot;
	}
	if (kom_verbose >= 3)
		fprintf(stderr, "[M::%s] # candidate unique pairs for (FF, FR, RF, RR): (%d, %d, %d, %d)\n", "mb_pestat", n[0], n[1], n[2], n[3]);
	for (d = 0, max = 0; d < 4; ++d)
		max = max > n[d] ? max : n[d];
	for (d = 0; d < 4; ++d) {              /* pe.c:95-122 */
		mb_pestat_t *r = &pes[d];
		uint64_t *q = a[d];
		int p25, p50, p75, x;
		if (n[d] < MIN_DIR_CNT || n[d] < max * MIN_DIR_RATIO) {
			r->failed = 1;
			free(q);
			continue;
		}
		radix_sort_mb64(q, q + n[d]);
		p25 = q[(int)(.25 * n[d] + .499)];
		p50 = q[(int)(.50 * n[d] + .499)];
		p75 = q[(int)(.75 * n[d] + .499)];
		r->lo = (int)(p25 - OUTLIER_BOUND * (p75 - p25) + .499);
		if (r->lo < 1) r->lo = 1;
		r->hi = (int)(p75 + OUTLIER_BOUND * (p75 - p25) + .499);
		for (i = x = 0, r->avg = 0; i < n[d]; ++i)
			if (q[i] >= r->lo && q[i] <= r->hi)
				r->avg += q[i], ++x;
		r->avg /= x;
		for (i = 0, r->std = 0; i < n[d]; ++i)
			if (q[i] >= r->lo && q[i] <= r->hi)
				r->std += (q[i] - r->avg) * (q[i] - r->avg);
		r->std = sqrt(r->std / x);
		if (kom_verbose >= 3)
			fprintf(stderr, "[M::%s::%c%c] (25, 50, 75) percentile: (%d, %d, %d); mean and std.dev: (%.2f, %.2f)\n",
			        "mb_pestat", "FR"[d >> 1 & 1], "FR"[d & 1], p25, p50, p75, r->avg, r->std);
		r->lo = (int)(p25 - MAPPING_BOUND * (p75 - p25) + .499);
		r->hi = (int)(p75 + MAPPING_BOUND * (p75 - p25) + .499);
		if (r->lo > r->avg - MAX_STDDEV * r->std) r->lo = (int)(r->avg - MAX_STDDEV * r->std + .499);
		if (r->hi < r->avg + MAX_STDDEV * r->std) r->hi = (int)(r->avg + MAX_STDDEV * r->std + .499);
		if (r->lo < 1) r->lo = 1;
		if (kom_verbose >= 3)
			fprintf(stderr, "[M::%s::%c%c] low and high boundaries for proper pairs: (%d, %d)\n", "mb_pestat", "FR"[d >> 1 & 1],
			        "FR"[d & 1], r->lo, r->hi);
		free(q);
	}
}
