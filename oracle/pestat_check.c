/*
 * pestat_check.c -- mbg_pestat (minibwa_amd/integration/mbg_pe.c) against the reference's own
 * mb_pestat (pe.c:68-124, compiled where it lies) on seeded synthetic hit sets: every mb_pestat_t
 * field of the four orientations must be identical, bit for bit, for 1, 3 and 16 threads.
 *
 * A set holds n_frag fragments, most of them pairs; each read has 0-3 hits, of which one or more are
 * primary (id == parent), with MAPQ around the uniqueness cut (10), on one of three contigs.  Mates
 * are placed at an insert size drawn around a set-specific mean, mostly in FR orientation, with
 * some FF/RF/RR pairs and far outliers, so that the outlier bounds, the minimum-count and the
 * minimum-ratio rules (pe.c:97) and the max_pe_ins cut all take effect across the sets.
 *
 *   pestat-check <n_sets> <seed>      prints "ok <sets> <fields compared> <orientations estimated>
 *                                     <orientations failed>" or the first difference
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mbpriv.h"
#include "mbg_pe.h"

static uint64_t g_s;
static uint64_t rnd(void)                 /* splitmix64 */
{
	uint64_t z = (g_s += 0x9e3779b97f4a7c15ULL);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}
static double unif(void) { return (rnd() >> 11) * (1.0 / 9007199254740992.0); }
static int64_t irand(int64_t n) { return n > 0 ? (int64_t)(rnd() % (uint64_t)n) : 0; }

static void make_hits(int32_t n, mb_hit_t *h, int64_t tid, int64_t pos, int rev, int unique)
{
	int32_t j;
	memset(h, 0, n * sizeof(mb_hit_t));
	for (j = 0; j < n; ++j) {
		h[j].id = j;
		h[j].parent = (j == 0 || (!unique && unif() < .5)) ? j : 0;
		h[j].mapq = j == 0 ? (int32_t)(unique ? 10 + irand(51) : irand(20)) : (int32_t)irand(60);
		h[j].tid = j == 0 ? tid : irand(3);
		h[j].rev = j == 0 ? rev : (uint32_t)irand(2);
		h[j].ts = j == 0 ? pos : irand(1000000);
		h[j].te = h[j].ts + 100 + irand(60);
	}
}

int main(int argc, char *argv[])
{
	int n_sets = argc > 1 ? atoi(argv[1]) : 40, set, nt_list[3] = {1, 3, 16}, t, d;
	long fields = 0, n_ok = 0, n_failed = 0;
	g_s = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
	kom_verbose = 0;
	for (set = 0; set < n_sets; ++set) {
		mb_opt_t opt;
		int32_t n_frag, n_seq = 0, i, *seg_off, *seg_cnt, *n_hit;
		mb_hit_t **hit;
		double mean = 150 + irand(800), sd = 5 + irand(120), p_other = unif() * (set % 3 == 0 ? .5 : .05);
		mb_pestat_t ref[4];
		mb_opt_init(&opt);
		opt.max_pe_ins = set % 5 == 0 ? (int32_t)(mean + irand(200)) : opt.max_pe_ins;
		n_frag = set % 7 == 0 ? (int32_t)irand(60) : (int32_t)(1000 + irand(60000));
		seg_off = (int32_t*)malloc((n_frag + 1) * sizeof(int32_t));
		seg_cnt = (int32_t*)malloc((n_frag + 1) * sizeof(int32_t));
		n_hit = (int32_t*)malloc((2 * n_frag + 1) * sizeof(int32_t));
		hit = (mb_hit_t**)calloc(2 * n_frag + 1, sizeof(mb_hit_t*));
		for (i = 0; i < n_frag; ++i) {
			int cnt = unif() < .03 ? 1 : 2, k, dir = unif() < p_other ? (int)irand(4) : 1;
			int64_t tid = irand(3), p0 = 10000 + irand(50000000), ins;
			ins = (int64_t)(mean + sd * (unif() + unif() + unif() - 1.5) * 2.0);
			if (unif() < .01) ins += irand(200000);                  /* far outliers */
			if (ins < 0) ins = -ins;
			seg_off[i] = n_seq, seg_cnt[i] = cnt;
			for (k = 0; k < cnt; ++k) {
				int32_t nh = unif() < .05 ? 0 : (int32_t)(1 + irand(3));
				int unique = unif() < .85, rev = k == 0 ? (dir >> 1 & 1) : (dir & 1);
				int64_t pos = k == 0 ? p0 : (unif() < .5 ? p0 + ins : p0 - ins);
				if (pos < 0) pos = 0;
				hit[n_seq] = (mb_hit_t*)malloc((nh + 1) * sizeof(mb_hit_t));
				make_hits(nh, hit[n_seq], unif() < .02 ? (tid + 1) % 3 : tid, pos, rev, unique);
				n_hit[n_seq++] = nh;
			}
		}
		mb_pestat(0, &opt, n_frag, seg_off, seg_cnt, n_hit, hit, ref);
		for (d = 0; d < 4; ++d) ref[d].failed ? ++n_failed : ++n_ok;
		for (t = 0; t < 3; ++t) {
			mb_pestat_t ours[4];
			mbg_pestat(&opt, n_frag, seg_off, seg_cnt, n_hit, hit, ours, nt_list[t]);
			for (d = 0; d < 4; ++d) {
				fields += 5;
				if (ours[d].failed != ref[d].failed || ours[d].lo != ref[d].lo || ours[d].hi != ref[d].hi ||
				    memcmp(&ours[d].avg, &ref[d].avg, sizeof(double)) != 0 || memcmp(&ours[d].std, &ref[d].std, sizeof(double)) != 0) {
					printf("diff set %d threads %d dir %d: ours failed %d lo %d hi %d avg %.17g std %.17g; ref failed %d lo %d hi %d avg %.17g std %.17g\n",
					       set, nt_list[t], d, ours[d].failed, ours[d].lo, ours[d].hi, ours[d].avg, ours[d].std, ref[d].failed,
					       ref[d].lo, ref[d].hi, ref[d].avg, ref[d].std);
					return 1;
				}
			}
		}
		for (i = 0; i < n_seq; ++i) free(hit[i]);
		free(hit); free(n_hit); free(seg_cnt); free(seg_off);
	}
	printf("ok %d %ld %ld %ld\n", n_sets, fields, n_ok, n_failed);
	return 0;
}
