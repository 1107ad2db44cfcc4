/*
 * l2b_check.c -- test infrastructure: target extraction checked against the reference's l2b_getseq
 * (l2bit.c:52-77), which this program links compiled from the reference's own l2bit.c.
 *
 *   l2b-check <idx.l2b> <n_random> [seed]
 *       compares mbg_l2b_getseq (minibwa_amd/integration/mbg_l2b.c, the mapper's host extraction)
 *       with l2b_getseq on every range around every ambiguous run of the index (both ends, inside,
 *       spanning, at 32-base word boundaries), on clipped and invalid ranges, and on n_random random
 *       ranges; prints "ok <ranges> <ambiguous runs>" or the first difference (exit 1).
 *   l2b-check -d <idx.l2b> <ranges.txt> <out.bin>
 *       writes, for each line "tid st en" of ranges.txt, l2b_getseq's return value (int64) and then
 *       max(ret, 0) bytes: the expected output for the device's mbg_getseq.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "l2bit.h"
#include "mbg_l2b.h"

static uint64_t g_rng = 11;
static uint64_t rnd(void)   /* splitmix64 */
{
	uint64_t z = (g_rng += 0x9e3779b97f4a7c15ULL);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

static uint8_t *g_a, *g_b;
static int64_t g_cap, g_n;

static int check(const l2b_t *l2b, int64_t tid, int64_t st, int64_t en)
{
	int64_t ra, rb;
	memset(g_a, 0xaa, g_cap), memset(g_b, 0xaa, g_cap);
	ra = l2b_getseq(l2b, tid, st, en, g_a);
	rb = mbg_l2b_getseq(l2b, tid, st, en, g_b);
	++g_n;
	if (ra != rb || memcmp(g_a, g_b, g_cap) != 0) {
		int64_t i;
		for (i = 0; i < g_cap && g_a[i] == g_b[i]; ++i);
		fprintf(stderr, "[l2b_check] differs: tid %ld st %ld en %ld: return %ld vs %ld, first byte %ld (%d vs %d)\n",
		        (long)tid, (long)st, (long)en, (long)ra, (long)rb, (long)i, i < g_cap ? g_a[i] : -1, i < g_cap ? g_b[i] : -1);
		return -1;
	}
	return 0;
}

static int64_t ctg_of(const l2b_t *l2b, uint64_t g)
{
	uint64_t k;
	for (k = 0; k < l2b->n_ctg; ++k)
		if (g >= l2b->ctg[k].off && g < l2b->ctg[k].off + l2b->ctg[k].len) return (int64_t)k;
	return -1;
}

int main(int argc, char *argv[])
{
	l2b_t *l2b;
	uint64_t k;
	int64_t i, n_rand;
	if (argc >= 5 && strcmp(argv[1], "-d") == 0) {
		FILE *fi, *fo;
		long tid, st, en;
		if ((l2b = l2b_load(argv[2])) == 0) return 2;
		fi = fopen(argv[3], "r"), fo = fopen(argv[4], "wb");
		if (fi == 0 || fo == 0) return 2;
		g_cap = (int64_t)l2b->tot_len + 64;
		g_a = (uint8_t*)malloc(g_cap);
		while (fscanf(fi, "%ld %ld %ld", &tid, &st, &en) == 3) {
			int64_t r = l2b_getseq(l2b, tid, st, en, g_a);
			fwrite(&r, 8, 1, fo);
			if (r > 0) fwrite(g_a, 1, r, fo);
		}
		fclose(fi), fclose(fo);
		l2b_destroy(l2b);
		return 0;
	}
	if (argc < 3) {
		fprintf(stderr, "Usage: l2b-check <idx.l2b> <n_random> [seed]\n       l2b-check -d <idx.l2b> <ranges.txt> <out.bin>\n");
		return 2;
	}
	if ((l2b = l2b_load(argv[1])) == 0) {
		fprintf(stderr, "[l2b_check] cannot load %s\n", argv[1]);
		return 2;
	}
	n_rand = atol(argv[2]);
	if (argc > 3) g_rng = strtoull(argv[3], 0, 10);
	g_cap = (int64_t)l2b->tot_len + 64;
	g_a = (uint8_t*)malloc(g_cap), g_b = (uint8_t*)malloc(g_cap);
	/* around every ambiguous run */
	for (k = 0; k < l2b->n_ambi; ++k) {
		static const int64_t d[] = {0, 1, 2, 3, 4, 5, 31, 32, 33, 63, 64, 65, 200};
		const l2b_intv_t *iv = &l2b->ambi[k];
		int64_t tid = ctg_of(l2b, iv->st), s, e, a, b;
		if (tid < 0) continue;
		s = (int64_t)(iv->st - l2b->ctg[tid].off), e = (int64_t)(iv->en - l2b->ctg[tid].off);
		for (a = 0; a < (int64_t)(sizeof(d) / sizeof(d[0])); ++a)
			for (b = 0; b < (int64_t)(sizeof(d) / sizeof(d[0])); ++b) {
				if (check(l2b, tid, s - d[a], e + d[b]) < 0) return 1;        /* spanning */
				if (check(l2b, tid, s + d[a], e + d[b]) < 0) return 1;        /* starting inside */
				if (check(l2b, tid, s - d[a], s + d[b]) < 0) return 1;        /* ending inside or at the start */
				if (check(l2b, tid, e - d[a], e + d[b]) < 0) return 1;        /* around the end */
			}
	}
	/* clipping and invalid arguments */
	for (k = 0; k < l2b->n_ctg; ++k) {
		int64_t L = (int64_t)l2b->ctg[k].len;
		if (check(l2b, k, 0, L) < 0 || check(l2b, k, -10, L + 10) < 0 || check(l2b, k, L - 5, L + 100) < 0) return 1;
		if (check(l2b, k, 10, 10) < 0 || check(l2b, k, 20, 10) < 0 || check(l2b, k, 5, -3) < 0 || check(l2b, k, -7, -3) < 0) return 1;
		if (check(l2b, k, L, L + 5) < 0 || check(l2b, k, L + 5, L + 9) < 0) return 1;
	}
	if (check(l2b, -1, 0, 10) < 0 || check(l2b, (int64_t)l2b->n_ctg, 0, 10) < 0) return 1;
	/* random ranges */
	for (i = 0; i < n_rand; ++i) {
		int64_t tid = (int64_t)(rnd() % l2b->n_ctg), L = (int64_t)l2b->ctg[tid].len;
		int64_t st = (int64_t)(rnd() % (uint64_t)(L + 100)) - 50;
		int64_t len = rnd() % 16 == 0 ? (int64_t)(rnd() % 5000) : (int64_t)(rnd() % 400);
		if (check(l2b, tid, st, st + len) < 0) return 1;
	}
	printf("ok %ld %lu\n", (long)g_n, (unsigned long)l2b->n_ambi);
	free(g_a), free(g_b);
	l2b_destroy(l2b);
	return 0;
}
