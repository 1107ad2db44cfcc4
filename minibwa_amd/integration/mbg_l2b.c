/*
 * mbg_l2b.c -- l2b_getseq for the batched mapper (interface in mbg_l2b.h).
 *
 * The packed reference (l2bit.h:52-55) holds 32 bases per 64-bit word, base i in bits 2(i%32)..+1 of
 * word i/32; on a little-endian host, byte j of that array therefore holds bases 4j..4j+3 in its bit
 * pairs from the lowest up.  A 256-entry table turns one such byte into four base codes.  The
 * mapper's callers (align.c:642,689,734,751,780; pe.c:350; seed.c:240) reach this function through
 * the linker's --wrap of l2b_getseq; the N runs come from the reference's own l2b_getambi.
 */
#include <string.h>
#include "mbg_l2b.h"

static uint32_t g_unpack4[256];

__attribute__((constructor)) static void unpack4_init(void)
{
	int b, k;
	for (b = 0; b < 256; ++b) {
		uint8_t o[4];
		for (k = 0; k < 4; ++k) o[k] = (uint8_t)(b >> (2 * k) & 3);
		memcpy(&g_unpack4[b], o, 4);
	}
}

static inline uint8_t base_at(const uint64_t *pac, int64_t i)
{
	return (uint8_t)(pac[i >> 5] >> ((i & 31) << 1) & 3);
}

int64_t mbg_l2b_getseq(const l2b_t *l2b, int64_t tid, int64_t st, int64_t en, uint8_t *seq)
{
	const l2b_ctg_t *ctg;
	const uint8_t *pb = (const uint8_t*)l2b->pac;
	int64_t i, aid;
	int32_t n_ambi;
	uint8_t *o;
	if (tid < 0 || (uint64_t)tid >= l2b->n_ctg) return -1;
	ctg = &l2b->ctg[tid];
	if (st < 0) st = 0;
	if ((uint64_t)en > ctg->len) en = (int64_t)ctg->len;   /* an unsigned comparison there: a negative en clips to the end */
	st += ctg->off;
	en += ctg->off;
	for (i = st, o = seq; i < en && (i & 3); ++i) *o++ = base_at(l2b->pac, i);
	for (; i + 4 <= en; i += 4, o += 4) memcpy(o, &g_unpack4[pb[i >> 2]], 4);
	for (; i < en; ++i) *o++ = base_at(l2b->pac, i);
	aid = l2b_getambi(l2b, tid, st - ctg->off, en - ctg->off, &n_ambi);
	if (aid >= 0) {
		for (i = 0; i < n_ambi; ++i) {
			const l2b_intv_t *iv = &l2b->ambi[aid + i];
			int64_t s = iv->st, e = iv->en;
			if (s < st) s = st;
			if (e > en) e = en;
			if (s < e) memset(&seq[s - st], 4, e - s);
		}
	}
	return en - st;
}

int64_t __wrap_l2b_getseq(const l2b_t *l2b, int64_t tid, int64_t st, int64_t en, uint8_t *seq);
int64_t __wrap_l2b_getseq(const l2b_t *l2b, int64_t tid, int64_t st, int64_t en, uint8_t *seq)
{
	return mbg_l2b_getseq(l2b, tid, st, en, seq);
}
