/*
 * mbsim.c -- seeded synthetic references and reads for the benchmark and parity workloads.
 *
 * All randomness comes from splitmix64 (the generator minibwa itself uses, kommon.h:66-72), so a
 * (seed, size) pair reproduces the same bytes on any machine.  Sequences are nt4 codes (A,C,G,T =
 * 0..3).  Workloads:
 *   reference   uniform ACGT, then a fraction of it overwritten by copies of earlier segments
 *               (length uniform in [rep_min, rep_max]) carrying rep_div substitutions per base;
 *   read pairs  fragment length ~ Normal(frag_mean, frag_sd), uniform position and strand, read 1
 *               from the fragment's start, read 2 the reverse complement of its end; per-base
 *               substitution rate `sub`, indel rate `indel` (insertion or deletion with equal
 *               probability, geometric length with mean 2);
 *   single-end  lengths log-uniform in [lmin, lmax], same error model;
 *   DP jobs     query uniform in [qmin, qmax]; target = mutated query with random flanks.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t sm64(uint64_t *x)
{
	uint64_t z = ((*x) += 0x9e3779b97f4a7c15ULL);
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

static inline double u01(uint64_t *x) { return (sm64(x) >> 11) * (1.0 / 9007199254740992.0); }

static double normal(uint64_t *x)
{
	double u1 = u01(x), u2 = u01(x);
	if (u1 < 1e-300) u1 = 1e-300;
	return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static int geom2(uint64_t *x) /* geometric on {1,2,...} with mean 2 */
{
	int l = 1;
	while (u01(x) < 0.5) ++l;
	return l;
}

void mbs_ref(uint64_t seed, int64_t len, double rep_frac, int32_t rep_min, int32_t rep_max, double rep_div, uint8_t *out)
{
	uint64_t x = seed;
	int64_t i, done = 0, target = (int64_t)(rep_frac * (double)len);
	for (i = 0; i + 32 <= len; i += 32) {
		uint64_t r = sm64(&x);
		int k;
		for (k = 0; k < 32; ++k) out[i + k] = (uint8_t)(r >> (2 * k) & 3);
	}
	for (; i < len; ++i) out[i] = (uint8_t)(sm64(&x) & 3);
	if (rep_max > len / 4) rep_max = (int32_t)(len / 4);
	if (rep_min > rep_max) rep_min = rep_max;
	while (done < target && rep_max > 0) {
		int64_t l = rep_min + (int64_t)(sm64(&x) % (uint64_t)(rep_max - rep_min + 1));
		int64_t src = (int64_t)(sm64(&x) % (uint64_t)(len - l + 1));
		int64_t dst = (int64_t)(sm64(&x) % (uint64_t)(len - l + 1));
		int64_t j;
		if (dst >= src && dst < src + l) continue;
		if (src >= dst && src < dst + l) continue;
		for (j = 0; j < l; ++j) {
			uint8_t c = out[src + j];
			if (u01(&x) < rep_div) c = (uint8_t)((c + 1 + sm64(&x) % 3) & 3);
			out[dst + j] = c;
		}
		done += l;
	}
}

/* copy the mutated version of src[0..) into dst until dst holds `len` bases; returns bases of src used */
static int64_t mutate_into(uint64_t *x, const uint8_t *src, int64_t src_avail, int32_t len, double sub, double indel, uint8_t *dst)
{
	int64_t i = 0;
	int32_t o = 0;
	while (o < len) {
		double r = u01(x);
		if (r < indel) {
			int l = geom2(x);
			if (u01(x) < 0.5) { /* insertion */
				for (; l > 0 && o < len; --l) dst[o++] = (uint8_t)(sm64(x) & 3);
			} else i += l;  /* deletion */
			continue;
		}
		{
			uint8_t c = i < src_avail ? src[i] : (uint8_t)(sm64(x) & 3);
			if (u01(x) < sub) c = (uint8_t)((c + 1 + sm64(x) % 3) & 3);
			dst[o++] = c;
			++i;
		}
	}
	return i;
}

/* n_pairs read pairs of length rlen; truth[i*2] = fragment start, truth[i*2+1] = strand<<32 | frag_len */
int mbs_pairs(uint64_t seed, const uint8_t *ref, int64_t ref_len, int64_t n_pairs, int32_t rlen, double frag_mean, double frag_sd,
              double sub, double indel, uint8_t *r1, uint8_t *r2, int64_t *truth)
{
	uint64_t x = seed ^ 0x5bd1e995ULL;
	int64_t i, j, fmax = (int64_t)(frag_mean + 20 * frag_sd) + 4 * (int64_t)rlen + 64;
	uint8_t *frag = (uint8_t*)malloc((size_t)fmax), *rc = (uint8_t*)malloc((size_t)fmax);
	if (!frag || !rc) { free(frag); free(rc); return -1; }
	for (i = 0; i < n_pairs; ++i) {
		int64_t fl = (int64_t)(frag_mean + frag_sd * normal(&x) + 0.5), st;
		int strand = (int)(sm64(&x) & 1);
		if (fl < rlen) fl = rlen;
		if (fl > fmax) fl = fmax;
		if (fl > ref_len) fl = ref_len;
		st = (int64_t)(sm64(&x) % (uint64_t)(ref_len - fl + 1));
		if (strand == 0) memcpy(frag, ref + st, (size_t)fl);
		else for (j = 0; j < fl; ++j) frag[j] = 3 - ref[st + fl - 1 - j];
		mutate_into(&x, frag, fl, rlen, sub, indel, r1 + i * rlen);
		for (j = 0; j < fl; ++j) rc[j] = 3 - frag[fl - 1 - j]; /* read 2: reverse complement of the end */
		mutate_into(&x, rc, fl, rlen, sub, indel, r2 + i * rlen);
		if (truth) truth[i * 2] = st, truth[i * 2 + 1] = (int64_t)strand << 32 | fl;
	}
	free(frag); free(rc);
	return 0;
}

/* n single-end reads, log-uniform lengths in [lmin, lmax]; out holds sum(len) bases, len[i] the lengths */
int64_t mbs_single(uint64_t seed, const uint8_t *ref, int64_t ref_len, int64_t n, int32_t lmin, int32_t lmax, double sub, double indel,
                   uint8_t *out, int64_t out_cap, int32_t *len)
{
	uint64_t x = seed ^ 0x27d4eb2fULL;
	int64_t i, j, o = 0;
	uint8_t *frag = (uint8_t*)malloc((size_t)lmax * 2 + 64);
	if (!frag) return -1;
	for (i = 0; i < n; ++i) {
		double lg = log((double)lmin) + u01(&x) * (log((double)lmax) - log((double)lmin));
		int32_t l = (int32_t)(exp(lg) + 0.5);
		int64_t span = (int64_t)l * 2 < ref_len ? (int64_t)l * 2 : ref_len, st;
		int strand = (int)(sm64(&x) & 1);
		if (l < lmin) l = lmin;
		if (l > lmax) l = lmax;
		if (o + l > out_cap) { free(frag); return -2; }
		st = (int64_t)(sm64(&x) % (uint64_t)(ref_len - span + 1));
		if (strand == 0) memcpy(frag, ref + st, (size_t)span);
		else for (j = 0; j < span; ++j) frag[j] = 3 - ref[st + span - 1 - j];
		mutate_into(&x, frag, span, l, sub, indel, out + o);
		len[i] = l;
		o += l;
	}
	free(frag);
	return o;
}

/* DP jobs: query uniform in [qmin,qmax]; target = mutated query plus random flanks of flank bases on
 * the right (and on the left too when two_sided).  Returns bases written to tbuf. */
int64_t mbs_dp_jobs(uint64_t seed, int64_t n, int32_t qmin, int32_t qmax, double sub, double indel, int32_t flank, int two_sided,
                    uint8_t *qbuf, int64_t qcap, int32_t *qlen, uint8_t *tbuf, int64_t tcap, int32_t *tlen)
{
	uint64_t x = seed ^ 0x9e3779b1ULL;
	int64_t i, j, qo = 0, to = 0;
	uint8_t *mut = (uint8_t*)malloc((size_t)qmax * 3 + 64);
	if (!mut) return -1;
	for (i = 0; i < n; ++i) {
		int32_t ql = qmin + (int32_t)(sm64(&x) % (uint64_t)(qmax - qmin + 1)), ml, tl;
		if (qo + ql > qcap) { free(mut); return -2; }
		for (j = 0; j < ql; ++j) qbuf[qo + j] = (uint8_t)(sm64(&x) & 3);
		/* mutated copy: walk the query with the error model */
		{
			int64_t k = 0;
			ml = 0;
			while (k < ql && ml < qmax * 3) {
				double r = u01(&x);
				if (r < indel) {
					int l = geom2(&x);
					if (u01(&x) < 0.5) { for (; l > 0 && ml < qmax * 3; --l) mut[ml++] = (uint8_t)(sm64(&x) & 3); }
					else k += l;
					continue;
				}
				{
					uint8_t c = qbuf[qo + k++];
					if (u01(&x) < sub) c = (uint8_t)((c + 1 + sm64(&x) % 3) & 3);
					mut[ml++] = c;
				}
			}
		}
		tl = ml + flank * (two_sided ? 2 : 1);
		if (to + tl > tcap) { free(mut); return -2; }
		{
			int64_t p = to;
			if (two_sided) for (j = 0; j < flank; ++j) tbuf[p++] = (uint8_t)(sm64(&x) & 3);
			memcpy(tbuf + p, mut, (size_t)ml); p += ml;
			for (j = 0; j < flank; ++j) tbuf[p++] = (uint8_t)(sm64(&x) & 3);
		}
		qlen[i] = ql, tlen[i] = tl;
		qo += ql, to += tl;
	}
	free(mut);
	return to;
}

/* write reads as FASTQ (quality 'I'); names <prefix><i>/<mate> */
int mbs_write_fastq(const char *fn, const char *prefix, int64_t n, const uint8_t *seq, const int64_t *off, const int32_t *len, int mate)
{
	FILE *fp = fopen(fn, "w");
	int64_t i;
	int32_t j;
	char *line;
	int32_t maxl = 0;
	if (fp == 0) return -1;
	for (i = 0; i < n; ++i) if (len[i] > maxl) maxl = len[i];
	line = (char*)malloc((size_t)maxl + 2);
	for (i = 0; i < n; ++i) {
		const uint8_t *s = seq + off[i];
		for (j = 0; j < len[i]; ++j) line[j] = "ACGTN"[s[j] < 4 ? s[j] : 4];
		line[len[i]] = 0;
		if (mate > 0) fprintf(fp, "@%s%lld/%d\n%s\n+\n", prefix, (long long)i, mate, line);
		else fprintf(fp, "@%s%lld\n%s\n+\n", prefix, (long long)i, line);
		memset(line, 'I', (size_t)len[i]);
		fputs(line, fp);
		fputc('\n', fp);
	}
	free(line);
	fclose(fp);
	return 0;
}

/* write a reference as FASTA with n_ctg equal contigs named <prefix><i> */
int mbs_write_fasta(const char *fn, const char *prefix, const uint8_t *ref, int64_t len, int32_t n_ctg)
{
	FILE *fp = fopen(fn, "w");
	int64_t c, i, st, en, per = (len + n_ctg - 1) / n_ctg;
	char buf[81];
	if (fp == 0) return -1;
	for (c = 0; c < n_ctg; ++c) {
		st = c * per, en = st + per < len ? st + per : len;
		fprintf(fp, ">%s%lld\n", prefix, (long long)(c + 1));
		for (i = st; i < en; i += 80) {
			int64_t k, e = i + 80 < en ? i + 80 : en;
			for (k = i; k < e; ++k) buf[k - i] = "ACGT"[ref[k] & 3];
			buf[e - i] = 0;
			fputs(buf, fp); fputc('\n', fp);
		}
	}
	fclose(fp);
	return 0;
}
