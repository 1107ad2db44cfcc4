/*
 * mbg_fmt.c -- SAM and PAF records, byte for byte the reference's (interface in mbg_fmt.h).
 *
 * The fields, their order and every condition follow mb_fmt_sam (format.c:196-353), its helpers
 * get_sam_pri, write_sam_cigar and sam_write_sq (format.c:145-194), and mb_fmt_paf with write_tags
 * (format.c:13-46).  Integer fields keep kom_sprintf_lite's conversions (kommon.c:66-138): "%d" reads
 * an int, so the int64 coordinates the reference passes there (r->ts + 1 and the like) are truncated
 * to 32 bits exactly as its va_arg(ap, int) does; "%u" an unsigned 32-bit value; "%ld" a long.
 *
 * Room is reserved once per group of fields with a bound on what the group can write (ROOM_INT per
 * integer), then bytes are stored directly.
 */
#include <assert.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "mbg_fmt.h"

#define ROOM_INT 24          /* the longest integer field: '-' and 20 digits */

static char g_rg_id[256];    /* "" when the run has no read group */

static void grow(kstring_t *s, size_t add)
{
	size_t m = s->l + add + 1;
	if (m <= s->m) return;
	m += m >> 1;
	s->s = (char*)realloc(s->s, m);
	if (s->s == 0) {
		fprintf(stderr, "[mbg_fmt] out of memory\n");
		abort();
	}
	s->m = m;
}

static inline void room(kstring_t *s, size_t add)
{
	if (s->l + add + 1 > s->m) grow(s, add);
}

static const char kDigits2[201] =
	"00010203040506070809101112131415161718192021222324252627282930313233343536373839"
	"40414243444546474849505152535455565758596061626364656667686970717273747576777879"
	"8081828384858687888990919293949596979899";

/* unsigned decimal, most significant digit first */
static const uint64_t kPow10[20] = {1ULL, 10ULL, 100ULL, 1000ULL, 10000ULL, 100000ULL, 1000000ULL, 10000000ULL, 100000000ULL,
	1000000000ULL, 10000000000ULL, 100000000000ULL, 1000000000000ULL, 10000000000000ULL, 100000000000000ULL,
	1000000000000000ULL, 10000000000000000ULL, 100000000000000000ULL, 1000000000000000000ULL, 10000000000000000000ULL};

/* decimal digits of x written in place, two at a time from the last */
static inline void put_u64(kstring_t *s, uint64_t x)
{
	int n = 1;
	char *p;
	while (n < 20 && x >= kPow10[n]) ++n;
	p = s->s + s->l + n;
	s->l += (size_t)n;
	while (x >= 100) {
		unsigned r = (unsigned)(x % 100);
		x /= 100;
		p -= 2, memcpy(p, &kDigits2[r * 2], 2);
	}
	if (x >= 10) p -= 2, memcpy(p, &kDigits2[x * 2], 2);
	else *--p = (char)('0' + x);
}

/* "%d": an int; the magnitude of a negative value is taken as unsigned (INT_MIN included) */
static inline void put_d(kstring_t *s, int32_t c)
{
	if (c < 0) {
		s->s[s->l++] = '-';
		put_u64(s, 0u - (uint32_t)c);
	} else put_u64(s, (uint32_t)c);
}

/* "%ld": a long */
static inline void put_ld(kstring_t *s, int64_t c)
{
	if (c < 0) {
		s->s[s->l++] = '-';
		put_u64(s, 0ull - (uint64_t)c);
	} else put_u64(s, (uint64_t)c);
}

static inline void put_c(kstring_t *s, char c) { s->s[s->l++] = c; }

static inline void put_n(kstring_t *s, const char *p, size_t n)
{
	memcpy(s->s + s->l, p, n);
	s->l += n;
}

#define PUT_LIT(s, lit) put_n((s), (lit), sizeof(lit) - 1)

/* a C string, with its own reservation */
static inline void put_str(kstring_t *s, const char *p)
{
	size_t n = strlen(p);
	room(s, n);
	put_n(s, p, n);
}

/* format.c:157-164 */
static const mb_hit_t *sam_pri(int n_hit, const mb_hit_t *hit)
{
	int i;
	for (i = 0; i < n_hit; ++i)
		if (hit[i].sam_pri) return &hit[i];
	assert(n_hit == 0);
	return 0;
}

/* write_sam_cigar (format.c:166-194) for its only use, in_tag == 0 */
static void put_cigar(kstring_t *s, int sam_flag, int qlen, const mb_hit_t *r, uint64_t opt_flag)
{
	if (r->p == 0) {
		room(s, 1);
		put_c(s, '*');
	} else {
		uint32_t k, clip0, clip1;
		const mb_extra_t *p = r->p;
		char clip_char;
		clip0 = r->rev ? qlen - r->qe : r->qs;
		clip1 = r->rev ? r->qs : qlen - r->qe;
		clip_char = (((sam_flag & 0x800) || ((sam_flag & 0x100) && (opt_flag & MB_F_2ND_SEQ))) && !(opt_flag & MB_F_SUPP_SOFT)) ? 'H' : 'S';
		assert(clip0 < (uint32_t)qlen && clip1 < (uint32_t)qlen);
		room(s, (size_t)(p->n_cigar + 2) * (ROOM_INT + 1));
		if (clip0) put_d(s, (int32_t)clip0), put_c(s, clip_char);
		for (k = 0; k < (uint32_t)p->n_cigar; ++k)
			put_d(s, (int32_t)(p->cigar[k] >> 4)), put_c(s, MB_CIGAR_STR[p->cigar[k] & 0xf]);
		if (clip1) put_d(s, (int32_t)clip1), put_c(s, clip_char);
	}
}

/* sam_write_sq (format.c:145-155); -1 for a reverse-complemented byte >= 128 */
static int put_sq(kstring_t *s, const char *seq, int l, int rev, int comp)
{
	room(s, l > 0 ? (size_t)l : 0);
	if (rev) {
		char *o = s->s + s->l;
		int i;
		unsigned hi = 0;
		if (comp) {
			for (i = 0; i < l; ++i) {
				uint8_t c = (uint8_t)seq[l - 1 - i];
				hi |= c;
				o[i] = (char)kom_comp_table[c & 0x7f];
			}
			if (hi & 0x80) return -1;
		} else
			for (i = 0; i < l; ++i) o[i] = seq[l - 1 - i];
		s->l += l;
	} else if (l > 0) put_n(s, seq, (size_t)l);
	return 0;
}

/* write_tags (format.c:13-17) */
static void put_tags(kstring_t *s, const mb_hit_t *p)
{
	int32_t nm = p->blen - p->mlen + p->p->n_ambi;
	room(s, 4 * (ROOM_INT + 8));
	PUT_LIT(s, "\tNM:i:"); put_d(s, nm);
	PUT_LIT(s, "\tAS:i:"); put_d(s, p->p->dp_score);
	PUT_LIT(s, "\tms:i:"); put_d(s, p->p->dp_max0);
	PUT_LIT(s, "\tmd:i:"); put_d(s, p->p->dp_max - p->p->dp_max2);
}

/* mb_fmt_paf (format.c:19-46) */
static void fmt_paf(kstring_t *s, const l2b_t *l2b, const mb_bseq1_t *t, const mb_hit_t *p, uint64_t opt_flag, int n_seg, int seg_idx)
{
	put_str(s, t->name);
	room(s, 2 * ROOM_INT + 4);
	if (n_seg > 1 && seg_idx >= 0) put_c(s, '/'), put_d(s, seg_idx + 1);
	put_c(s, '\t'), put_ld(s, (long)t->l_seq);
	if (p == 0) {
		room(s, 32);
		PUT_LIT(s, "\t*\t*\t*\t*\t*\t*\t*\t0\t0\t0\n");
		s->s[s->l] = 0;
		return;
	}
	room(s, 12 * (ROOM_INT + 8));
	put_c(s, '\t'), put_d(s, p->qs);
	put_c(s, '\t'), put_d(s, p->qe);
	put_c(s, '\t'), put_c(s, p->rev ? '-' : '+');
	put_c(s, '\t');
	put_str(s, l2b->ctg[p->tid].name);
	room(s, 12 * (ROOM_INT + 8));
	put_c(s, '\t'), put_ld(s, (long)l2b->ctg[p->tid].len);
	put_c(s, '\t'), put_ld(s, (long)p->ts);
	put_c(s, '\t'), put_ld(s, (long)p->te);
	put_c(s, '\t'), put_d(s, p->mlen);
	put_c(s, '\t'), put_d(s, p->blen);
	put_c(s, '\t'), put_d(s, p->mapq);
	PUT_LIT(s, "\ttp:A:"), put_c(s, p->parent == p->id ? 'P' : 'S');
	PUT_LIT(s, "\ts1:i:"), put_d(s, p->score);
	PUT_LIT(s, "\tcm:i:"), put_d(s, p->cnt);
	if (p->parent == p->id) PUT_LIT(s, "\ts2:i:"), put_d(s, p->subsc >= 0 ? p->subsc : 0);
	if (p->p) {
		put_tags(s, p);
		if (p->p->n_cigar > 0) {
			int32_t i;
			room(s, 6 + (size_t)p->p->n_cigar * (ROOM_INT + 1));
			PUT_LIT(s, "\tcg:Z:");
			for (i = 0; i < p->p->n_cigar; ++i)
				put_d(s, (int32_t)(p->p->cigar[i] >> 4)), put_c(s, MB_CIGAR_STR[p->p->cigar[i] & 0xf]);
		}
		if (p->p->cs) {
			room(s, 1);
			put_c(s, '\t');
			put_str(s, (const char*)&p->p->cigar[p->p->n_cigar]);
		}
	}
	if ((opt_flag & MB_F_COPY_COMMENT) && t->comment) {
		room(s, 1);
		put_c(s, '\t');
		put_str(s, t->comment);
	}
	room(s, 1);
	put_c(s, '\n');
	s->s[s->l] = 0;
}

/* mb_fmt_sam (format.c:196-353) */
static int fmt_sam(kstring_t *s, const l2b_t *l2b, const mb_bseq1_t *t, int32_t n_seg, const int32_t *n_hit, mb_hit_t *const *hit,
                   int32_t hit_idx, const mb_opt_t *opt, int seg_idx, int32_t mate_qlen)
{
	int flag, n_h = n_hit[seg_idx];
	int this_tid = -1, this_pos = -1;
	const int qlen = (int)t->l_seq;
	const mb_hit_t *h = hit[seg_idx], *r_prev, *r_next;
	const mb_hit_t *r = n_h > 0 && hit_idx < n_h && hit_idx >= 0 ? &h[hit_idx] : 0;

	assert(n_seg == 1 || n_seg == 2);
	if (n_seg > 1) {
		int next_sid = (seg_idx + 1) % n_seg;
		r_prev = r_next = sam_pri(n_hit[next_sid], hit[next_sid]);
	} else r_prev = r_next = 0;

	/* QNAME and FLAG */
	put_str(s, t->name);
	flag = n_seg > 1 ? 0x1 : 0x0;
	if (r == 0) flag |= 0x4;
	else {
		if (r->rev) flag |= 0x10;
		if (r->parent != r->id) flag |= 0x100;
		else if (!r->sam_pri) flag |= 0x800;
	}
	if (n_seg > 1) {
		if (r && r->proper_pair) flag |= 0x2;
		if (seg_idx == 0) flag |= 0x40;
		else if (seg_idx == n_seg - 1) flag |= 0x80;
		if (r_next == 0) flag |= 0x8;
		else if (r_next->rev) flag |= 0x20;
	}
	room(s, ROOM_INT + 1);
	put_c(s, '\t'), put_d(s, flag);

	/* RNAME, POS, MAPQ and CIGAR */
	if (r == 0) {
		if (r_prev) {
			this_tid = (int)r_prev->tid, this_pos = (int)r_prev->ts;
			room(s, 1);
			put_c(s, '\t');
			put_str(s, l2b->ctg[this_tid].name);
			room(s, ROOM_INT + 8);
			put_c(s, '\t'), put_d(s, this_pos + 1);
			PUT_LIT(s, "\t0\t*");
		} else {
			room(s, 8);
			PUT_LIT(s, "\t*\t0\t0\t*");
		}
	} else {
		this_tid = (int)r->tid, this_pos = (int)r->ts;
		room(s, 1);
		put_c(s, '\t');
		put_str(s, l2b->ctg[r->tid].name);
		room(s, 2 * ROOM_INT + 4);
		put_c(s, '\t'), put_d(s, (int32_t)(r->ts + 1));
		put_c(s, '\t'), put_d(s, r->mapq);
		put_c(s, '\t');
		put_cigar(s, flag, qlen, r, opt->flag);
	}

	/* RNEXT, PNEXT and TLEN */
	if (n_seg > 1) {
		int tlen = 0;
		if (this_tid >= 0 && r_next) {
			if (this_tid == r_next->tid) {
				if (r) {
					int this_pos5 = r->rev ? (int)(r->te - 1) : this_pos;
					int next_pos5 = r_next->rev ? (int)(r_next->te - 1) : (int)r_next->ts;
					tlen = next_pos5 - this_pos5;
				}
				room(s, 3);
				PUT_LIT(s, "\t=\t");
			} else {
				room(s, 1);
				put_c(s, '\t');
				put_str(s, l2b->ctg[r_next->tid].name);
				room(s, 1);
				put_c(s, '\t');
			}
			room(s, ROOM_INT + 1);
			put_d(s, (int32_t)(r_next->ts + 1)), put_c(s, '\t');
		} else if (r_next) {
			room(s, 1);
			put_c(s, '\t');
			put_str(s, l2b->ctg[r_next->tid].name);
			room(s, ROOM_INT + 2);
			put_c(s, '\t'), put_d(s, (int32_t)(r_next->ts + 1)), put_c(s, '\t');
		} else if (this_tid >= 0) {    /* the mate is unmapped: placed at this segment's primary */
			const mb_hit_t *r_pri = sam_pri(n_h, h);
			if (r->tid == r_pri->tid) {
				room(s, ROOM_INT + 4);
				PUT_LIT(s, "\t=\t"), put_d(s, (int32_t)(r_pri->ts + 1)), put_c(s, '\t');
			} else {
				room(s, 1);
				put_c(s, '\t');
				put_str(s, l2b->ctg[r_pri->tid].name);
				room(s, ROOM_INT + 2);
				put_c(s, '\t'), put_d(s, (int32_t)(r_pri->ts + 1)), put_c(s, '\t');
			}
		} else {
			room(s, 6);
			PUT_LIT(s, "\t*\t0\t");
		}
		if (tlen > 0) ++tlen;
		else if (tlen < 0) --tlen;
		room(s, ROOM_INT + 1);
		put_d(s, tlen), put_c(s, '\t');
	} else {
		room(s, 8);
		PUT_LIT(s, "\t*\t0\t0\t");
	}

	/* SEQ and QUAL */
	if (r == 0) {
		put_sq(s, t->seq, qlen, 0, 0);
		room(s, 2);
		put_c(s, '\t');
		if (t->qual) put_sq(s, t->qual, qlen, 0, 0);
		else put_c(s, '*');
	} else {
		if ((flag & 0x900) == 0 || (opt->flag & MB_F_SUPP_SOFT)) {
			if (put_sq(s, t->seq, qlen, r->rev, r->rev) < 0) return -1;
			room(s, 2);
			put_c(s, '\t');
			if (t->qual) put_sq(s, t->qual, qlen, r->rev, 0);
			else put_c(s, '*');
		} else if ((flag & 0x100) && !(opt->flag & MB_F_2ND_SEQ)) {
			room(s, 3);
			PUT_LIT(s, "*\t*");
		} else {
			if (put_sq(s, t->seq + r->qs, r->qe - r->qs, r->rev, r->rev) < 0) return -1;
			room(s, 2);
			put_c(s, '\t');
			if (t->qual) put_sq(s, t->qual + r->qs, r->qe - r->qs, r->rev, 0);
			else put_c(s, '*');
		}
	}

	/* tags */
	if (g_rg_id[0]) {
		room(s, 6);
		PUT_LIT(s, "\tRG:Z:");
		put_str(s, g_rg_id);
	}
	if (r) {
		put_tags(s, r);
		if (n_seg > 1 && r_next && r_next->p && r_next->p->n_cigar > 0 && mate_qlen > 0) {
			room(s, 6);
			PUT_LIT(s, "\tMC:Z:");
			put_cigar(s, 0, mate_qlen, r_next, opt->flag);
			room(s, ROOM_INT + 6);
			PUT_LIT(s, "\tMQ:i:"), put_d(s, r_next->mapq);
		}
		if (r->p->cs) {
			room(s, 1);
			put_c(s, '\t');
			put_str(s, (const char*)&r->p->cigar[r->p->n_cigar]);
		}
		if (r->parent == r->id && r->p && n_h > 1 && h && r >= h && r - h < n_h) {   /* supplementary alignments may exist */
			int i, n_sa = 0;
			for (i = 0; i < n_h; ++i)
				if (i != r - h && h[i].parent == h[i].id && h[i].p) ++n_sa;
			if (n_sa > 0) {
				room(s, 6);
				PUT_LIT(s, "\tSA:Z:");
				for (i = 0; i < n_h; ++i) {
					const mb_hit_t *q = &h[i];
					int l_M, l_I = 0, l_D = 0, clip5, clip3;
					if (r == q || q->parent != q->id || q->p == 0) continue;
					if (q->qe - q->qs < q->te - q->ts) l_M = q->qe - q->qs, l_D = (int)((q->te - q->ts) - l_M);
					else l_M = (int)(q->te - q->ts), l_I = (q->qe - q->qs) - l_M;
					clip5 = q->rev ? (int)(t->l_seq - q->qe) : q->qs;
					clip3 = q->rev ? q->qs : (int)(t->l_seq - q->qe);
					put_str(s, l2b->ctg[q->tid].name);
					room(s, 8 * (ROOM_INT + 2));
					put_c(s, ','), put_d(s, (int32_t)(q->ts + 1)), put_c(s, ','), put_c(s, "+-"[q->rev]), put_c(s, ',');
					if (clip5) put_d(s, clip5), put_c(s, 'S');
					if (l_M) put_d(s, l_M), put_c(s, 'M');
					if (l_I) put_d(s, l_I), put_c(s, 'I');
					if (l_D) put_d(s, l_D), put_c(s, 'D');
					if (clip3) put_d(s, clip3), put_c(s, 'S');
					put_c(s, ','), put_d(s, q->mapq), put_c(s, ','), put_d(s, q->blen - q->mlen + q->p->n_ambi), put_c(s, ';');
				}
			}
			if (opt->xa_max > 0) {
				int n_xa = 0;
				for (i = 0; i < n_h; ++i)
					if (i != r - h && h[i].parent == r - h && h[i].p->dp_max >= (double)opt->out_s * r->p->dp_max) ++n_xa;
				if (n_xa > 0) {
					room(s, ROOM_INT + 6);
					PUT_LIT(s, "\tn2:i:"), put_d(s, n_xa);
				}
				if (n_xa > 0 && n_xa <= opt->xa_max) {
					room(s, 6);
					PUT_LIT(s, "\tXA:Z:");
					for (i = 0; i < n_h; ++i) {
						const mb_hit_t *q = &h[i];
						if (i != r - h && q->parent == r - h && q->p->dp_max >= (double)opt->out_s * r->p->dp_max) {
							put_str(s, l2b->ctg[q->tid].name);
							room(s, ROOM_INT + 3);
							put_c(s, ','), put_c(s, "+-"[q->rev]), put_d(s, (int32_t)(q->ts + 1)), put_c(s, ',');
							put_cigar(s, 0, qlen, q, opt->flag);
							room(s, ROOM_INT + 2);
							put_c(s, ','), put_d(s, q->blen - q->mlen + q->p->n_ambi), put_c(s, ';');
						}
					}
				}
			}
		}
	}
	if ((opt->flag & MB_F_COPY_COMMENT) && t->comment) {
		room(s, 1);
		put_c(s, '\t');
		put_str(s, t->comment);
	}
	room(s, 1);
	put_c(s, '\n');
	s->s[s->l] = 0;
	return 0;
}

int mbg_fmt_record(kstring_t *s, const l2b_t *l2b, const mb_bseq1_t *t, int32_t n_seg, const int32_t *n_hit,
                   mb_hit_t *const *hit, int32_t hit_idx, const mb_opt_t *opt, int seg_idx, int32_t mate_qlen)
{
	if (!(opt->flag & MB_F_PAF)) {
		size_t l0 = s->l;
		if (fmt_sam(s, l2b, t, n_seg, n_hit, hit, hit_idx, opt, seg_idx, mate_qlen) < 0) {
			s->l = l0;
			if (s->s) s->s[l0] = 0;
			return -1;
		}
	} else
		fmt_paf(s, l2b, t, hit_idx >= 0 ? &hit[seg_idx][hit_idx] : 0, opt->flag, n_seg, seg_idx);
	return 0;
}

int mbg_fmt_learn_rg(const mb_opt_t *opt)
{
	/* an unmapped single-end probe: "q\t4\t*\t0\t0\t*\t*\t0\t0\tA\t*" then "\tRG:Z:<id>" when a read
	 * group is set, then the comment (none here) and the newline */
	static const char head[] = "q\t4\t*\t0\t0\t*\t*\t0\t0\tA\t*";
	char name[] = "q", seq[] = "A";
	mb_bseq1_t t;
	int32_t n_hit = 0;
	mb_hit_t *hit = 0;
	mb_opt_t o = *opt;
	kstring_t s = {0, 0, 0};
	const char *p;
	size_t n;
	memset(&t, 0, sizeof(t));
	t.l_seq = 1, t.name = name, t.seq = seq;
	o.flag &= ~(uint64_t)(MB_F_PAF | MB_F_COPY_COMMENT);
	g_rg_id[0] = 0;
	mb_format(0, &s, 0, &t, 1, &n_hit, &hit, -1, &o, 0, 0);   /* SAM: mb_fmt_sam */
	if (s.s == 0 || strncmp(s.s, head, sizeof(head) - 1) != 0) {
		free(s.s);
		return -1;
	}
	p = s.s + sizeof(head) - 1;
	if (strncmp(p, "\tRG:Z:", 6) == 0) {
		p += 6;
		n = strlen(p);
		if (n == 0 || p[n - 1] != '\n' || n > sizeof(g_rg_id)) {
			free(s.s);
			return -1;
		}
		memcpy(g_rg_id, p, n - 1);
		g_rg_id[n - 1] = 0;
	} else if (strcmp(p, "\n") != 0) {
		free(s.s);
		return -1;
	}
	free(s.s);
	return 0;
}
