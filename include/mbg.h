/*
 * mbg.h -- C ABI of libmbg.so, the MI355X (gfx950) seeding + extension library for minibwa.
 *
 * Plain pointers and sizes only.  Host-pointer entry points are drop-in batch replacements for
 * the reference's hot-path functions (minibwa 0.7-r421); *_dev entry points take device pointers
 * and a hipStream_t (passed as void*) for callers that keep inputs resident in HBM.
 *
 * Every entry point returns MBG_OK (0) or a negative MBG_E* code; mbg_strerror() names it.
 * A context serialises its own calls (one mutex), so one context may be shared by threads.  The
 * calls share the context's scratch on the device; a *_dev call may return with its work still
 * running on the caller's stream, and the next call on a different stream (or on the context's own)
 * first waits for it on the device, so calls from different streams never overlap on that scratch.
 */
#ifndef MBG_H
#define MBG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MBG_OK            0
#define MBG_EARG        (-1)   /* invalid argument */
#define MBG_EHIP        (-2)   /* HIP runtime error (detail in mbg_last_hip_error) */
#define MBG_ENOMEM      (-3)   /* device or host allocation failed */
#define MBG_EIO         (-4)   /* index file missing or unreadable */
#define MBG_EFORMAT     (-5)   /* bad magic / inconsistent index file */
#define MBG_ENOIDX      (-6)   /* no index uploaded to the context */
#define MBG_EUNSUPPORTED (-7)  /* job outside the kernels' supported range */
#define MBG_EOVERFLOW   (-8)   /* caller-provided output capacity too small */

typedef struct mbg_ctx_s mbg_ctx_t;

/* SA interval, layout-identical to mb_sai_t (bwt.h:9-12): x[0] forward row, x[1] reverse-complement
 * row, size, info = qs<<32 | qe (half-open query range). */
typedef struct { uint64_t x[2]; uint64_t size, info; } mbg_sai_t;

/* One seeding output record of the device entry point: the interval and the read it belongs to. */
typedef struct { mbg_sai_t p; uint32_t rid, pass; } mbg_seed_rec_t;

/* FM-index in the reference's in-memory layout (mb_bwt_t, bwt.h:14-30; 64-byte occ blocks,
 * bwt.c:60-65).  Host pointers; mbg_bwt_upload copies them to HBM. */
typedef struct {
	uint64_t primary;       /* row of '$' */
	uint64_t L2[5];         /* cumulative base counts, L2[0]=0 */
	uint64_t seq_len;       /* BWT length without '$' */
	const uint64_t *data;   /* occ blocks + packed BWT, data_len = ((n+127)/128)*8 + 4 words */
	uint32_t sa_bit;        /* SA sampled at rows with row % (1<<sa_bit) == 0 */
	uint64_t n_sa;
	const uint64_t *sa;     /* sampled SA, sa[0] = UINT64_MAX */
} mbg_bwt_desc_t;

/* 2-bit reference (l2b_t, l2bit.h:20-32). */
typedef struct {
	uint64_t tot_len, n_ctg;
	const uint64_t *ctg_len;     /* n_ctg contig lengths; offsets are their prefix sums */
	uint64_t n_pac;
	const uint64_t *pac;         /* 32 bases per word, base i at bits 2*(i&31) of pac[i>>5] */
	uint64_t n_ambi;
	const uint64_t *ambi;        /* n_ambi sorted [st,en) pairs of N runs, concatenated coordinates */
} mbg_l2b_desc_t;

typedef struct {
	uint64_t seq_len, primary, tot_len, n_ctg;
	int32_t pre_len, sa_bit;
	uint64_t device_bytes;       /* HBM held by the resident index */
} mbg_idx_info_t;

/* One DP job (the arguments of ksw_extd2_sse / ksw_extz2_sse, ksw2.h:73-80). Sequences are bytes
 * 0..m-1 stored in the batch's concatenated query/target buffers at qoff/toff. */
typedef struct {
	int64_t qoff, toff;
	int32_t qlen, tlen;
	int8_t q, e, q2, e2;          /* q2/e2 ignored by extz2 */
	int32_t w, zdrop, end_bonus, flag;
	int64_t cigar_off;            /* offset (in uint32) of this job's CIGAR slots, capacity qlen+tlen */
} mbg_dp_job_t;

/* ksw_extz_t (ksw2.h:36-45) fields, plus the number of exact-band cells the kernel processed. */
typedef struct {
	int32_t max, zdropped, max_q, max_t, mqe, mqe_t, mte, mte_q, score, reach_end;
	int32_t n_cigar;
	int32_t cells;
} mbg_dp_result_t;

const char *mbg_strerror(int err);
const char *mbg_last_hip_error(void);

/* Context on HIP device `device`.  Allocates a stream and grow-only scratch pools. */
int mbg_init(int device, mbg_ctx_t **ctx);
void mbg_destroy(mbg_ctx_t *ctx);

/* Index residency.  Replaces mb_bwt_load + mb_bwt_cache (bwt.c:585-591,648-683; map-algo.c:19-42)
 * and l2b_load (l2bit.c:275-330).  pre_len is the k-mer interval cache length, 0..16 (10 in the
 * reference, 0 for none; 4^pre_len x 32 bytes of HBM: 8.6 GB at 14, 34 GB at 15, 137 GB at 16),
 * computed on the device.  Seeding results do not depend on it: an interval only shrinks as its
 * pattern grows, so a cached k-mer interval is the state stepwise extension would reach, and a
 * too-small one falls back to stepwise (bwt.c:403-406).  A longer cache skips more dependent
 * steps.  The load also builds the suffix array of every 4th row (rows/4 x 8 bytes, by the same
 * walk mb_bwt_sa_batch takes), which shortens locates without changing their results; it is
 * skipped when the device cannot hold it. */
int mbg_idx_load(mbg_ctx_t *ctx, const char *prefix, int32_t pre_len);
int mbg_bwt_upload(mbg_ctx_t *ctx, const mbg_bwt_desc_t *bwt, int32_t pre_len);
int mbg_l2b_upload(mbg_ctx_t *ctx, const mbg_l2b_desc_t *l2b);
int mbg_idx_info(mbg_ctx_t *ctx, mbg_idx_info_t *info);
/* Use the index resident in src (same device) instead of a copy of one's own, so several contexts
 * (each with its own stream and pools) serve from one index in HBM.  src keeps ownership: it must
 * outlive ctx and must not load another index while ctx uses it.  No counterpart in the reference,
 * whose index is shared by reference between threads (map-algo.c, mb_map_batch). */
int mbg_idx_attach(mbg_ctx_t *ctx, mbg_ctx_t *src);
/* Device address of the k-mer cache (4^pre_len mbg_sai_t), for checking it against the host's. */
int mbg_pre_copy(mbg_ctx_t *ctx, mbg_sai_t *out, uint64_t n);

/* Occurrence rank of all four bases at rows k[i] and l[i]: mb_bwt_rank2a (bwt.c:195-231). */
int mbg_rank2a(mbg_ctx_t *ctx, int64_t n, const uint64_t *k, const uint64_t *l, uint64_t *cntk, uint64_t *cntl);
/* Bidirectional extension of ik[i] by all four bases: mb_bwt_extend (bwt.c:237-250); ok has 4n. */
int mbg_extend(mbg_ctx_t *ctx, int64_t n, const mbg_sai_t *ik, const uint8_t *is_back, mbg_sai_t *ok);

/* Two-pass SMEM seeding of n_seq nt4-encoded reads: mb_seed_intv_batch (seed.c:56-103).
 * On return *out is a malloc'd array (caller free()s it) holding read i's intervals at
 * [off[i], off[i+1]); off has n_seq+1 entries.  Within a read the order is unspecified, as in
 * the reference (the caller sorts them, seed.c:109-138). */
int mbg_seed_batch(mbg_ctx_t *ctx, int32_t n_seq, const int32_t *len, const uint8_t *const *seq,
                   int32_t min_len, int32_t max_sub_occ, mbg_sai_t **out, int64_t *off);

/* In-place, order-preserving SA locate of n rows: mb_bwt_sa_batch (bwt.c:511-541). */
int mbg_sa_batch(mbg_ctx_t *ctx, int64_t n, uint64_t *x);

/* Batched banded DP: ksw_extd2_sse (ksw2_extd2_sse.c:27-403) / ksw_extz2_sse
 * (ksw2_extz2_sse.c:16-296), one result and one CIGAR per job.  m and mat (m*m) are shared by the
 * batch.  CIGARs are written at cigar + jobs[i].cigar_off, oriented as the reference leaves them. */
int mbg_extd2_batch(mbg_ctx_t *ctx, int32_t n_jobs, const mbg_dp_job_t *jobs, const uint8_t *qseq, int64_t qbytes,
                    const uint8_t *tseq, int64_t tbytes, int8_t m, const int8_t *mat,
                    mbg_dp_result_t *res, uint32_t *cigar, int64_t cigar_cap);
int mbg_extz2_batch(mbg_ctx_t *ctx, int32_t n_jobs, const mbg_dp_job_t *jobs, const uint8_t *qseq, int64_t qbytes,
                    const uint8_t *tseq, int64_t tbytes, int8_t m, const int8_t *mat,
                    mbg_dp_result_t *res, uint32_t *cigar, int64_t cigar_cap);

/* The mb_opt_t fields mb_align_pair reads (align.c:329-378): match a, mismatch b, transition
 * mismatch b_ts, the two gap pairs, meth = (opt->flag & MB_F_METH) != 0, and max_sw_mat. */
typedef struct {
	int32_t a, b, b_ts, q, e, q2, e2, meth;
	int64_t max_sw_mat;
} mbg_align_opt_t;

/* One mb_align_pair call: query qseq[qoff, qoff+qlen) against tseq[toff, toff+tlen) (nt4 codes
 * 0..4), with the band w, end_bonus, zdrop (before the reference's scaling by opt->a) and ksw_flag
 * as align.c passes them.  cigar_off: the call's CIGAR slots, capacity qlen+tlen+1. */
typedef struct {
	int64_t qoff, toff;
	int32_t qlen, tlen;
	int32_t w, end_bonus, zdrop, flag;
	int64_t cigar_off;
} mbg_align_job_t;

/* Batched mb_align_pair (align.c:329-390, static there; its callers are mb_align1's extensions and
 * gap fills, align.c:645-735, and mb_align1_inv, :792): the KSW_EZ_GENERIC_SC switch for a
 * transition mismatch or methylation, the ungapped extension fast path (EXTZ_ONLY and tlen >= qlen:
 * at most 2 mismatches and the end within end_bonus of the maximum), the ungapped global fast path
 * (qlen == tlen: at most 3 mismatches or a score above the best gapped one), the band cap from the
 * mismatch count when the shorter side is under 100, the max_sw_mat skip (zdropped set, no CIGAR),
 * and the DP the rest reach: ksw_extz2 when q == q2 and e == e2, else ksw_extd2, with zdrop * a.
 * mat is the 5x5 nt4 matrix the caller built (ksw_gen_nt4_mat, ksw2.h:199-219).  res[i] holds the
 * ksw_extz_t fields the call leaves (cells: the DP's exact-band cells, 0 when no DP ran); its CIGAR
 * is at cigar + jobs[i].cigar_off. */
int mbg_align_pair_batch(mbg_ctx_t *ctx, const mbg_align_opt_t *opt, int32_t n, const mbg_align_job_t *jobs, const uint8_t *qseq,
                         int64_t qbytes, const uint8_t *tseq, int64_t tbytes, const int8_t *mat, mbg_dp_result_t *res,
                         uint32_t *cigar, int64_t cigar_cap);

/* One ksw_ll job: query qseq[qoff, qoff+qlen), target tseq[toff, toff+tlen), bytes 0..m-1.  size
 * selects the core as ksw_ll_qinit's size does: 1 = ksw_ll_u8_core, 2 = ksw_ll_i16_core. */
typedef struct {
	int64_t qoff, toff;
	int32_t qlen, tlen;
	int32_t size, pad;
} mbg_ll_job_t;

/* ksw_llrst_t (ksw2.h:47-51): best score, its target and query end, second best and its end. */
typedef struct { int32_t score, te, qe, score2, te2; } mbg_ll_result_t;

/* Batched striped local Smith-Waterman as mate rescue runs it (pe.c:260-272): per job,
 * ksw_ll_qinit (ksw2_ll_sse.c:49-96) then ksw_ll_u8_core or ksw_ll_i16_core (ksw2_ll_sse.c:121-333;
 * declared ksw2.h:101-103), results bit-identical to them.  m, mat (m*m), gapo, gape and xtra
 * (KSW_LL_SUBO / KSW_LL_STOP | score) are shared by the batch.  Supported range: gapo >= 1 (with
 * gapo == 0 the reference's lazy-F exit depends on its SSE lane order), gapo + gape <= 255,
 * qlen <= 1024, a matrix with a positive entry; u8 jobs need a negative entry and
 * max - min <= 255, i16 jobs max(mat) * min(qlen, tlen) <= 32767 (no saturation).  Outside it:
 * MBG_EUNSUPPORTED, and no job runs. */
int mbg_ll_batch(mbg_ctx_t *ctx, int32_t n_jobs, const mbg_ll_job_t *jobs, const uint8_t *qseq, int64_t qbytes,
                 const uint8_t *tseq, int64_t tbytes, int8_t m, const int8_t *mat, int32_t gapo, int32_t gape, int32_t xtra,
                 mbg_ll_result_t *res);

/* One mate-rescue candidate of mb_matesw_core (pe.c:319-378): the mate's bases as they would align
 * (reverse-complemented for a reverse-strand candidate, pe.c:345) at qseq[qoff, qoff+qlen), and
 * the window [ts, te) on contig tid where the insert-size model puts it (pe.c:338-339; clipped to
 * the contig here as pe.c:340-341 does). */
typedef struct {
	int64_t qoff;
	int32_t qlen, tid;
	int64_t ts, te;
	int32_t min_sc, pad;          /* mb_matesw's min_sc, already divided by a */
	int64_t cigar_off;            /* CIGAR slots at cigar + cigar_off, capacity qlen + te - ts */
} mbg_matesw_job_t;

/* The mb_opt_t fields mate rescue reads (pe.c:249-312); mt is the methylation conversion of the
 * batch (0 none, 1 C->T, 2 G->A) after l2b_meth_rev for reverse-strand candidates. */
typedef struct {
	int32_t a, b, b_ts, b_ambi, q, e, bw, zdrop, end_bonus, min_dp_max, min_len, mt;
} mbg_matesw_opt_t;

/* Per candidate: the vote, the window mb_matesw_align saw, and its results before mb_update_extra. */
typedef struct {
	int32_t max_ug, max_i, n_good, n_kmer;   /* mb_ungap over the window (pe.c:209-247) */
	int32_t ts2, te2;                        /* aligned window, relative to the clipped ts (pe.c:343-348) */
	int32_t aligned;                         /* the vote passed: mb_matesw_align ran (pe.c:349) */
	int32_t rescued;                         /* a hit was made (pe.c:293) */
	mbg_ll_result_t ll;                      /* its ksw_ll; all -1 when it did not run */
	mbg_dp_result_t ez;                      /* its ksw_extz2 on the reversed prefixes; zero when not run */
	int32_t qs, qe, ts, te;                  /* the rescued hit as mb_update_extra leaves it, relative to the
	                                            aligned window (pe.c:295-296,309) */
	int32_t dp_max2;                         /* pe.c:298-299 */
	int32_t ran_dp;                          /* the extension DP ran (pe.c:273) */
	int32_t n_cigar;                         /* the rescued hit's CIGAR after mb_fix_cigar */
	int32_t blen, mlen, n_ambi, dp_max;      /* mb_update_extra's (dp_max0 = dp_max; dp_score = ez.max) */
} mbg_matesw_result_t;

/* Batched mate rescue on the resident reference: per candidate, l2b_getseq of the window, mb_ungap,
 * the narrowing to a clear diagonal and mb_matesw_align (pe.c:342-349, 249-312): ksw_ll, ksw_extz2
 * of the reversed prefixes, and for a rescued hit mb_fix_cigar + mb_update_extra's scores.  Results
 * equal the reference's.  The host keeps the rest of mb_matesw_core (candidate windows,
 * orientation, the hit's coordinates on the contig) and the cs/MD strings.  A window
 * not longer than the read after clipping is no candidate (aligned 0).  Supported: qlen <= 1024,
 * windows <= 16384 bases, q/a >= 1 (ksw_ll); else MBG_EUNSUPPORTED.  Candidate i's CIGAR is at
 * cigar + jobs[i].cigar_off: the rescued hit's (n_cigar entries), else the extension DP's as
 * ksw_extz2 leaves it (ez.n_cigar entries).  stats
 * (NULL: not wanted) receives counts and the device time of each phase. */
typedef struct {
	int64_t n_aligned, n_rescued;
	int64_t ll_cells;             /* ksw_ll: read length x window rows scanned */
	int64_t dp_cells;             /* extension DP: band cells */
	float ms_window, ms_ungap, ms_ll, ms_dp, ms_finish, ms_total;   /* device time per phase */
} mbg_matesw_stats_t;
int mbg_matesw_batch(mbg_ctx_t *ctx, int32_t n_jobs, const mbg_matesw_job_t *jobs, const uint8_t *qseq, int64_t qbytes,
                     const mbg_matesw_opt_t *opt, mbg_matesw_result_t *res, uint32_t *cigar, int64_t cigar_cap,
                     mbg_matesw_stats_t *stats);

/* One hit to finish: its CIGAR and the bases it aligns. */
typedef struct {
	int64_t qoff, toff;           /* the hit's first query and target bases: qseq + qoff, tseq + toff */
	int32_t qs, qe;               /* in; out: a dropped leading insertion moves qs (qe when rev) */
	int64_t ts, te;               /* in; out: a dropped leading deletion moves ts */
	int32_t rev, n_cigar;         /* mb_hit_t.rev; n_cigar in and out */
	int64_t cigar_off;            /* the CIGAR at cigar + cigar_off, edited in place */
	int32_t blen, mlen, n_ambi, dp_max;   /* out: mb_hit_t.blen/mlen, mb_extra_t.n_ambi, dp_max (= dp_max0) */
} mbg_extra_job_t;

/* Batched mb_fix_cigar (align.c:70-142) + the scores of mb_update_extra (align.c:219-258; decl
 * mbpriv.h:98): indels left-aligned, adjacent I/D runs merged, zero-length runs dropped, a leading
 * gap removed, then blen, mlen, the ambiguous-base count and the best prefix score, with gaps at
 * q+e (log_gap 0) or q+e*mb_log2(1+len) (log_gap 1).  mat is the 5x5 matrix.  The =/X, cs and MD
 * strings stay with the host.  Results equal the reference's. */
int mbg_update_extra_batch(mbg_ctx_t *ctx, int64_t n, mbg_extra_job_t *jobs, const uint8_t *qseq, int64_t qbytes,
                           const uint8_t *tseq, int64_t tbytes, const int8_t *mat, int8_t q, int8_t e, int32_t log_gap,
                           uint32_t *cigar, int64_t cigar_cap);

/* mb_anchor_t (mbpriv.h:26-32): sid = contig<<1 | reverse strand; len bases ending at query position
 * qpos (on the anchor's strand) and at contig position tpos; flag is 0 on output (its top bit is
 * mb_anchor_t::flt). */
typedef struct { int32_t sid, len, qpos; uint32_t flag; int64_t tpos; } mbg_anchor_t;

/* Anchors of n_reads reads from their SMEM intervals: mb_anchor (seed.c:268-339; decl mbpriv.h:58)
 * for reads without a methylation conversion (mt L2B_METH_NONE): mb_seed_sort_dedup
 * (seed.c:109-138), the SA of every row of an interval group of size <= max_occ or of max_occ evenly
 * stepped rows (seed.c:296-318), the contig and strand of each hit (l2b_intv2cid, l2bit.c:23-34),
 * klib's radix sort by concatenated position (seed.c:331) and mb_anchor_dedup (seed.c:140-166).
 * Read r's intervals are intv[intv_off[r] .. intv_off[r+1]) (intv_off[0] = 0); they come back sorted
 * and deduplicated as the reference leaves its mb_sai_v, n_intv[r] of them (n_intv may be null).  Its
 * anchors are written to anc[anc_off[r] ...] (anc_off[0] = 0), n_anc[r] of them, in the reference's
 * order; anc_off[r+1] - anc_off[r] must be at least the sum of min(size, max_occ) over read r's
 * intervals (the reference's preallocation, seed.c:292-293).  The FM-index and the l2b must be
 * resident.  Results equal the reference's. */
int mbg_anchor_batch(mbg_ctx_t *ctx, int32_t n_reads, mbg_sai_t *intv, const int64_t *intv_off, const int32_t *qlen,
                     int32_t max_occ, mbg_anchor_t *anc, const int64_t *anc_off, int64_t *n_anc, int64_t *n_intv);

/* The parameters of mb_lchain_dp (lchain.c:145; decl mbpriv.h:62-63): max_dist_x/max_dist_y (raised
 * to bw when smaller), bw, max_skip, max_iter, min_sc, and pen_gap = chn_pen_gap, which the mapper
 * sets to chain_gap_scale * .01 * min_len (map-algo.c:607). */
typedef struct {
	int32_t max_dist_x, max_dist_y, bw, max_skip, max_iter, min_sc;
	float pen_gap;
	int32_t pad;
} mbg_chain_opt_t;

/* Chaining of n_reads reads' anchors: mb_lchain_dp (lchain.c:145-240) -- the DP with its skip and
 * iteration limits, the backtrack (lchain.c:44-92) and the compaction (lchain.c:94-123).  Read r's
 * anchors are anc[anc_off[r] .. anc_off[r] + n_anc[r]) in mb_anchor's order (anc_off has n_reads+1
 * entries, anc_off[0] = 0, n_anc[r] <= anc_off[r+1] - anc_off[r], n_anc[r] < 2^31 as lchain.c's int32
 * indices require).  Its chains (score<<32 | anchor
 * count, ordered by the first anchor's concatenated position) go to u[anc_off[r] ...], n_u[r] of them,
 * and their anchors, first to last, to out[anc_off[r] ...], n_out[r] of them.  Anchors' contigs must
 * be those of the resident l2b.  Results equal the reference's. */
int mbg_lchain_batch(mbg_ctx_t *ctx, int32_t n_reads, const mbg_anchor_t *anc, const int64_t *anc_off, const int64_t *n_anc,
                     const mbg_chain_opt_t *opt, uint64_t *u, int32_t *n_u, mbg_anchor_t *out, int64_t *n_out);

/* ---- the mapper's front pass: seeding -> anchors -> chains of one mini-batch -------------------- */

/* Options of mbg_map_front: seeding's min_len and max_sub_occ (mb_seed_intv_batch), mb_anchor's
 * max_occ, and n_copt sets of mb_lchain_dp parameters, read r using copt[copt_idx[r]] (the mapper
 * derives them from mb_opt_adap for each read length, map-algo.c:606-612).  no_chain (may be NULL):
 * reads r with no_chain[r] != 0 are seeded only (methylation-converted reads, whose anchors the
 * reference builds with a conversion-aware mb_anchor). */
typedef struct {
	int32_t min_len, max_sub_occ, max_occ, n_copt;
	const mbg_chain_opt_t *copt;
	const int32_t *copt_idx;
	const uint8_t *no_chain;
} mbg_front_opt_t;

/* Results of mbg_map_front, in host memory the context owns (pinned), valid until its next
 * mbg_map_front call.  Read r:
 *   intervals  intv[intv_off[r] .. + n_intv[r]): mb_seed_intv_batch's set for the read, sorted and
 *              deduplicated as mb_seed_sort_dedup leaves it (seed.c:109-138);
 *   anchors    anc[anc_off[r] .. + n_anc[r]): mb_anchor's result in its order (n_anc -1 for no_chain);
 *   chains     u[anc_off[r] .. + n_u[r]) and chn[anc_off[r] .. + n_chn[r]): mb_lchain_dp's chain list
 *              (score<<32 | anchor count, by first anchor's position) and chained anchors. */
typedef struct {
	int64_t n_rec, n_slots;
	const int64_t *intv_off, *n_intv;
	const mbg_sai_t *intv;
	const int64_t *anc_off, *n_anc;
	const mbg_anchor_t *anc;
	const int32_t *n_u;
	const uint64_t *u;
	const int64_t *n_chn;
	const mbg_anchor_t *chn;
	float ms_seed, ms_group, ms_anchor, ms_chain, ms_copy, ms_total;   /* device time per phase */
	float ms_seed1, ms_seed2;            /* SMEM pass 1 kernel; pass 2 with its task step and compaction */
	uint64_t seed1_occ_blocks, seed1_kmer_lookups;   /* pass 1's work, while mbg_count_work is on (else 0) */
} mbg_front_t;

/* The front of mb_map_sai (map-algo.c:588-631) for n_reads reads whose nt4 bases are at
 * seq + off[r], len[r] of them (host memory): two-pass seeding (mb_seed_intv_batch, seed.c:56-103),
 * mb_anchor (seed.c:268-339) and the first mb_lchain_dp (lchain.c:145-231), in one device pass with
 * no host step between them.  Results equal the reference's. */
int mbg_map_front(mbg_ctx_t *ctx, int32_t n_reads, const uint8_t *seq, const int64_t *off, const int32_t *len,
                  const mbg_front_opt_t *opt, mbg_front_t *out);

/* Cumulative device time (HIP events on the context's stream, in ms) and work of the context's
 * calls since mbg_init or the last reset: mbg_map_front's phases, the DP batches (mbg_extd2_batch /
 * mbg_extz2_batch: kernel time and the band cells processed) and the local-SW batches (mbg_ll_batch).
 * seed1_occ_blocks / seed1_kmer_lookups (the algorithmic bytes of SMEM pass 1: 64 B per occ block
 * fetched, 32 B per k-mer cache lookup) are counted only while mbg_count_work(ctx, 1) is on; counting
 * adds atomics to the seeding kernels, so timed runs leave it off.  No counterpart in the reference. */
typedef struct {
	int64_t fronts, front_reads, front_records, front_slots;
	double ms_seed1, ms_seed2, ms_group, ms_anchor, ms_chain, ms_front_copy, ms_front;
	uint64_t seed1_occ_blocks, seed1_kmer_lookups;
	int64_t dp_launches, dp_jobs, dp_cells;
	double ms_dp;
	int64_t ll_launches, ll_jobs, ll_cells;
	double ms_ll;
	/* launch chains run as captured HIP graphs, how many of those needed a new executable graph
	 * (the rest updated the last one in place), and chains launched one by one (graphs off) */
	int64_t graph_launches, graph_instantiations, direct_chains;
} mbg_counters_t;
int mbg_counters(mbg_ctx_t *ctx, mbg_counters_t *out, int reset);
int mbg_count_work(mbg_ctx_t *ctx, int on);
/* Launch chains without a host step inside (the seeding phases around SMEM pass 1; the whole
 * seed -> locate -> extend pass of mbg_map_step_dev) run as HIP graphs captured from the stream and
 * updated in place from call to call when on (off by default; MBG_GRAPH=1 in the environment of
 * mbg_init turns it on).  Results do not depend on it.  While a context captures, no other thread of
 * the process should allocate device memory or create streams (the capture fails with MBG_EHIP).
 * With graphs on, mbg_map_step_dev reports only its total device time, not its phases.  No
 * counterpart in the reference. */
int mbg_use_graphs(mbg_ctx_t *ctx, int on);

/* ---- reads at the edge: FASTA/FASTQ (host code) ------------------------------------------------ */

/* A reader of one FASTA/FASTQ file, plain or gzip ("-" or NULL: standard input): mb_bseq_open
 * (bseq.c:41-52).  A read-ahead thread decompresses while the caller parses; n_threads (1-64) encode
 * each chunk's bases.  One reader is used by one thread at a time. */
typedef struct mbg_fq_s mbg_fq_t;
/* Storage of one chunk, reused from call to call (the views below point into it). */
typedef struct mbg_fq_buf_s mbg_fq_buf_t;
/* One chunk: record i's name is name[name_off[i] .. name_off[i+1]), its comment comment[com_off[i] ..
 * com_off[i+1]) (empty when not kept), its bases seq[seq_off[i] .. seq_off[i+1]) (U/u turned into T/t as
 * kseq2bseq does, bseq.c:64-78), their nt4 codes (kom_nt4_table, kommon.c:162-179) at the same
 * offsets of nt4, and its quality at the same offsets of qual when has_qual[i].  status: 0, or -2 when
 * a record with a missing or short quality string ended the chunk (kseq_read's -2; the reference
 * warns and continues), or MBG_EIO on a decompression error. */
typedef struct {
	int32_t n, status;
	const int64_t *name_off, *com_off, *seq_off;
	const char *name, *comment, *seq, *qual;
	const uint8_t *has_qual, *nt4;
} mbg_fq_chunk_t;
int mbg_fq_open(const char *path, int32_t n_threads, mbg_fq_t **fq);
void mbg_fq_close(mbg_fq_t *fq);                                     /* mb_bseq_close (bseq.c:54-59) */
int mbg_fq_buf_new(mbg_fq_buf_t **buf);
void mbg_fq_buf_free(mbg_fq_buf_t *buf);
/* The next chunk: mb_bseq_read (bseq.c:85-131) -- records until chunk_size bases and min_cnt records,
 * or max_chunk_size bases; in frag_mode the records after that whose names equal the last one's
 * (mb_qname_same, bseq.h:26-39) join it, and the first that does not opens the next chunk.  n = 0 at
 * the end of the input. */
int mbg_fq_read(mbg_fq_t *fq, int64_t chunk_size, int32_t with_qual, int32_t with_comment, int32_t frag_mode, int32_t min_cnt,
                int64_t max_chunk_size, mbg_fq_buf_t *buf, mbg_fq_chunk_t *chunk);
/* mb_bseq_read_frag (bseq.c:133-161): one record from each of n_fp files per step, interleaved, until
 * chunk_size bases or until a file ends. */
int mbg_fq_read_frag(int32_t n_fp, mbg_fq_t **fq, int64_t chunk_size, int32_t with_qual, int32_t with_comment, mbg_fq_buf_t *buf,
                     mbg_fq_chunk_t *chunk);

/* Target extraction l2b_getseq (l2bit.c:52-77) for n windows [st[i],en[i]) of contig tid[i]
 * (clipped to the contig, N runs restored as 4), written at out + out_off[i].  len[i] receives
 * the clipped length (or -1 for a bad tid). */
int mbg_getseq(mbg_ctx_t *ctx, int64_t n, const int64_t *tid, const int64_t *st, const int64_t *en,
               const int64_t *out_off, uint8_t *out, int64_t out_bytes, int64_t *len);

/* ---- device-resident entry points (all pointers are device addresses) ---------------------- */

/* Seeding over reads already in HBM: seq bytes at seq+off[i], length len[i].  Records go to
 * out[0..cap); *n_out (device int64) receives the total count (> cap means overflow: records past
 * cap were dropped and MBG_EOVERFLOW is returned).  Synchronises the stream before returning. */
int mbg_seed_batch_dev(mbg_ctx_t *ctx, void *stream, int32_t n_seq, const int64_t *off, const int32_t *len,
                       const uint8_t *seq, int32_t min_len, int32_t max_sub_occ,
                       mbg_seed_rec_t *out, int64_t cap, int64_t *n_out_host);
/* Same as mbg_seed_batch_dev but asynchronous: *n_out_dev is a device int64 the caller reads
 * later; pass-2 task generation happens on the device. */
int mbg_seed_batch_dev_async(mbg_ctx_t *ctx, void *stream, int32_t n_seq, const int64_t *off, const int32_t *len,
                             const uint8_t *seq, int32_t min_len, int32_t max_sub_occ,
                             mbg_seed_rec_t *out, int64_t cap, int64_t *n_out_dev);
int mbg_sa_batch_dev(mbg_ctx_t *ctx, void *stream, int64_t n, uint64_t *x);
/* mbg_sa_batch_dev without any allocation: no counterpart in the reference.  It is for callers that
 * cannot allocate device memory -- code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The kernel's work-queue counter lives in a 256-byte array of static device memory
 * (one per GPU, held until the process ends) instead of the context's pools; nothing is allocated,
 * freed or set by hipMemset.  Up to MBG_SA_NOALLOC_MAX_ROWS rows (else MBG_EUNSUPPORTED).  All calls on
 * one GPU share that array: calls that overlap in time on one GPU (two streams, or two contexts on the
 * same device) race, so a caller serialises them.  Asynchronous on the stream, like mbg_sa_batch_dev.
 * On the device entry points a row outside the index (above seq_len) comes back as 2^64-1. */
#define MBG_SA_NOALLOC_MAX_ROWS ((int64_t)1 << 31)
int mbg_sa_batch_dev_noalloc(mbg_ctx_t *ctx, void *stream, int64_t n, uint64_t *x);
/* mbg_anchor_batch over buffers in HBM: intv_off, anc_off (n_reads+1 each, host-computed as there,
 * anc_off[n_reads] = n_slots), qlen, n_anc and n_intv (n_reads each, both required) and rows, a
 * workspace of n_slots uint64.  Every row x[0] + size - 1 of every interval must be a row of the
 * resident index.  A read whose slots do not suffice gets n_anc -1.  Asynchronous on the stream. */
int mbg_anchor_batch_dev(mbg_ctx_t *ctx, void *stream, int32_t n_reads, int64_t n_slots, mbg_sai_t *intv,
                         const int64_t *intv_off, const int32_t *qlen, int32_t max_occ, uint64_t *rows, mbg_anchor_t *anc,
                         const int64_t *anc_off, int64_t *n_anc, int64_t *n_intv);
/* Workspace bytes mbg_lchain_batch_dev needs for n_slots anchor slots. */
#define MBG_LCHAIN_WORK_BYTES(n_slots) ((size_t)(n_slots) * 72 + 2048)
/* mbg_lchain_batch over buffers in HBM (anc_off[n_reads] = n_slots) with a caller's workspace of
 * work_bytes >= MBG_LCHAIN_WORK_BYTES(n_slots).  Asynchronous on the stream. */
int mbg_lchain_batch_dev(mbg_ctx_t *ctx, void *stream, int32_t n_reads, int64_t n_slots, const mbg_anchor_t *anc,
                         const int64_t *anc_off, const int64_t *n_anc, const mbg_chain_opt_t *opt, void *work, size_t work_bytes,
                         uint64_t *u, int32_t *n_u, mbg_anchor_t *out, int64_t *n_out);
/* mbg_ll_batch over jobs, sequences and results in HBM.  Every job is run with the registers of
 * max_qlen (<= 1024); rows_total bounds the sum of the jobs' tlen (the per-row maxima lists).  A job
 * outside the supported range (qlen > max_qlen, an i16 job that could saturate) gets score -1 and
 * its other fields -1; the matrix must suit the u8 core (a negative entry, max - min <= 255).
 * Synchronises the stream. */
int mbg_ll_batch_dev(mbg_ctx_t *ctx, void *stream, int32_t n_jobs, const mbg_ll_job_t *jobs, const uint8_t *qseq,
                     const uint8_t *tseq, int8_t m, const int8_t *mat_host, int32_t gapo, int32_t gape, int32_t xtra,
                     mbg_ll_result_t *res, int32_t max_qlen, int64_t rows_total);
/* mbg_extd2_batch over device-resident jobs, sequences, results and CIGAR slots, asynchronous on the
 * stream.  The launch is planned for jobs with qlen <= max_qlen, tlen <= max_tlen and w <= max_w
 * (backtrack slabs and row rings are sized from them); a job outside that plan is not run: its result
 * is an empty job's with cells -1. */
int mbg_extd2_batch_dev(mbg_ctx_t *ctx, void *stream, int32_t n_jobs, const mbg_dp_job_t *jobs,
                        const uint8_t *qseq, const uint8_t *tseq, int8_t m, const int8_t *mat_host,
                        mbg_dp_result_t *res, uint32_t *cigar, int32_t max_qlen, int32_t max_tlen, int32_t max_w);

/* Device memory accounting for admission control: *bytes receives the HBM held by the work buffers
 * of every context of this process (they grow to the largest batch seen and are kept; the resident
 * index is not counted), *dev_free and *dev_total the current device's free and total memory (each
 * pointer may be null).  No counterpart in the reference, which has no device: for the serving loop's
 * memory-gated admission. */
int mbg_work_bytes(int64_t *bytes, int64_t *dev_free, int64_t *dev_total);

/* Device memory for callers that keep inputs in HBM without another runtime: allocate, free,
 * and synchronous copies in either direction.  A copy first waits for the work queued on the
 * context's own stream (entry points called with stream == NULL); work the caller queued on a
 * stream of its own is the caller's to wait for. */
int mbg_dev_alloc(mbg_ctx_t *ctx, size_t bytes, void **ptr);
int mbg_dev_free(mbg_ctx_t *ctx, void *ptr);
int mbg_copy_to_device(mbg_ctx_t *ctx, void *dst_dev, const void *src, size_t bytes);
int mbg_copy_to_host(mbg_ctx_t *ctx, void *dst, const void *src_dev, size_t bytes);
/* Page-locked host memory, for staging buffers that the host-pointer entry points copy from and to
 * at full PCIe rate. */
int mbg_host_alloc(mbg_ctx_t *ctx, size_t bytes, void **ptr);
int mbg_host_free(mbg_ctx_t *ctx, void *ptr);

/* A plain streaming copy kernel (16 B per lane per iteration), used to calibrate HBM byte
 * counters: its FETCH/WRITE counters are the reference for `bytes` read and written. */
int mbg_stream_copy(mbg_ctx_t *ctx, void *stream, void *dst_dev, const void *src_dev, size_t bytes);

/* Random-access ceiling of the resident occ table: n_chains independent lanes each fetch `steps`
 * 64-byte blocks in sequence, the next block index depending on the bytes just loaded (the access
 * pattern of one rank per extension step).  Reports the kernel time from HIP events. */
int mbg_gather_bench(mbg_ctx_t *ctx, int64_t n_chains, int32_t steps, int32_t block_threads, float *ms, uint64_t *checksum);

/* ---- seed -> locate -> extend step over device-resident reads --------------------------------- */

/* Options of one step; defaults are minibwa's (options.c:5-46) with the adaptive band/z-drop for the
 * batch's read length (options.c:90-116): ksw_w = (int)(bw*1.5+1) (align.c:564), zdrop = zdrop*a. */
typedef struct {
	int32_t min_len, max_sub_occ, max_occ;
	int32_t a, b, q, e, q2, e2, end_bonus;
	int32_t ksw_w, zdrop;
	int8_t mat[25];                /* ksw_gen_nt4_mat(a, b, b_ts, b_ambi) */
} mbg_map_opt_t;

/* Per extension job: which read, which end (0 left, 1 right), strand, contig and anchor position. */
typedef struct { uint32_t rid; int8_t side, rev; int16_t pad; int32_t tid; int64_t tpos; } mbg_job_meta_t;

typedef struct {
	int64_t n_reads, n_rec, n_rows, n_anchored, n_fast, n_jobs;
	int64_t q_bytes, t_bytes, cells;
	/* algorithmic work, filled when collect_stats != 0: 64-byte occ blocks fetched and 32-byte
	 * k-mer cache lookups by seeding pass 1 / pass 2, extension steps, LF steps of SA locate */
	uint64_t occ_blocks[2], kmer_lookups[2], extend_steps[2], lf_steps;
	/* device time per phase from HIP events on the step's stream (ms) */
	float ms_seed1, ms_seed2, ms_rows, ms_sa, ms_jobs, ms_dp, ms_total;
} mbg_map_stats_t;

/* Device addresses owned by the context, valid until its next call. */
typedef struct {
	const mbg_seed_rec_t *recs;
	const uint64_t *pos;           /* located SA positions, n_rows */
	const uint32_t *pos_rec;       /* record index of each position */
	const mbg_dp_job_t *jobs;
	const mbg_job_meta_t *meta;
	const mbg_dp_result_t *res;
	const uint32_t *cigar;
	const uint8_t *qbuf, *tbuf;
} mbg_map_out_t;

/* One pass of the hot path over n_reads reads in HBM (seq+off[i], len[i] <= max_len): two-pass
 * seeding, SA locate of every interval's anchor rows (seed.c:304-315), and for each read the left
 * and right extension from its longest SMEM (align.c mb_align1/mb_align_pair rules; ungapped fast
 * path first, ksw_extd2 -- or ksw_extz2 when q==q2 and e==e2 -- for the rest).  Synchronises the
 * stream; grows the context's pools and repeats the step if a pool overflowed.  A read longer than
 * max_len overflows the job pools sized from it: its jobs run empty and the step returns
 * MBG_EOVERFLOW (the other results stand). */
int mbg_map_step_dev(mbg_ctx_t *ctx, void *stream, int32_t n_reads, const int64_t *off, const int32_t *len, const uint8_t *seq,
                     int32_t max_len, const mbg_map_opt_t *opt, int collect_stats, mbg_map_stats_t *stats, mbg_map_out_t *out);

#ifdef __cplusplus
}
#endif
#endif
