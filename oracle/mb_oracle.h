/*
 * mb_oracle.h -- CPU restatement of minibwa 0.7-r421's seeding + extension hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP library is compared against; it is
 * linked only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
 * the product (minibwa_amd/libmbg.so).  Every function cites the reference file:line it restates.
 */
#ifndef MB_ORACLE_H
#define MB_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t x[2], size, info; } orc_sai_t;   /* bwt.h:9-12 */

typedef struct {                                            /* bwt.h:14-30 (the fields the path reads) */
	uint64_t primary, L2[5], seq_len, data_len;
	uint64_t *data;
	uint32_t sa_bit;
	uint64_t n_sa;
	uint64_t *sa;
	int32_t pre_len;
	orc_sai_t *pre;
} orc_bwt_t;

typedef struct {                                            /* ksw2.h:36-45, bitfields widened */
	int32_t max, zdropped;
	int32_t max_q, max_t, mqe, mqe_t, mte, mte_q, score, reach_end;
	int32_t n_cigar;
} orc_ez_t;

orc_bwt_t *orc_bwt_load(const char *fn);
void orc_bwt_destroy(orc_bwt_t *b);
void orc_bwt_cache(orc_bwt_t *b, int32_t len);
void orc_rank1a(const orc_bwt_t *b, uint64_t k, uint64_t cnt[4]);
void orc_rank2a(const orc_bwt_t *b, uint64_t k, uint64_t l, uint64_t ck[4], uint64_t cl[4]);
void orc_extend(const orc_bwt_t *b, const orc_sai_t *ik, orc_sai_t ok[4], int is_back);
int64_t orc_smem(const orc_bwt_t *b, uint32_t len, const uint8_t *q, int64_t x, int64_t min_len, int64_t min_occ, orc_sai_t *p);
int64_t orc_seed_intv(const orc_bwt_t *b, int32_t len, const uint8_t *seq, int32_t min_len, int32_t max_sub_occ, orc_sai_t *out, int64_t cap);
uint64_t orc_sa(const orc_bwt_t *b, uint64_t k);
void orc_sa_batch(const orc_bwt_t *b, int64_t n, uint64_t *x);

int orc_extd2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
              int8_t q, int8_t e, int8_t q2, int8_t e2, int w, int zdrop, int end_bonus, int flag,
              orc_ez_t *ez, uint32_t *cigar, int cigar_cap);
int orc_extz2(int qlen, const uint8_t *query, int tlen, const uint8_t *target, int8_t m, const int8_t *mat,
              int8_t q, int8_t e, int w, int zdrop, int end_bonus, int flag,
              orc_ez_t *ez, uint32_t *cigar, int cigar_cap);
int64_t orc_dp_cells(int qlen, int tlen, int w);
int orc_align_pair(const int32_t p[8], int64_t max_sw_mat, int qlen, const uint8_t *qseq, int tlen, const uint8_t *tseq,
                   const int8_t *mat, int w, int end_bonus, int zdrop, int flag, orc_ez_t *ez, uint32_t *cigar, int cigar_cap);

typedef struct { int32_t score, te, qe, score2, te2; } orc_llrst_t;   /* ksw_llrst_t, ksw2.h:47-51 */
int orc_ll(int size, int qlen, const uint8_t *query, int m, const int8_t *mat, int tlen, const uint8_t *target,
           int gapo, int gape, int xtra, orc_llrst_t *r);

typedef struct {               /* mb_matesw_align's results before mb_update_extra */
	int32_t ran_ll, ran_dp, rescued;
	orc_llrst_t ll;
	orc_ez_t ez;
	int32_t qs, qe, ts, te, dp_max2;    /* window-relative */
} orc_rescue_t;
int32_t orc_ungap(int32_t qlen, const uint8_t *qseq, int32_t tlen, const uint8_t *tseq, int32_t kmer, int32_t mt,
                  int32_t *max_i, int32_t *n_good, int32_t *n_kmer);
int orc_matesw_align(const int32_t p[11], int32_t qlen, const uint8_t *qseq, int32_t tlen, const uint8_t *tseq, int32_t min_sc,
                     int32_t mt, orc_rescue_t *r, uint32_t *cigar, int cigar_cap);
typedef struct {               /* mb_anchor_t (mbpriv.h:26-32) */
	int32_t sid, len, qpos;
	uint32_t flag:31, flt:1;
	int64_t tpos;
} orc_anchor_t;
int64_t orc_anchor(const orc_bwt_t *b, int64_t n_ctg, const uint64_t *ctg_off, const uint64_t *ctg_len, orc_sai_t *u, int64_t n_u,
                   int32_t qlen, int32_t max_occ, orc_anchor_t *v, int64_t cap);
int32_t orc_lchain(int64_t n_ctg, const uint64_t *ctg_off, const uint64_t *ctg_len, int32_t max_dist_x, int32_t max_dist_y,
                   int32_t bw, int32_t max_skip, int32_t max_iter, int32_t min_sc, float pen_gap, int64_t n, const orc_anchor_t *a,
                   uint64_t *u_out, orc_anchor_t *b_out, int64_t *n_b);

typedef struct {               /* the mb_hit_t / mb_extra_t fields mb_fix_cigar and mb_update_extra set */
	int32_t qs, qe, rev, n_cigar;
	int64_t ts, te;
	int32_t blen, mlen, n_ambi, dp_max;
} orc_extra_t;
void orc_update_extra(orc_extra_t *r, uint32_t *cigar, const uint8_t *qseq, const uint8_t *tseq, const int8_t *mat, int8_t q, int8_t e,
                      int log_gap);
int orc_matesw_cand(const int32_t p[11], int32_t qlen, const uint8_t *qseq, int32_t wlen, const uint8_t *win, int32_t min_sc,
                    int32_t mt, int32_t ug[4], int32_t win2[2], int32_t *aligned, orc_rescue_t *r, uint32_t *cigar, int cigar_cap);
uint64_t orc_bench(const orc_bwt_t *b, int type, int64_t n, int intv);

#ifdef __cplusplus
}
#endif
#endif
