// mbg_internal.h -- types shared by the kernels and the C-ABI layer of libmbg.so
#pragma once
#include "mbg_dev.h"

namespace mbg {

// one seeding output record; layout-identical to mbg_seed_rec_t
struct SeedRec { uint64_t x0, x1, size, info; uint32_t rid, pass; };
struct SeedTask { uint32_t rid; int32_t st, en, min_len, min_occ; };

struct SeedArgs {
	const uint8_t *seq;     // concatenated nt4 reads
	const int64_t *off;     // read i at seq + off[i]
	const int32_t *len;
	const uint32_t *packed; // read i's bases as 2-bit codes at packed + i*kPackWords (16 a word), or null
	const uint32_t *qflag;  // read i: nonzero when packed[] does not hold it (an N, or longer than 384)
	const SeedTask *tasks;  // pass 2 only
	uint32_t n_tasks;
	const uint32_t *n_tasks_dev;  // pass 2: task count produced on the device (clamped to task_cap)
	int64_t task_cap;
	int32_t min_len;
	uint32_t *next;         // task counter (zeroed before launch)
	// output: task t's records at slots[t*kslot ...], slot_cnt[t] of them; beyond kslot they go to
	// the spill list (one atomic each)
	SeedRec *slots;
	uint32_t *slot_cnt;
	int32_t kslot;
	SeedRec *spill;
	unsigned long long *n_spill;
	int64_t spill_cap;
	unsigned long long *stats;  // optional: [0] occ blocks fetched, [1] k-mer cache lookups, [2] extend steps
	// segmented pass 1 for reads of at least 2*seg_len bases (seg_len 0: off).  Read r's segments are
	// [k*seg_len, (k+1)*seg_len), the last one running to the read's end (K = len / seg_len of them).
	// Segment 0 is the read's own pass-1 task; segment k >= 1 is task segbase[r] + k - 1 of the
	// segment launch.
	int32_t seg_len;
	const uint32_t *segbase;    // per read: its first segment task (exclusive scan of K - 1)
	const uint32_t *seg_rid;    // per segment task: its read (segment k found from segbase)
	uint32_t *hand0;            // per read: segment 0's handoff (first chain value >= seg_len)
	uint32_t *hand;             // per segment task: its handoff (the read's end when its chain ends)
	uint32_t *chain;            // per segment task: its chain values, chain_cap each
	uint32_t *chain_n;
	int32_t chain_cap;
	uint32_t *conv;             // per segment task: where it joined the true chain (0xffffffff: never)
	uint32_t *rflag;            // per read: 1 when a segment or fix-up task overflowed (the read falls back)
	const uint32_t *list;       // fix-up and fallback launches: their reads
	const uint32_t *n_list;     // their count (on the device)
};

constexpr int kPackWords = 24;          // 2-bit words per packed read: reads up to 384 bases

struct L2b {
	const uint64_t *pac;
	const uint64_t *ambi;   // [st,en) pairs
	const uint64_t *ctg_off, *ctg_len;
	uint64_t n_ctg, n_ambi, tot_len;
};

// seed -> locate -> extend step (mbg_pipe.hip)
struct MapOpt {
	int32_t a, b, q, e, q2, e2, end_bonus;
	int32_t ksw_w, zdrop;      // band (int)(bw*1.5+1) and zdrop*a for the batch's read length
	int8_t mat[28];
};
struct JobMeta { uint32_t rid; int8_t side, rev; int16_t pad; int32_t tid; int64_t tpos; };

// DP batch arguments (kernels in mbg_dp.hip)
struct DpJob { int64_t qoff, toff; int32_t qlen, tlen; int8_t q, e, q2, e2; int32_t w, zdrop, end_bonus, flag; int64_t cigar_off; };
struct DpRes { int32_t max, zdropped, max_q, max_t, mqe, mqe_t, mte, mte_q, score, reach_end, n_cigar, cells; };
static_assert(sizeof(DpJob) == 56, "DpJob must match mbg_dp_job_t");
static_assert(sizeof(DpRes) == 48, "DpRes must match mbg_dp_result_t");
static_assert(sizeof(SeedRec) == 40, "SeedRec must match mbg_seed_rec_t");

struct DpArgs {
	const DpJob *jobs;
	int32_t n_jobs;
	const unsigned long long *n_jobs_dev;   // when set: job count produced on the device (clamped to n_jobs)
	const uint8_t *qseq, *tseq;
	DpRes *res;
	uint32_t *cigar;
	uint8_t *pbuf;          // backtrack slabs, one per resident block
	int64_t slab;           // bytes per slab
	int32_t cap;            // circular-buffer length (power of two)
	int32_t qt_cap;         // LDS bytes after the 12*cap region for staging a job's query and target
	uint8_t *gstate;        // when gcap > 0: one 12*gcap-byte row-state region per resident block
	int32_t gcap;           // ring length (power of two) for jobs whose band needs more than cap - 17
	int32_t tcap;           // ring length of dp_team_kernel (0: no job needs it)
	unsigned int *next;     // five zeroed job counters (jobs are taken one at a time)
	int32_t *defer;         // n_jobs entries: jobs the fast kernel leaves to the general ones
	int32_t inline_general; // the fast kernel runs the LDS-ring general path itself (set when cap fits its LDS)
	int32_t mat_last, mat_max, mat_min;  // mat[m*m-1] and the extremes (ksw2_extd2_sse.c:77,87-90), as
	                                     // dwords so that they load as scalars
	int8_t m;
	int8_t mat[32];
};

// fills DpArgs' matrix fields from an m x m scoring matrix
inline void dp_set_matrix(DpArgs &a, int8_t m, const int8_t *mat)
{
	a.m = m;
	for (int i = 0; i < 32; ++i) a.mat[i] = i < m * m ? mat[i] : 0;
	a.mat_last = m > 0 ? a.mat[m * m - 1] : 0;
	a.mat_max = a.mat[0], a.mat_min = a.mat[1];     // entries past m*m read as 0
	for (int t = 1; t < m * m; ++t) {
		a.mat_max = a.mat_max > a.mat[t] ? a.mat_max : a.mat[t];
		a.mat_min = a.mat_min < a.mat[t] ? a.mat_min : a.mat[t];
	}
}

// ksw_ll batch (mbg_ll.hip); layouts of mbg_ll_job_t / mbg_ll_result_t
struct LlJob { int64_t qoff, toff; int32_t qlen, tlen, size, pad; };
struct LlRes { int32_t score, te, qe, score2, te2; };
static_assert(sizeof(LlJob) == 32, "LlJob must match mbg_ll_job_t");
static_assert(sizeof(LlRes) == 20, "LlRes must match mbg_ll_result_t");

struct LlArgs {
	const LlJob *jobs;
	const int32_t *order;   // the launch's job indices (null: jobs 0..n-1)
	int32_t n;              // entries in order
	const uint8_t *qseq, *tseq;
	LlRes *res;
	uint64_t *bbuf;         // per-row maxima list of job j at bbuf + boff[j] (tlen entries)
	const int64_t *boff;
	int32_t gapo, gape, minsc, endsc;
	int32_t shift, qmax;    // u8 bias of the matrix (ksw2_ll_sse.c:67-74) and its largest entry
	int32_t m;
	const int8_t *sc;       // scores sc[t*8 + c] for target base t, query base c; c = 7: padding (-1)
	int64_t rows_cap;       // entries in bbuf; a job whose span would pass it is refused
	bool check;             // refuse jobs outside the launch's range on the device (score -1)
};
constexpr int kLlMaxQlen = 1024;        // query positions a wave holds in registers (16 a lane)
// slots per lane (1, 2, 3, 4, 6, 8, 12 or 16) for a job of qlen bases and `size` (1: u8, 2: i16)
int ll_slots(int qlen, int size);
hipError_t launch_ll(const LlArgs &a, int slots, hipStream_t st);
hipError_t scan_ll_rows(void *temp, size_t *temp_bytes, const LlJob *jobs, int64_t *out, int64_t n, hipStream_t st);

// hit finishing (mbg_extra.hip); layout of mbg_extra_job_t
struct ExtraJob {
	int64_t qoff, toff;
	int32_t qs, qe;
	int64_t ts, te;
	int32_t rev, n_cigar;
	int64_t cigar_off;
	int32_t blen, mlen, n_ambi, dp_max;
};
static_assert(sizeof(ExtraJob) == 72, "ExtraJob must match mbg_extra_job_t");
struct ExtraArgs {
	ExtraJob *jobs;
	int64_t n;
	const uint8_t *qseq, *tseq;
	uint32_t *cigar;
	int8_t mat[28];
	int32_t q, e, log_gap;
};
hipError_t launch_extra(const ExtraArgs &a, hipStream_t st);

// mate rescue (mbg_rescue.hip); layouts of mbg_matesw_job_t / mbg_matesw_result_t
struct MsJob { int64_t qoff; int32_t qlen, tid; int64_t ts, te; int32_t min_sc, pad; int64_t cigar_off; };
struct MsRes {
	int32_t max_ug, max_i, n_good, n_kmer, ts2, te2, aligned, rescued;
	LlRes ll;
	DpRes ez;
	int32_t qs, qe, ts, te, dp_max2, ran_dp;
	int32_t n_cigar, blen, mlen, n_ambi, dp_max;
};
static_assert(sizeof(MsJob) == 48, "MsJob must match mbg_matesw_job_t");
static_assert(sizeof(MsRes) == 144, "MsRes must match mbg_matesw_result_t");

struct MsArgs {
	const MsJob *jobs;
	int32_t n;
	const uint8_t *qseq;
	const uint8_t *win;     // candidate windows: c's at win + woff[c], wlen[c] bases (clipped to its contig)
	const int64_t *woff, *wlen;
	uint8_t *rq, *rt;       // reversed prefixes for the extension DP, laid out as qseq and win
	LlJob *ll_jobs;
	const LlRes *ll_res;
	DpJob *dp_jobs;
	const DpRes *dp_res;
	MsRes *res;
	const uint32_t *cigar;  // the extension DP's CIGAR slots (at each job's cigar_off)
	uint32_t *ccnt;         // per candidate: CIGAR operations to bring back
	ExtraJob *extra;        // per candidate: the rescued hit's mb_fix_cigar + mb_update_extra job
	int32_t tab_size;       // k-mer table slots of the vote (a power of two >= 2 * the longest read)
	int32_t mt, a, b_mm, q, e, bw, zdrop, end_bonus, min_dp_max;
};
// mb_log2 (mbpriv.h:105-113), with the float operations kept separate as the host computes them
__device__ __forceinline__ float mb_log2_dev(float x)
{
#pragma clang fp contract(off)
	uint32_t i = __float_as_uint(x);
	float l = (float)(((i >> 23) & 255u) - 128u);
	i &= ~(255u << 23);
	i += 127u << 23;
	const float f = __uint_as_float(i);
	const float a = -0.34484843f * f;
	const float b = a + 2.02466578f;
	const float c = b * f;
	l += c - 0.67487759f;
	return l;
}

// mb_anchor and mb_lchain_dp (kernels in mbg_anchor.hip); AncRec is layout-identical to mbg_anchor_t / mb_anchor_t
struct AncRec { int32_t sid, len, qpos; uint32_t flag; int64_t tpos; };
static_assert(sizeof(AncRec) == 24, "AncRec must match mbg_anchor_t");
struct AncArgs {
	int32_t n_reads, max_occ;
	Sai *u;                  // read r's intervals at u[uoff[r] .. uoff[r+1]), sorted and deduplicated in place
	const int64_t *uoff;
	const int32_t *qlen;
	const int64_t *aoff;     // read r's row and anchor slots: [aoff[r], aoff[r+1])
	uint64_t *rows;          // rows to locate, then their SA positions
	AncRec *anc;
	int64_t *nu, *nanc;      // per read: intervals after dedup, anchors; -1 when the slots do not suffice
	L2b lb;
};
hipError_t launch_anc_rows(const AncArgs &a, hipStream_t st);
hipError_t launch_anc_build(const AncArgs &a, hipStream_t st);
struct U128 { uint64_t x, y; };           // mb128_t (mbpriv.h:42)
// one set of mb_lchain_dp's parameters; layout of mbg_chain_opt_t
struct ChainOpt { int32_t max_dist_x, max_dist_y, bw, max_skip, max_iter, min_sc; float pen_gap; int32_t pad; };
static_assert(sizeof(ChainOpt) == 32, "ChainOpt must match mbg_chain_opt_t");
struct ChainArgs {
	int32_t n_reads;
	const AncRec *anc;       // read r's anchors at anc[aoff[r] .. aoff[r] + nanc[r]), mb_anchor's order
	const int64_t *aoff, *nanc;
	int32_t max_dist_x, max_dist_y, bw, max_skip, max_iter, min_sc;
	float pen_gap;
	const ChainOpt *opts;    // when set: read r uses opts[opt_idx[r]] instead of the fields above
	const int32_t *opt_idx;
	const uint64_t *ctg_off, *ctg_len;
	// per-read workspace at offset aoff[r] of each array
	int32_t *f, *t, *v;
	int64_t *p;
	U128 *z;
	uint64_t *ut;
	AncRec *bb;
	// output at offset aoff[r]: n_u[r] chains (score<<32 | count), n_b[r] chained anchors
	uint64_t *u;
	int32_t *n_u;
	AncRec *b;
	int64_t *n_b;
};
hipError_t launch_lchain(const ChainArgs &a, hipStream_t st);

constexpr int kMsMaxWindow = 16384;    // rescue window bases: 16-bit vote counters, 32 KB of LDS
hipError_t launch_ungap(const MsArgs &a, int max_wlen, hipStream_t st);
hipError_t launch_rescue_dp(const MsArgs &a, hipStream_t st);
hipError_t launch_rescue_finish(const MsArgs &a, hipStream_t st);
hipError_t launch_rescue_cigar(const MsArgs &a, const uint32_t *coff, uint32_t *packed, hipStream_t st);
hipError_t launch_rescue_final(const MsArgs &a, hipStream_t st);


hipError_t launch_kmer_cache(const Idx &ix, Sai *pre, int len, hipStream_t st);
hipError_t launch_sa_fine(const Idx &ix, uint64_t *fine, uint64_t n_fine, hipStream_t st);
hipError_t launch_rank2a(const Idx &ix, int64_t n, const uint64_t *k, const uint64_t *l, uint64_t *ck, uint64_t *cl, hipStream_t st);
hipError_t launch_extend(const Idx &ix, int64_t n, const Sai *ik, const uint8_t *is_back, Sai *ok, hipStream_t st);
hipError_t launch_smem(const Idx &ix, const SeedArgs &a, uint32_t pass, unsigned grid, hipStream_t st);
// segmented pass 1: segment counts per read (long[r] = K - 1), the segment tasks and the long-read list
hipError_t launch_seg_count(const int32_t *len, int32_t n_reads, int32_t seg_len, uint32_t *nseg, hipStream_t st);
hipError_t launch_seg_build(const int32_t *len, int32_t n_reads, int32_t seg_len, const uint32_t *segbase, int64_t seg_cap,
                            uint32_t *seg_rid, uint32_t *longl, uint32_t *n_long, uint32_t *rflag, hipStream_t st);
// after the segment and fix-up launches: the flagged reads (fallback list) and the valid records of
// the others appended to the spill list
hipError_t launch_seg_flagged(const uint32_t *longl, const uint32_t *n_long, const uint32_t *rflag, uint32_t *fbl, uint32_t *n_fb,
                              int64_t cap, hipStream_t st);
hipError_t launch_seg_merge(const int32_t *len, int32_t seg_len, const uint32_t *longl, const uint32_t *n_long, const uint32_t *rflag,
                            const uint32_t *segbase, const SeedRec *segslots, const uint32_t *segcnt, int32_t kseg,
                            const uint32_t *conv, const SeedRec *fixslots, const uint32_t *fixcnt, int32_t kfix, SeedRec *spill,
                            unsigned long long *n_spill, int64_t spill_cap, int64_t cap, hipStream_t st);
hipError_t launch_subtask(const SeedRec *slots, const uint32_t *slot_cnt, int32_t kslot, int32_t n_reads, const SeedRec *spill,
                          const unsigned long long *n_spill, int64_t spill_cap, int32_t min_len, int32_t max_sub_occ,
                          SeedTask *tasks, uint32_t *n_tasks, int64_t task_cap, hipStream_t st);
hipError_t launch_compact(const SeedRec *slots1, const SeedRec *slots2, int32_t k1, int32_t k2, int64_t n1, int64_t n_cnt,
                          const uint32_t *cnt, const uint32_t *off, const SeedRec *spill, const unsigned long long *n_spill,
                          int64_t spill_cap, SeedRec *out, int64_t cap, int64_t *n_out, hipStream_t st);
hipError_t scan_u32(void *temp, size_t *temp_bytes, const uint32_t *in, uint32_t *out, int64_t n, hipStream_t st);
// exclusive prefix sums into 64-bit offsets
hipError_t scan_u32_i64(void *temp, size_t *temp_bytes, const uint32_t *in, int64_t *out, int64_t n, hipStream_t st);
hipError_t scan_u64_i64(void *temp, size_t *temp_bytes, const unsigned long long *in, int64_t *out, int64_t n, hipStream_t st);
// CIGARs of a DP batch packed back to back (mbg_front.hip); poff[i]: job i's first operation
hipError_t launch_cigar_pack(const DpJob *jobs, const DpRes *res, int32_t n, const int64_t *poff, const uint32_t *cigar, uint32_t *packed,
                             hipStream_t st);
// grouping of seeding records by read (mbg_front.hip)
hipError_t launch_grp_count(const SeedRec *recs, const int64_t *n_rec, int64_t cap, int32_t max_occ, uint32_t *cnt,
                            unsigned long long *slots, unsigned grid, hipStream_t st);
hipError_t launch_grp_scatter(const SeedRec *recs, const int64_t *n_rec, int64_t cap, const int64_t *uoff, uint32_t *cursor, Sai *intv,
                              unsigned grid, hipStream_t st);
hipError_t launch_sa(const Idx &ix, int64_t n, const int64_t *n_dev, uint64_t *x, uint32_t *next, unsigned long long *stats,
                     unsigned grid, hipStream_t st);
hipError_t launch_getseq(const L2b &lb, int64_t n, const int64_t *tid, const int64_t *st_, const int64_t *en_, const int64_t *out_off,
                         uint8_t *out, int64_t *len, hipStream_t st);
hipError_t launch_pack_reads(const uint8_t *seq, const int64_t *off, const int32_t *len, int32_t n_reads, uint32_t *packed,
                             uint32_t *qflag, hipStream_t st);
hipError_t launch_dp(bool dual, const DpArgs &a, unsigned grid, hipStream_t st);
// mb_align_pair's decisions before the DP (mbg_align.hip): one call of align.c:329-390 per job
struct AlignJob { int64_t qoff, toff; int32_t qlen, tlen; int32_t w, end_bonus, zdrop, flag; int64_t cigar_off; };
static_assert(sizeof(AlignJob) == 48, "AlignJob must match mbg_align_job_t");
struct AlignArgs {
	const AlignJob *jobs;
	int64_t n;
	const uint8_t *qseq, *tseq;
	const int8_t *mat;                   // 5x5
	int32_t sc_a, sc_b, q, e, q2, e2;    // opt->a, b, q, e, q2, e2
	int32_t generic;                     // KSW_EZ_GENERIC_SC for every call (transition mismatch or methylation)
	int64_t max_sw_mat;
	DpJob *dp;                           // out: the DP job, or an empty one (qlen 0) for a settled call
	uint8_t *kind;                       // out: 0 DP, 1 fast path, 2 max_sw_mat skip
	DpRes *res;                          // out: the settled calls' results
	uint32_t *cigar;                     // out: the fast paths' CIGARs at cigar_off
};
hipError_t launch_align_plan(const AlignArgs &a, hipStream_t st);
hipError_t launch_align_merge(const uint8_t *kind, const DpRes *decided, DpRes *res, int64_t n, hipStream_t st);
// ring length of dp_team_kernel for a launch whose widest band is nb_max cells (0: not needed)
int dp_team_cap(int nb_max);
hipError_t launch_gather_chain(const Idx &ix, int64_t n_chains, int32_t steps, int32_t bs, unsigned long long *sum, hipStream_t st);
hipError_t launch_stream_copy(void *dst, const void *src, int64_t n16, unsigned grid, hipStream_t st);
hipError_t launch_row_counts(const SeedRec *recs, const int64_t *n_rec, int64_t rec_cap, int32_t max_occ, uint32_t *cnt, hipStream_t st);
hipError_t launch_anchor_rows(const SeedRec *recs, const int64_t *n_rec, int64_t rec_cap, int32_t max_occ, const uint32_t *cnt,
                              const uint32_t *off, uint64_t *rows, uint32_t *row_rec, uint32_t *first_slot,
                              unsigned long long *n_rows, int64_t row_cap, hipStream_t st);
hipError_t launch_best_seed(const SeedRec *recs, const int64_t *n_rec, int64_t rec_cap, unsigned long long *best, unsigned grid, hipStream_t st);
hipError_t launch_ext_jobs(const L2b &lb, const MapOpt &o, int32_t n_reads, const int64_t *off, const int32_t *len, const uint8_t *seq,
                           const SeedRec *recs, const uint32_t *first_slot, const uint64_t *pos, const unsigned long long *best,
                           DpJob *jobs, JobMeta *meta, uint8_t *qbuf, uint8_t *tbuf, int64_t job_cap, int64_t qcap, int64_t tcap,
                           unsigned long long *ctr, hipStream_t st);

// geometry of one DP job (ksw2_extd2_sse.c:82-86): the band width after `w<0` handling and the
// number of 16-cell columns per anti-diagonal row
__host__ __device__ static inline void dp_geom(int qlen, int tlen, int w, int *w_out, int *n_col_)
{
	if (w < 0) w = tlen > qlen ? tlen : qlen;
	int n = qlen < tlen ? qlen : tlen;
	n = ((n < w + 1 ? n : w + 1) + 15) / 16 + 1;
	*w_out = w, *n_col_ = n;
}

} // namespace mbg
