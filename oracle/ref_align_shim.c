/*
 * ref_align_shim.c -- exposes the reference's mb_align_pair (align.c:329-390), which align.c keeps
 * static, to the tests.  TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles it with align.c
 * included from where it lies in the reference tree (nothing of it is copied here) into
 * oracle/_ref/libmbref_pe.so, beside ref_pe_shim.c.
 */
#include "align.c"

/* mb_align_pair on one query/target pair.  p = {a, b, b_ts, q, e, q2, e2, meth}: the mb_opt_t fields
 * the function reads (meth sets MB_F_METH); max_sw_mat as opt->max_sw_mat.  ez_out = {max, zdropped,
 * max_q, max_t, mqe, mqe_t, mte, mte_q, score, reach_end, n_cigar}; the CIGAR (at most cap
 * operations) goes to cigar. */
void mbx_align_pair(const int32_t p[8], int64_t max_sw_mat, int32_t qlen, const uint8_t *qseq, int32_t tlen, const uint8_t *tseq,
                    const int8_t *mat, int32_t w, int32_t end_bonus, int32_t zdrop, int32_t flag, int32_t ez_out[11],
                    uint32_t *cigar, int32_t cap)
{
	mb_opt_t opt;
	ksw_extz_t ez;
	mb_opt_init(&opt);
	opt.a = p[0], opt.b = p[1], opt.b_ts = p[2], opt.q = p[3], opt.e = p[4], opt.q2 = p[5], opt.e2 = p[6];
	if (p[7]) opt.flag |= MB_F_METH;
	else opt.flag &= ~MB_F_METH;
	opt.max_sw_mat = max_sw_mat;
	memset(&ez, 0, sizeof(ez));
	mb_align_pair(0, &opt, qlen, qseq, tlen, tseq, mat, w, end_bonus, zdrop, flag, &ez);
	ez_out[0] = ez.max, ez_out[1] = ez.zdropped, ez_out[2] = ez.max_q, ez_out[3] = ez.max_t, ez_out[4] = ez.mqe;
	ez_out[5] = ez.mqe_t, ez_out[6] = ez.mte, ez_out[7] = ez.mte_q, ez_out[8] = ez.score, ez_out[9] = ez.reach_end;
	ez_out[10] = ez.n_cigar;
	if (ez.n_cigar > 0) memcpy(cigar, ez.cigar, 4 * (size_t)(ez.n_cigar < cap ? ez.n_cigar : cap));
	free(ez.cigar);
}
