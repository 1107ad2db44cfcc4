/*
 * mbg_fmt.h -- SAM and PAF records for the batched mapper's output step.
 *
 * mbg_fmt_record() writes the same bytes as the reference's mb_format (format.c:355-361), which
 * dispatches to mb_fmt_sam (format.c:196-353) or mb_fmt_paf (format.c:19-46), for the same arguments.
 * The reference builds every field with kom_sprintf_lite (kommon.c:66-138), which parses its format
 * string and checks the buffer's room once per field; here each record reserves room once per group
 * of fields and writes integers with a two-digit table, which makes the output step several times
 * cheaper on the host threads that the mapping step also needs.
 */
#ifndef MBG_FMT_H
#define MBG_FMT_H

#include <stdint.h>
#include "mbpriv.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The read group ID that mb_fmt_sam appends as RG:Z (format.c:7,77-112), as set by the header
 * formatting of the run (mb_fmt_sam_hdr): learnt once per run by formatting a probe record with the
 * reference's own mb_fmt_sam.  Returns 0, or -1 when it cannot tell. */
int mbg_fmt_learn_rg(const mb_opt_t *opt);

/* Appends one record to s, as mb_format(km, s, l2b, t, n_seg, n_hit, hit, hit_idx, opt, seg_idx,
 * mate_qlen) would.  Returns 0, or -1 for a record it leaves to mb_format (a sequence byte >= 128 on
 * a reverse-complemented SEQ, where the reference indexes its complement table with a negative
 * char); s is then as it was. */
int mbg_fmt_record(kstring_t *s, const l2b_t *l2b, const mb_bseq1_t *t, int32_t n_seg, const int32_t *n_hit,
                   mb_hit_t *const *hit, int32_t hit_idx, const mb_opt_t *opt, int seg_idx, int32_t mate_qlen);

#ifdef __cplusplus
}
#endif
#endif
