/*
 * mbg_l2b.h -- target extraction for the mapper's host code.
 */
#ifndef MBG_L2B_H
#define MBG_L2B_H

#include <stdint.h>
#include "l2bit.h"

#ifdef __cplusplus
extern "C" {
#endif

/* l2b_getseq (l2bit.c:52-77): bases [st,en) of contig tid as 0..3, ambiguous bases (N runs) as 4,
 * with the same clipping and return value (en - st after clipping, -1 for a bad tid).  Four bases
 * per table lookup instead of one shift per base. */
int64_t mbg_l2b_getseq(const l2b_t *l2b, int64_t tid, int64_t st, int64_t en, uint8_t *seq);

#ifdef __cplusplus
}
#endif
#endif
