/*
 * mbg_pe.h -- the insert-size model of a mini-batch, mb_pestat (pe.c:68-124, declared mbpriv.h:101)
 * with its pass over the pairs on n_threads threads; the same pes[4] and the same log lines.
 */
#ifndef MBG_PE_H
#define MBG_PE_H

#include <stdint.h>
#include "mbpriv.h"

#ifdef __cplusplus
extern "C" {
#endif

void mbg_pestat(const mb_opt_t *opt, int32_t n_frag, const int32_t *seg_off, const int32_t *seg_cnt, const int32_t *n_hit,
                mb_hit_t *const *hit, mb_pestat_t pes[4], int n_threads);

#ifdef __cplusplus
}
#endif
#endif
