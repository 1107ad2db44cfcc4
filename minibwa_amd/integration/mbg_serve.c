/*
 * mbg_serve.c -- continuous batching with memory-gated admission over the batched mapper (interface
 * and rationale in mbg_serve.h).
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "mbg_exec.h"
#include "mbg_serve.h"

/* the output step's records of one fragment, then its hits freed (mbg_map.c) */
void mbg_map_fmt_fragment(kstring_t *out, const l2b_t *l2b, const mb_opt_t *opt, mb_bseq1_t *t, int32_t n_seg, int32_t *n_hit,
                          mb_hit_t **hit);

typedef struct req_s {
	int64_t id;
	int paired, collected, done;
	int32_t n;
	char **name, **seq;
	int32_t *qlen;
	int64_t bases;
	double t_start, t_done;
	char *sam;
	int64_t sam_len;
	struct req_s *next;      /* the queue */
} req_t;

/* a mapped batch waiting for its records */
typedef struct fbatch_s {
	req_t *reqs;
	int64_t n_req, n_seq, bases;
	int paired;
	int32_t *n_hit;
	mb_hit_t **hit;
	double t0;
	struct fbatch_s *next;
} fbatch_t;

struct mbg_server_s {
	mb_opt_t opt_pe, opt_se;
	const mb_idx_t *idx;
	mbg_serve_opt_t so;
	pthread_t th, th_fmt;
	pthread_mutex_t fmu;     /* the formatting queue */
	pthread_cond_t fcv;
	fbatch_t *fhead, *ftail;
	int fstop;
	pthread_mutex_t mu;
	pthread_cond_t cv_work, cv_done;
	req_t *head, *tail;
	req_t **all;             /* by id */
	int64_t n_all, m_all;
	int stop;
	double bpb;              /* device bytes per base beyond the largest batch so far */
	int64_t work;            /* the executor's work buffers after the last batch */
	int64_t n_grow;          /* batches that grew the largest batch */
	mbg_serve_stats_t st;
};

double mbg_server_now(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void die(const char *msg)
{
	fprintf(stderr, "[mbg_serve] %s\n", msg);
	abort();
}

static void *xmalloc(size_t n)
{
	void *p = malloc(n ? n : 1);
	if (p == 0) die("out of memory");
	return p;
}

static void free_req(req_t *r)
{
	int32_t i;
	for (i = 0; i < r->n; ++i) free(r->name[i]), free(r->seq[i]);
	free(r->name), free(r->seq), free(r->qlen), free(r->sam);
	free(r);
}

/* the requests of the next batch, unlinked from the queue (called with the lock held) */
static req_t *take_batch(mbg_server_t *s, int64_t *n_req, int *paired)
{
	req_t *b = 0, **bt = &b, **pp = &s->head, *q;
	const int kind = s->head->paired;
	/* the work buffers grow to the largest batch and stay: a batch up to that size needs no new
	 * device memory, a larger one about bpb bytes per extra base, within what the budget has left */
	const double left = (double)(s->st.dev_budget - s->work);
	const double cap = (double)s->st.max_batch_bases + (left > 0 ? left / s->bpb : 0.0);
	int64_t bases = 0, n = 0;
	int gated = 0;
	while (*pp) {
		req_t *r = *pp;
		if (r->paired != kind) { pp = &r->next; continue; }
		if (n > 0 && ((double)(bases + r->bases) > cap || (s->so.max_reqs > 0 && n >= s->so.max_reqs) || s->so.solo)) {
			if ((double)(bases + r->bases) > cap) gated = 1;
			break;
		}
		*pp = r->next;                    /* unlink, keep the others in order */
		r->next = 0, *bt = r, bt = &r->next;
		bases += r->bases, ++n;
	}
	s->tail = 0;                          /* the last request still queued */
	for (q = s->head; q; q = q->next) s->tail = q;
	if (gated) ++s->st.gated;
	*n_req = n, *paired = kind;
	return b;
}

/* the mapping of a batch: one mb_map_batch call; its records are written by the formatting thread
 * while the next batch maps */
static void run_batch(mbg_server_t *s, req_t *b, int64_t n_req, int paired)
{
	const mb_opt_t *opt = paired ? &s->opt_pe : &s->opt_se;
	int64_t k, n_seq = 0, bases = 0, i;
	int32_t *qlen;
	const char **seq, **name;
	fbatch_t *fb = (fbatch_t*)calloc(1, sizeof(fbatch_t));
	req_t *r;
	double t0 = mbg_server_now();
	if (fb == 0) die("out of memory");
	for (r = b; r; r = r->next) n_seq += r->n, bases += r->bases;
	if (n_seq > INT32_MAX) die("a batch of more than 2^31 reads");
	qlen = (int32_t*)xmalloc(n_seq * sizeof(int32_t));
	seq = (const char**)xmalloc(n_seq * sizeof(char*));
	name = (const char**)xmalloc(n_seq * sizeof(char*));
	fb->n_hit = (int32_t*)xmalloc(n_seq * sizeof(int32_t));
	for (r = b, k = 0; r; r = r->next)
		for (i = 0; i < r->n; ++i, ++k) qlen[k] = r->qlen[i], seq[k] = r->seq[i], name[k] = r->name[i];
	for (r = b; r; r = r->next) r->t_start = t0;
	fb->hit = mb_map_batch(opt, s->idx, (int32_t)n_seq, qlen, seq, fb->n_hit, 0, name);
	if (fb->hit == 0) die("mb_map_batch failed");
	free(qlen), free(seq), free(name);
	fb->reqs = b, fb->n_req = n_req, fb->n_seq = n_seq, fb->bases = bases, fb->paired = paired, fb->t0 = t0;
	pthread_mutex_lock(&s->fmu);
	if (s->ftail) s->ftail->next = fb;
	else s->fhead = fb;
	s->ftail = fb;
	pthread_cond_signal(&s->fcv);
	pthread_mutex_unlock(&s->fmu);
}

/* each request's records, fragment by fragment, requests in parallel; then the batch is done */
static void format_batch(mbg_server_t *s, fbatch_t *fb)
{
	const mb_opt_t *opt = fb->paired ? &s->opt_pe : &s->opt_se;
	int64_t q, k;
	req_t *r, **rv = (req_t**)xmalloc(fb->n_req * sizeof(req_t*));
	int64_t *first = (int64_t*)xmalloc(fb->n_req * sizeof(int64_t));
	for (r = fb->reqs, q = 0, k = 0; r; r = r->next, ++q) rv[q] = r, first[q] = k, k += r->n;
#pragma omp parallel for schedule(dynamic, 1)
	for (q = 0; q < fb->n_req; ++q) {
		req_t *rq = rv[q];
		kstring_t out = {0, 0, 0};
		const int32_t step = rq->paired ? 2 : 1;
		int64_t i, kk = first[q];
		for (i = 0; i < rq->n; i += step, kk += step) {
			mb_bseq1_t t[2];
			int32_t j;
			memset(t, 0, sizeof(t));
			for (j = 0; j < step; ++j) t[j].l_seq = (uint64_t)rq->qlen[i + j], t[j].name = rq->name[i + j], t[j].seq = rq->seq[i + j];
			mbg_map_fmt_fragment(&out, s->idx->l2b, opt, t, step, &fb->n_hit[kk], &fb->hit[kk]);
		}
		rq->sam = out.s, rq->sam_len = (int64_t)out.l;
	}
	free(rv), free(first);
	free(fb->hit), free(fb->n_hit);
	{
		int64_t work = 0, fr = 0, tot = 0;
		mbx_mem(&work, &fr, &tot);
		pthread_mutex_lock(&s->mu);
		if (fb->bases > s->st.max_batch_bases) {
			/* the growth per extra base; the first batch's growth also holds the buffers whose size
			 * does not depend on the batch, so it sets no rate */
			if (s->n_grow > 0 && work > s->work) {
				const double obs = (double)(work - s->work) / (double)(fb->bases - s->st.max_batch_bases);
				if (obs > s->bpb) s->bpb = obs;
			}
			++s->n_grow;
			s->st.max_batch_bases = fb->bases;
		}
		s->work = work;
		s->st.work_bytes = work;
		++s->st.batches, s->st.reqs += fb->n_req, s->st.reads += fb->n_seq, s->st.bases += fb->bases;
		s->st.busy_s += mbg_server_now() - fb->t0;
		s->st.bytes_per_base = s->bpb;
		for (r = fb->reqs; r; r = r->next) r->t_done = mbg_server_now(), r->done = 1;
		pthread_cond_broadcast(&s->cv_done);
		pthread_mutex_unlock(&s->mu);
	}
	free(fb);
}

static void *format_main(void *arg)
{
	mbg_server_t *s = (mbg_server_t*)arg;
	for (;;) {
		fbatch_t *fb;
		pthread_mutex_lock(&s->fmu);
		while (s->fhead == 0 && !s->fstop) pthread_cond_wait(&s->fcv, &s->fmu);
		if (s->fhead == 0) {
			pthread_mutex_unlock(&s->fmu);
			break;
		}
		fb = s->fhead;
		s->fhead = fb->next;
		if (s->fhead == 0) s->ftail = 0;
		pthread_mutex_unlock(&s->fmu);
		format_batch(s, fb);
	}
	return 0;
}

static void *server_main(void *arg)
{
	mbg_server_t *s = (mbg_server_t*)arg;
	for (;;) {
		req_t *b;
		int64_t n_req;
		int paired;
		pthread_mutex_lock(&s->mu);
		while (s->head == 0 && !s->stop) pthread_cond_wait(&s->cv_work, &s->mu);
		if (s->head == 0 && s->stop) {
			pthread_mutex_unlock(&s->mu);
			break;
		}
		b = take_batch(s, &n_req, &paired);
		pthread_mutex_unlock(&s->mu);
		run_batch(s, b, n_req, paired);
	}
	return 0;
}

mbg_server_t *mbg_server_start(const mb_opt_t *opt_pe, const mb_opt_t *opt_se, const mb_idx_t *idx, const mbg_serve_opt_t *so)
{
	mbg_server_t *s = (mbg_server_t*)calloc(1, sizeof(mbg_server_t));
	int64_t work = 0, fr = 0, tot = 0;
	if (s == 0) return 0;
	s->opt_pe = *opt_pe, s->opt_se = *opt_se, s->idx = idx;
	if (so) s->so = *so;
	s->bpb = s->so.bytes_per_base0 > 0 ? s->so.bytes_per_base0 : 4096.0;
	mbx_mem(&work, &fr, &tot);
	s->work = work;
	s->st.dev_budget = s->so.dev_budget > 0 ? s->so.dev_budget : fr > 0 ? work + fr / 2 : (int64_t)1 << 34;
	s->st.bytes_per_base = s->bpb;
	pthread_mutex_init(&s->mu, 0);
	pthread_cond_init(&s->cv_work, 0);
	pthread_cond_init(&s->cv_done, 0);
	pthread_mutex_init(&s->fmu, 0);
	pthread_cond_init(&s->fcv, 0);
	if (pthread_create(&s->th_fmt, 0, format_main, s) != 0) {
		free(s);
		return 0;
	}
	if (pthread_create(&s->th, 0, server_main, s) != 0) {
		pthread_mutex_lock(&s->fmu);
		s->fstop = 1;
		pthread_cond_signal(&s->fcv);
		pthread_mutex_unlock(&s->fmu);
		pthread_join(s->th_fmt, 0);
		free(s);
		return 0;
	}
	return s;
}

int64_t mbg_server_submit(mbg_server_t *s, int paired, int32_t n, const char *const *name, const char *const *seq)
{
	req_t *r;
	int32_t i;
	int64_t id;
	if (n <= 0 || (paired && (n & 1))) return -1;
	r = (req_t*)calloc(1, sizeof(req_t));
	if (r == 0) die("out of memory");
	r->paired = !!paired, r->n = n;
	r->name = (char**)xmalloc(n * sizeof(char*));
	r->seq = (char**)xmalloc(n * sizeof(char*));
	r->qlen = (int32_t*)xmalloc(n * sizeof(int32_t));
	for (i = 0; i < n; ++i) {
		size_t l = strlen(seq[i]);
		r->name[i] = strdup(name[i]), r->seq[i] = strdup(seq[i]);
		if (r->name[i] == 0 || r->seq[i] == 0) die("out of memory");
		r->qlen[i] = (int32_t)l, r->bases += (int64_t)l;
	}
	pthread_mutex_lock(&s->mu);
	if (s->stop) {
		pthread_mutex_unlock(&s->mu);
		free_req(r);
		return -1;
	}
	id = r->id = s->n_all;
	if (s->n_all == s->m_all) {
		s->m_all = s->m_all ? s->m_all * 2 : 1024;
		s->all = (req_t**)realloc(s->all, s->m_all * sizeof(req_t*));
		if (s->all == 0) die("out of memory");
	}
	s->all[s->n_all++] = r;
	if (s->tail) s->tail->next = r;
	else s->head = r;
	s->tail = r;
	pthread_cond_signal(&s->cv_work);
	pthread_mutex_unlock(&s->mu);
	return id;
}

int mbg_server_wait(mbg_server_t *s, int64_t id, char **sam, int64_t *len, double *t_start, double *t_done)
{
	req_t *r;
	pthread_mutex_lock(&s->mu);
	if (id < 0 || id >= s->n_all || s->all[id] == 0 || s->all[id]->collected) {
		pthread_mutex_unlock(&s->mu);
		return -1;
	}
	r = s->all[id];
	while (!r->done) pthread_cond_wait(&s->cv_done, &s->mu);
	r->collected = 1;
	s->all[id] = 0;
	pthread_mutex_unlock(&s->mu);
	*sam = r->sam, *len = r->sam_len;
	if (t_start) *t_start = r->t_start;
	if (t_done) *t_done = r->t_done;
	r->sam = 0;
	free_req(r);
	return 0;
}

void mbg_server_stop(mbg_server_t *s)
{
	int64_t i;
	if (s == 0) return;
	pthread_mutex_lock(&s->mu);
	s->stop = 1;
	pthread_cond_signal(&s->cv_work);
	pthread_mutex_unlock(&s->mu);
	pthread_join(s->th, 0);
	pthread_mutex_lock(&s->fmu);          /* the mapping thread has queued its last batch */
	s->fstop = 1;
	pthread_cond_signal(&s->fcv);
	pthread_mutex_unlock(&s->fmu);
	pthread_join(s->th_fmt, 0);
	for (i = 0; i < s->n_all; ++i)
		if (s->all[i]) free_req(s->all[i]);
	free(s->all);
	pthread_mutex_destroy(&s->mu);
	pthread_cond_destroy(&s->cv_work);
	pthread_cond_destroy(&s->cv_done);
	pthread_mutex_destroy(&s->fmu);
	pthread_cond_destroy(&s->fcv);
	free(s);
}

void mbg_server_stats(mbg_server_t *s, mbg_serve_stats_t *st)
{
	pthread_mutex_lock(&s->mu);
	*st = s->st;
	pthread_mutex_unlock(&s->mu);
}
