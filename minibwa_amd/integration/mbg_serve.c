/*
 * mbg_serve.c -- continuous batching with memory-gated admission over the batched mapper (interface
 * and rationale in mbg_serve.h).
 *
 * One open run of the mapper's worker pool lasts as long as the server.  The admission thread turns
 * the queued requests into groups (mbg_map.h): it runs a group's front pass on a free front slot and
 * adds one task per read pair (or single read) to the running pool; the tasks map, pair and format
 * their fragment, so the reads of every admitted group map at the same time and share device batches.
 * A request is done when its last fragment is written, a group when its last task ends, which frees
 * its front slot and its share of the memory budget for the next group.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "mbg_exec.h"
#include "mbg_map.h"
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
	kstring_t *frag;         /* each fragment's records, written by its task */
	int32_t n_frag;
	atomic_int left;         /* fragments not written yet */
	struct req_s *next;      /* the queue */
} req_t;

typedef struct group_s {
	mbg_group_t *g;
	int64_t n_req, n_seq, bases;
	int paired, slot;
	int32_t *qlen;
	const char **seq, **name;
	atomic_llong left;       /* tasks not finished */
	double t0;
} group_t;

/* task i of the open run: fragment `frag` of request `req`, reads r0 (and r0 + 1) of its group */
typedef struct { group_t *grp; req_t *req; int32_t r0, frag; } task_t;

#define TASK_SHIFT 16
#define TASK_CHUNK (1 << TASK_SHIFT)
#define TASK_MAX_CHUNKS (1 << 20)

struct mbg_server_s {
	mb_opt_t opt_pe, opt_se;
	const mb_idx_t *idx;
	mbg_serve_opt_t so;
	mbs_pool_t *pool;
	pthread_t th;
	pthread_mutex_t mu;
	pthread_cond_t cv_work, cv_done;
	req_t *head, *tail;
	req_t **all;             /* by id */
	int64_t n_all, m_all;
	int stop;
	int slot_busy[MBX_FRONT_SLOTS];
	int64_t n_live;          /* groups admitted and not finished */
	int64_t inflight;        /* their bases */
	int64_t peak_upd;        /* the in-flight peak when the rate was last updated */
	double bpb;              /* device bytes per base beyond the in-flight peak so far */
	int64_t work;            /* the executor's work buffers at the last update */
	int64_t n_grow;          /* updates that saw the peak grow */
	/* the task table: chunks of TASK_CHUNK tasks, each freed when its last task ends */
	task_t **tab;
	atomic_int *tab_left;
	int64_t n_task;          /* tasks written (by the admission thread only) */
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

static void *xcalloc(size_t n, size_t sz)
{
	void *p = calloc(n ? n : 1, sz);
	if (p == 0) die("out of memory");
	return p;
}

static void free_req(req_t *r)
{
	int32_t i;
	for (i = 0; i < r->n; ++i) free(r->name[i]), free(r->seq[i]);
	if (r->frag)
		for (i = 0; i < r->n_frag; ++i) free(r->frag[i].s);
	free(r->frag);
	free(r->name), free(r->seq), free(r->qlen), free(r->sam);
	free(r);
}

/* the bases the budget lets the in-flight groups reach: the work buffers grow to the largest
 * in-flight load so far and stay, a larger load needs about bpb bytes per extra base (lock held) */
static double inflight_cap(const mbg_server_t *s)
{
	const double left = (double)(s->st.dev_budget - s->work);
	return (double)s->st.max_batch_bases + (left > 0 ? left / s->bpb : 0.0);
}

static int free_slot(const mbg_server_t *s)
{
	int k;
	for (k = 0; k < MBX_FRONT_SLOTS; ++k)
		if (!s->slot_busy[k]) return k;
	return -1;
}

/* whether the oldest request can start a group now (lock held) */
static int admissible(const mbg_server_t *s)
{
	if (s->head == 0 || free_slot(s) < 0) return 0;
	if (s->n_live == 0) return 1;          /* the oldest request always runs, alone if it must */
	if (s->so.solo) return 0;
	return (double)(s->inflight + s->head->bases) <= inflight_cap(s);
}

/* the requests of the next group, unlinked from the queue: the oldest request's kind, oldest first,
 * within what the budget allows (lock held) */
static req_t *take_group(mbg_server_t *s, int64_t *n_req, int *paired, int64_t *bases_out)
{
	req_t *b = 0, **bt = &b, **pp = &s->head, *q;
	const int kind = s->head->paired;
	const double cap = inflight_cap(s);
	int64_t bases = 0, n = 0;
	int gated = 0;
	while (*pp) {
		req_t *r = *pp;
		if (r->paired != kind) { pp = &r->next; continue; }
		if (n > 0 && ((double)(s->inflight + bases + r->bases) > cap || (s->so.max_reqs > 0 && n >= s->so.max_reqs) || s->so.solo)) {
			if ((double)(s->inflight + bases + r->bases) > cap) gated = 1;
			break;
		}
		*pp = r->next;                    /* unlink, keep the others in order */
		r->next = 0, *bt = r, bt = &r->next;
		bases += r->bases, ++n;
	}
	s->tail = 0;                          /* the last request still queued */
	for (q = s->head; q; q = q->next) s->tail = q;
	if (gated) ++s->st.gated;
	*n_req = n, *paired = kind, *bases_out = bases;
	return b;
}

/* the last fragment of a request was written: its records, in fragment order */
static void req_finish(mbg_server_t *s, req_t *r)
{
	int64_t tot = 0, o = 0;
	int32_t f;
	char *sam;
	for (f = 0; f < r->n_frag; ++f) tot += (int64_t)r->frag[f].l;
	sam = (char*)xmalloc(tot + 1);
	for (f = 0; f < r->n_frag; ++f) {
		if (r->frag[f].l) memcpy(sam + o, r->frag[f].s, r->frag[f].l);
		o += (int64_t)r->frag[f].l;
		free(r->frag[f].s);
	}
	sam[tot] = 0;
	free(r->frag);
	r->frag = 0;
	pthread_mutex_lock(&s->mu);
	r->sam = sam, r->sam_len = tot;
	r->t_done = mbg_server_now(), r->done = 1;
	pthread_cond_broadcast(&s->cv_done);
	pthread_mutex_unlock(&s->mu);
}

/* the last task of a group ended: its slot and budget go back; the rate of the memory model */
static void group_finish(mbg_server_t *s, group_t *G)
{
	int64_t work = 0, fr = 0, tot = 0;
	mbx_mem(&work, &fr, &tot);
	pthread_mutex_lock(&s->mu);
	if (s->st.max_batch_bases > s->peak_upd) {
		/* growth per extra in-flight base; the first update also holds the buffers whose size does
		 * not depend on the load, so it sets no rate */
		if (s->n_grow > 0 && work > s->work) {
			const double obs = (double)(work - s->work) / (double)(s->st.max_batch_bases - s->peak_upd);
			if (obs > s->bpb) s->bpb = obs;
		}
		++s->n_grow;
		s->peak_upd = s->st.max_batch_bases;
	}
	s->work = work;
	s->st.work_bytes = work, s->st.bytes_per_base = s->bpb;
	++s->st.batches, s->st.reqs += G->n_req, s->st.reads += G->n_seq, s->st.bases += G->bases;
	s->st.busy_s += mbg_server_now() - G->t0;
	s->slot_busy[G->slot] = 0;
	s->inflight -= G->bases, --s->n_live;
	pthread_cond_signal(&s->cv_work);
	pthread_mutex_unlock(&s->mu);
	mbg_group_free(G->g);
	free(G->qlen), free(G->seq), free(G->name);
	free(G);
}

/* one pool task: map (and pair) a fragment, write its records, and finish what it was the last of */
static void serve_task(void *data, int64_t i, int tid)
{
	mbg_server_t *s = (mbg_server_t*)data;
	const task_t t = s->tab[i >> TASK_SHIFT][i & (TASK_CHUNK - 1)];
	group_t *G = t.grp;
	req_t *r = t.req;
	const int step = G->paired ? 2 : 1;
	const mb_opt_t *opt = G->paired ? &s->opt_pe : &s->opt_se;
	mb_bseq1_t bs[2];
	kstring_t out = {0, 0, 0};
	int j;
	(void)tid;
	mbg_group_map(G->g, t.r0, step);
	memset(bs, 0, sizeof(bs));
	for (j = 0; j < step; ++j)
		bs[j].l_seq = G->qlen[t.r0 + j], bs[j].name = (char*)G->name[t.r0 + j], bs[j].seq = (char*)G->seq[t.r0 + j];
	mbg_map_fmt_fragment(&out, s->idx->l2b, opt, bs, step, &mbg_group_n_hit(G->g)[t.r0], &mbg_group_hit(G->g)[t.r0]);
	r->frag[t.frag] = out;
	if (atomic_fetch_sub(&r->left, 1) == 1) req_finish(s, r);
	if (atomic_fetch_sub(&G->left, 1) == 1) group_finish(s, G);
	if (atomic_fetch_sub(&s->tab_left[i >> TASK_SHIFT], 1) == 1) {
		free(s->tab[i >> TASK_SHIFT]);
		s->tab[i >> TASK_SHIFT] = 0;
	}
}

/* a group of requests: its front pass on the device, then its tasks into the running pool */
static void start_group(mbg_server_t *s, req_t *b, int64_t n_req, int paired, int64_t bases, int slot)
{
	const mb_opt_t *opt = paired ? &s->opt_pe : &s->opt_se;
	group_t *G = (group_t*)xcalloc(1, sizeof(group_t));
	int64_t n_seq = 0, n_task = 0, k, first;
	req_t *r;
	for (r = b; r; r = r->next) n_seq += r->n;
	if (n_seq > INT32_MAX) die("a group of more than 2^31 reads");
	G->n_req = n_req, G->n_seq = n_seq, G->bases = bases, G->paired = paired, G->slot = slot;
	G->qlen = (int32_t*)xmalloc(n_seq * sizeof(int32_t));
	G->seq = (const char**)xmalloc(n_seq * sizeof(char*));
	G->name = (const char**)xmalloc(n_seq * sizeof(char*));
	G->t0 = mbg_server_now();
	for (r = b, k = 0; r; r = r->next) {
		int32_t i;
		for (i = 0; i < r->n; ++i, ++k) G->qlen[k] = r->qlen[i], G->seq[k] = r->seq[i], G->name[k] = r->name[i];
		r->t_start = G->t0;
		r->n_frag = paired ? r->n / 2 : r->n;
		r->frag = (kstring_t*)xcalloc(r->n_frag, sizeof(kstring_t));
		atomic_init(&r->left, r->n_frag);
		n_task += r->n_frag;
	}
	G->g = mbg_group_new(opt, s->idx, (int32_t)n_seq, G->qlen, G->seq, G->name, slot);
	if (G->g == 0) die("a group could not be mapped (a methylation option without a methylation index?)");
	atomic_init(&G->left, n_task);
	for (r = b, k = 0; r; ) {              /* the tasks, in request and fragment order */
		req_t *nx = r->next;
		int32_t f;
		for (f = 0; f < r->n_frag; ++f) {
			const int64_t id = s->n_task++;
			const int64_t c = id >> TASK_SHIFT;
			task_t *t;
			if (c >= TASK_MAX_CHUNKS) die("task table full");
			if ((id & (TASK_CHUNK - 1)) == 0) {
				s->tab[c] = (task_t*)xmalloc(TASK_CHUNK * sizeof(task_t));
				atomic_init(&s->tab_left[c], TASK_CHUNK);
			}
			t = &s->tab[c][id & (TASK_CHUNK - 1)];
			t->grp = G, t->req = r, t->r0 = (int32_t)(k + (int64_t)f * (paired ? 2 : 1)), t->frag = f;
		}
		k += r->n;
		r->next = 0;
		r = nx;
	}
	first = mbs_add(s->pool, n_task);
	if (first != s->n_task - n_task) die("the task table and the run disagree");
}

static void *server_main(void *arg)
{
	mbg_server_t *s = (mbg_server_t*)arg;
	mbs_open(s->pool, 4, serve_task, mbg_map_round, s, mbg_map_flush());
	for (;;) {
		req_t *b;
		int64_t n_req, bases;
		int paired, slot;
		pthread_mutex_lock(&s->mu);
		while (!(s->head == 0 && s->stop) && !admissible(s)) pthread_cond_wait(&s->cv_work, &s->mu);
		if (s->head == 0) {                /* stopping, and nothing queued */
			pthread_mutex_unlock(&s->mu);
			break;
		}
		slot = free_slot(s);
		b = take_group(s, &n_req, &paired, &bases);
		s->slot_busy[slot] = 1, ++s->n_live;
		s->inflight += bases;
		if (s->inflight > s->st.max_batch_bases) s->st.max_batch_bases = s->inflight;
		pthread_mutex_unlock(&s->mu);
		start_group(s, b, n_req, paired, bases, slot);
	}
	mbs_close(s->pool);                    /* every admitted group's tasks finish first */
	return 0;
}

mbg_server_t *mbg_server_start(const mb_opt_t *opt_pe, const mb_opt_t *opt_se, const mb_idx_t *idx, const mbg_serve_opt_t *so)
{
	mbg_server_t *s = (mbg_server_t*)calloc(1, sizeof(mbg_server_t));
	int64_t work = 0, fr = 0, tot = 0;
	const char *env = getenv("MBG_MAP_THREADS");
	if (s == 0) return 0;
	s->opt_pe = *opt_pe, s->opt_se = *opt_se, s->idx = idx;
	if (so) s->so = *so;
	s->bpb = s->so.bytes_per_base0 > 0 ? s->so.bytes_per_base0 : 4096.0;
	s->pool = mbg_map_pool(env ? atoi(env) : opt_pe->n_thread);
	mbg_map_fmt_init(opt_pe);
	s->tab = (task_t**)xcalloc(TASK_MAX_CHUNKS, sizeof(task_t*));
	s->tab_left = (atomic_int*)xcalloc(TASK_MAX_CHUNKS, sizeof(atomic_int));
	mbx_mem(&work, &fr, &tot);
	s->work = work;
	s->st.dev_budget = s->so.dev_budget > 0 ? s->so.dev_budget : fr > 0 ? work + fr / 2 : (int64_t)1 << 34;
	s->st.bytes_per_base = s->bpb;
	pthread_mutex_init(&s->mu, 0);
	pthread_cond_init(&s->cv_work, 0);
	pthread_cond_init(&s->cv_done, 0);
	if (pthread_create(&s->th, 0, server_main, s) != 0) {
		free(s->tab), free(s->tab_left), free(s);
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
	r = (req_t*)xcalloc(1, sizeof(req_t));
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
	pthread_join(s->th, 0);               /* after the run closed: every task has ended */
	for (i = 0; i < s->n_all; ++i)
		if (s->all[i]) free_req(s->all[i]);
	for (i = 0; i < TASK_MAX_CHUNKS && i <= (s->n_task >> TASK_SHIFT); ++i) free(s->tab[i]);
	free(s->all), free(s->tab), free(s->tab_left);
	pthread_mutex_destroy(&s->mu);
	pthread_cond_destroy(&s->cv_work);
	pthread_cond_destroy(&s->cv_done);
	free(s);
}

void mbg_server_stats(mbg_server_t *s, mbg_serve_stats_t *st)
{
	pthread_mutex_lock(&s->mu);
	*st = s->st;
	pthread_mutex_unlock(&s->mu);
}
