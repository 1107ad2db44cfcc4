/*
 * mbg_serve_bench.c -- the serving loop (mbg_serve.c) driven by a request schedule:
 *
 *   mbg-serve-bench [-t threads] [-I avg,std] [-b budget_MiB] [-s] [-o out_prefix] <idx> <r1.fa> <r2.fa> <long.fa> <schedule>
 *
 * The schedule has one request per line, "<arrival seconds> P <n>" for n read pairs or
 * "<arrival seconds> L 1" for one single read; requests take their reads from the pools in order
 * (pairs from r1.fa/r2.fa, single reads from long.fa; FASTA with one line per sequence), wrapping
 * around at a pool's end.  A client thread submits each request at its arrival time; the main thread
 * collects the requests in order and measures each one's latency from its arrival to its records
 * being back on the host.  Pairs use predefined insert-size statistics (-I avg,std, default 400,50,
 * converted as map-main.c:424-433 converts `minibwa map -I`), single reads the defaults without
 * pairing.  -s runs every request in a batch of its own (the batch-invariance reference).  With -o the
 * records of the pair requests go to <prefix>.pairs.sam and those of the single reads to
 * <prefix>.long.sam, each in request order.  One JSON line on stdout: offered and achieved request
 * rates, read pairs per second, latency percentiles, and the server's batching statistics.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "kommon.h"
#include "mbpriv.h"
#include "ketopt.h"
#include "mbg_serve.h"

typedef struct { int64_t n; char **name, **seq; } pool_t;
typedef struct { double t; int paired; int32_t n; } sched_t;

static void die(const char *m)
{
	fprintf(stderr, "[mbg_serve_bench] %s\n", m);
	exit(1);
}

static void read_pool(const char *fn, pool_t *p)
{
	FILE *fp = fopen(fn, "r");
	char *line = 0;
	size_t cap = 0;
	ssize_t l;
	int64_t m = 0;
	if (fp == 0) die("cannot open a read pool");
	memset(p, 0, sizeof(*p));
	while ((l = getline(&line, &cap, fp)) > 0) {
		while (l > 0 && (line[l - 1] == '\n' || line[l - 1] == '\r')) line[--l] = 0;
		if (line[0] == '>') {
			char *e = line + 1;
			while (*e && *e != ' ' && *e != '\t') ++e;
			*e = 0;
			if (p->n == m) {
				m = m ? m * 2 : 1024;
				p->name = (char**)realloc(p->name, m * sizeof(char*));
				p->seq = (char**)realloc(p->seq, m * sizeof(char*));
			}
			p->name[p->n] = strdup(line + 1), p->seq[p->n] = 0, ++p->n;
		} else if (p->n > 0 && p->seq[p->n - 1] == 0) p->seq[p->n - 1] = strdup(line);
	}
	free(line);
	fclose(fp);
}

typedef struct {
	mbg_server_t *srv;
	const sched_t *sc;
	int64_t n_sc;
	pool_t *r1, *r2, *lr;
	double t0;
	int solo;
	int64_t *ids;
	pthread_mutex_t mu;
	pthread_cond_t cv;
	int64_t n_submitted;
} client_t;

static void sleep_until(double t)
{
	for (;;) {
		double d = t - mbg_server_now();
		struct timespec ts;
		if (d <= 0) return;
		ts.tv_sec = (time_t)d, ts.tv_nsec = (long)((d - (double)ts.tv_sec) * 1e9);
		nanosleep(&ts, 0);
	}
}

static void *client_main(void *arg)
{
	client_t *c = (client_t*)arg;
	int64_t i, cp = 0, cl = 0, m = 0;
	const char **name = 0, **seq = 0;
	for (i = 0; i < c->n_sc; ++i) {
		const sched_t *s = &c->sc[i];
		int32_t j, n = s->paired ? 2 * s->n : s->n;
		int64_t id;
		if (n > m) {
			m = n;
			name = (const char**)realloc(name, m * sizeof(char*));
			seq = (const char**)realloc(seq, m * sizeof(char*));
		}
		for (j = 0; j < s->n; ++j) {
			if (s->paired) {
				const int64_t k = cp++ % c->r1->n;
				name[2 * j] = c->r1->name[k], seq[2 * j] = c->r1->seq[k];
				name[2 * j + 1] = c->r2->name[k], seq[2 * j + 1] = c->r2->seq[k];
			} else {
				const int64_t k = cl++ % c->lr->n;
				name[j] = c->lr->name[k], seq[j] = c->lr->seq[k];
			}
		}
		if (!c->solo) sleep_until(c->t0 + s->t);
		id = mbg_server_submit(c->srv, s->paired, n, name, seq);
		if (id != i) die("request ids out of step");
		pthread_mutex_lock(&c->mu);
		c->n_submitted = i + 1;
		pthread_cond_signal(&c->cv);
		pthread_mutex_unlock(&c->mu);
		if (c->solo) {   /* the next request only after this one is done: one request per batch, no queue */
			pthread_mutex_lock(&c->mu);
			while (c->ids[i] == 0) pthread_cond_wait(&c->cv, &c->mu);
			pthread_mutex_unlock(&c->mu);
		}
	}
	free(name), free(seq);
	return 0;
}

static int cmp_d(const void *a, const void *b)
{
	double x = *(const double*)a, y = *(const double*)b;
	return x < y ? -1 : x > y;
}

int main(int argc, char *argv[])
{
	ketopt_t o = KETOPT_INIT;
	int c, solo = 0;
	double avg = 400, sd = 50, budget_mib = 0;
	const char *prefix = 0;
	mb_opt_t ope, ose;
	mb_idx_t *idx;
	pool_t r1, r2, lr;
	sched_t *sc = 0;
	int64_t n_sc = 0, m_sc = 0, i, n_pairs = 0, n_long = 0;
	FILE *fs, *fo_p = 0, *fo_l = 0;
	char kind;
	double t, *lat;
	int32_t n;
	mbg_serve_opt_t so;
	mbg_server_t *srv;
	client_t cl;
	pthread_t th;
	mbg_serve_stats_t st;

	kom_realtime();
	mb_opt_init(&ope);
	while ((c = ketopt(&o, argc, argv, 1, "t:I:b:so:", 0)) >= 0) {
		if (c == 't') ope.n_thread = atoi(o.arg);
		else if (c == 'I') {
			char *q;
			avg = strtod(o.arg, &q);
			if (*q == ',') sd = strtod(q + 1, &q);
		} else if (c == 'b') budget_mib = atof(o.arg);
		else if (c == 's') solo = 1;
		else if (c == 'o') prefix = o.arg;
	}
	if (argc - o.ind < 5) {
		fprintf(stderr, "Usage: mbg-serve-bench [-t threads] [-I avg,std] [-b budget_MiB] [-s] [-o out_prefix] <idx> <r1.fa> <r2.fa> <long.fa> <schedule>\n");
		return 1;
	}
	/* -I avg,std as set_ins_size (map-main.c:424-433) sets it */
	ope.pe_avg = (int32_t)(.499 + avg);
	ope.pe_std = (int32_t)(.499 + sd);
	ope.pe_hi = ope.pe_avg + ope.pe_std * 4;
	ope.pe_lo = ope.pe_avg - ope.pe_std * 4;
	if (ope.pe_lo < 1) ope.pe_lo = 1;
	ope.flag |= MB_F_PE | MB_F_PE_PREDEF;
	ose = ope;
	ose.flag &= ~(uint64_t)(MB_F_PE | MB_F_PE_PREDEF);

	idx = mb_idx_load(argv[o.ind], 0);
	if (idx == 0) die("cannot load the index");
	read_pool(argv[o.ind + 1], &r1), read_pool(argv[o.ind + 2], &r2), read_pool(argv[o.ind + 3], &lr);
	if (r1.n == 0 || r1.n != r2.n || lr.n == 0) die("empty or unequal read pools");
	if ((fs = fopen(argv[o.ind + 4], "r")) == 0) die("cannot open the schedule");
	while (fscanf(fs, "%lf %c %d", &t, &kind, &n) == 3) {
		if (n_sc == m_sc) m_sc = m_sc ? m_sc * 2 : 1024, sc = (sched_t*)realloc(sc, m_sc * sizeof(sched_t));
		sc[n_sc].t = t, sc[n_sc].paired = kind == 'P', sc[n_sc].n = n;
		if (kind == 'P') n_pairs += n;
		else n_long += n;
		++n_sc;
	}
	fclose(fs);
	if (n_sc == 0) die("empty schedule");
	if (prefix) {
		kstring_t f = {0, 0, 0};
		kom_sprintf_lite(&f, "%s.pairs.sam", prefix);
		fo_p = fopen(f.s, "w");
		f.l = 0;
		kom_sprintf_lite(&f, "%s.long.sam", prefix);
		fo_l = fopen(f.s, "w");
		free(f.s);
		if (fo_p == 0 || fo_l == 0) die("cannot open the outputs");
	}

	memset(&so, 0, sizeof(so));
	so.dev_budget = (int64_t)(budget_mib * 1048576.0);
	so.solo = solo;
	srv = mbg_server_start(&ope, &ose, idx, &so);
	if (srv == 0) die("cannot start the server");
	memset(&cl, 0, sizeof(cl));
	cl.srv = srv, cl.sc = sc, cl.n_sc = n_sc, cl.r1 = &r1, cl.r2 = &r2, cl.lr = &lr, cl.solo = solo;
	cl.ids = (int64_t*)calloc(n_sc, sizeof(int64_t));
	pthread_mutex_init(&cl.mu, 0);
	pthread_cond_init(&cl.cv, 0);
	lat = (double*)malloc(n_sc * sizeof(double));
	cl.t0 = mbg_server_now() + 0.05;
	if (pthread_create(&th, 0, client_main, &cl) != 0) die("pthread_create");
	for (i = 0; i < n_sc; ++i) {
		char *sam;
		int64_t len;
		double ts, td;
		pthread_mutex_lock(&cl.mu);
		while (cl.n_submitted <= i) pthread_cond_wait(&cl.cv, &cl.mu);
		pthread_mutex_unlock(&cl.mu);
		if (mbg_server_wait(srv, i, &sam, &len, &ts, &td) != 0) die("a request went missing");
		lat[i] = td - (solo ? ts : cl.t0 + sc[i].t);
		if (fo_p && sam && len > 0) fwrite(sam, 1, len, sc[i].paired ? fo_p : fo_l);
		free(sam);
		if (solo) {
			pthread_mutex_lock(&cl.mu);
			cl.ids[i] = 1;
			pthread_cond_broadcast(&cl.cv);
			pthread_mutex_unlock(&cl.mu);
		}
	}
	{
		const double el = mbg_server_now() - cl.t0;
		double *sl = (double*)malloc(n_sc * sizeof(double));
		pthread_join(th, 0);
		mbg_server_stats(srv, &st);
		memcpy(sl, lat, n_sc * sizeof(double));
		qsort(sl, n_sc, sizeof(double), cmp_d);
		printf("{\"requests\": %ld, \"pairs\": %ld, \"long_reads\": %ld, \"span_s\": %.4f, \"offered_qps\": %.2f, "
		       "\"achieved_qps\": %.2f, \"pairs_per_s\": %.1f, \"p50_ms\": %.3f, \"p90_ms\": %.3f, \"p99_ms\": %.3f, \"max_ms\": %.3f, "
		       "\"batches\": %ld, \"reqs_per_batch\": %.2f, \"max_batch_bases\": %ld, \"gated_batches\": %ld, "
		       "\"bytes_per_base\": %.1f, \"dev_budget\": %ld, \"work_bytes\": %ld, \"busy_s\": %.4f, \"solo\": %d}\n",
		       (long)n_sc, (long)n_pairs, (long)n_long, sc[n_sc - 1].t, n_sc / (sc[n_sc - 1].t > 0 ? sc[n_sc - 1].t : 1.0),
		       n_sc / el, n_pairs / el, 1e3 * sl[n_sc / 2], 1e3 * sl[(int64_t)(0.9 * (n_sc - 1))],
		       1e3 * sl[(int64_t)(0.99 * (n_sc - 1))], 1e3 * sl[n_sc - 1], (long)st.batches, st.batches ? (double)st.reqs / st.batches : 0.0,
		       (long)st.max_batch_bases, (long)st.gated, st.bytes_per_base, (long)st.dev_budget, (long)st.work_bytes, st.busy_s, solo);
		free(sl);
	}
	mbg_server_stop(srv);
	if (fo_p) fclose(fo_p);
	if (fo_l) fclose(fo_l);
	mb_idx_destroy(idx);
	return 0;
}
