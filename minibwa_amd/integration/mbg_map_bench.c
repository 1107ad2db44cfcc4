/*
 * mbg_map_bench.c -- the batched GPU mapper timed over several passes of the same inputs, with the
 * index loaded (and made resident in HBM) once:
 *
 *   mbg-map-bench [-f] [-t threads] [-P passes] [-o first.sam] [-O rest.sam] <idx> <reads.fq> [mates.fq]
 *
 * Each pass is one call of the reference's mb_map_file (map-main.c:276-299) -- the function
 * `minibwa map` runs after parsing its options and loading the index -- with minibwa's default options
 * (mb_opt_init: the adaptive preset) and -t threads, so a pass is exactly a `minibwa map` run without
 * the index load: FASTQ reading and mini-batching, mapping, pairing and SAM output (PAF with -f, as
 * `minibwa map -f`, which writes no header: map-main.c:564-569).  The first pass
 * writes to -o (default /dev/null), later passes to -O (default /dev/null).  The wall time of each pass
 * is printed to stderr as "[mbg_bench] pass <i> <seconds>", the index load as "[mbg_bench] load <s>".
 *
 * MBG_PROF_OUT=<file>: sample the program counter of every thread the mapper creates (and the main
 * thread) once per millisecond of that thread's CPU time, during the passes after the first, and write
 * the samples and /proc/self/maps to <file> at the end, for an offline flat profile of the host side.
 * Each sample also holds the word at the top of the stack, which inside a leaf function (libc's
 * memcpy and the like) is its return address, so that such time can be charged to the caller.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <signal.h>
#include <unistd.h>
#include <stdatomic.h>
#include <sys/time.h>
#include <sys/resource.h>
#include <ucontext.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include "kommon.h"
#include "mbpriv.h"
#include "ketopt.h"

int32_t mb_map_file(const mb_opt_t *opt, const mb_idx_t *idx, int32_t n, const char **fn, const char *hdr, const char *fn_out);

#define PROF_CAP (1 << 24)
static uintptr_t *g_prof, *g_prof_ret;
static atomic_long g_nprof;
static atomic_int g_prof_on;
static int g_prof_req;                   /* MBG_PROF_OUT is set: every thread gets a CPU-time timer */

static void prof_handler(int sig, siginfo_t *si, void *uc_)
{
	ucontext_t *uc = (ucontext_t*)uc_;
	long k;
	(void)sig, (void)si;
	if (!atomic_load(&g_prof_on)) return;
	k = atomic_fetch_add(&g_nprof, 1);
	if (k < PROF_CAP) {              /* the PC, and the word at the top of the stack: a leaf function's return address */
		g_prof[k] = (uintptr_t)uc->uc_mcontext.gregs[REG_RIP];
		g_prof_ret[k] = *(const uintptr_t*)uc->uc_mcontext.gregs[REG_RSP];
	}
}

/* a timer on the calling thread's own CPU clock, signalling that thread every millisecond it runs */
static void prof_thread_timer(void)
{
	timer_t tm;
	struct sigevent ev;
	struct itimerspec it;
	memset(&ev, 0, sizeof(ev));
	ev.sigev_notify = SIGEV_THREAD_ID;
	ev.sigev_signo = SIGPROF;
	ev._sigev_un._tid = gettid();
	if (timer_create(CLOCK_THREAD_CPUTIME_ID, &ev, &tm) != 0) return;
	it.it_interval.tv_sec = 0, it.it_interval.tv_nsec = 1000000;
	it.it_value = it.it_interval;
	timer_settime(tm, 0, &it, 0);
}

typedef struct { void *(*fn)(void*); void *arg; } tramp_t;
static void *prof_tramp(void *a)
{
	tramp_t t = *(tramp_t*)a;
	free(a);
	prof_thread_timer();
	return t.fn(t.arg);
}

/* threads created by the mapper (kt_pipeline, kt_for, the worker pools) are linked to come here */
int __real_pthread_create(pthread_t *th, const pthread_attr_t *attr, void *(*fn)(void*), void *arg);
int __wrap_pthread_create(pthread_t *th, const pthread_attr_t *attr, void *(*fn)(void*), void *arg)
{
	tramp_t *t;
	if (!g_prof_req || (t = (tramp_t*)malloc(sizeof(tramp_t))) == 0) return __real_pthread_create(th, attr, fn, arg);
	t->fn = fn, t->arg = arg;
	return __real_pthread_create(th, attr, prof_tramp, t);
}

static void prof_init(void)
{
	struct sigaction sa;
	g_prof_req = 1;
	g_prof = (uintptr_t*)calloc(PROF_CAP, sizeof(uintptr_t));
	g_prof_ret = (uintptr_t*)calloc(PROF_CAP, sizeof(uintptr_t));
	memset(&sa, 0, sizeof(sa));
	sa.sa_sigaction = prof_handler;
	sa.sa_flags = SA_SIGINFO | SA_RESTART;
	sigaction(SIGPROF, &sa, 0);
	prof_thread_timer();
}

static void prof_stop(const char *fn)
{
	FILE *fp, *maps;
	long i, n;
	char line[4096];
	atomic_store(&g_prof_on, 0);
	n = atomic_load(&g_nprof);
	if (n > PROF_CAP) n = PROF_CAP;
	if ((fp = fopen(fn, "w")) == 0) return;
	maps = fopen("/proc/self/maps", "r");
	while (maps && fgets(line, sizeof(line), maps)) fprintf(fp, "M %s", line);
	if (maps) fclose(maps);
	for (i = 0; i < n; ++i) fprintf(fp, "P %lx %lx\n", (unsigned long)g_prof[i], (unsigned long)g_prof_ret[i]);
	fclose(fp);
}

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* user + system CPU time of the whole process */
static double cpu_s(void)
{
	struct rusage ru;
	getrusage(RUSAGE_SELF, &ru);
	return ru.ru_utime.tv_sec + ru.ru_stime.tv_sec + 1e-6 * (ru.ru_utime.tv_usec + ru.ru_stime.tv_usec);
}

int main(int argc, char *argv[])
{
	ketopt_t o = KETOPT_INIT;
	int c, i, passes = 1;
	const char *out0 = "/dev/null", *out1 = "/dev/null";
	mb_opt_t opt;
	mb_idx_t *idx;
	kstring_t hdr = {0, 0, 0};
	double t0;
	kom_realtime();
	if (getenv("MBG_PROF_OUT")) prof_init();
	mb_opt_init(&opt);
	while ((c = ketopt(&o, argc, argv, 1, "ft:P:o:O:", 0)) >= 0) {
		if (c == 'f') opt.flag |= MB_F_PAF;
		else if (c == 't') opt.n_thread = atoi(o.arg);
		else if (c == 'P') passes = atoi(o.arg);
		else if (c == 'o') out0 = o.arg;
		else if (c == 'O') out1 = o.arg;
	}
	if (argc - o.ind < 2 || passes < 1) {
		fprintf(stderr, "Usage: mbg-map-bench [-f] [-t threads] [-P passes] [-o first.sam] [-O rest.sam] <idx> <reads.fq> [mates.fq]\n");
		return 1;
	}
	t0 = now_s();
	idx = mb_idx_load(argv[o.ind], 0);
	if (idx == 0) {
		fprintf(stderr, "[mbg_bench] failed to load the index '%s'\n", argv[o.ind]);
		return 1;
	}
	fprintf(stderr, "[mbg_bench] load %.3f\n", now_s() - t0);
	if (!(opt.flag & MB_F_PAF) && mb_fmt_sam_hdr(&hdr, idx->l2b, 0, MB_VERSION, argc, argv) < 0) return 1;
	for (i = 0; i < passes; ++i) {
		double t1 = now_s(), c1 = cpu_s();
		if (i == 1 && g_prof_req) atomic_store(&g_prof_on, 1);
		if (mb_map_file(&opt, idx, argc - (o.ind + 1), (const char**)&argv[o.ind + 1], hdr.s, i == 0 ? out0 : out1) != 0) {
			fprintf(stderr, "[mbg_bench] mapping failed\n");
			return 1;
		}
		fprintf(stderr, "[mbg_bench] pass %d %.4f cpu %.3f\n", i, now_s() - t1, cpu_s() - c1);
	}
	if (g_prof_req) prof_stop(getenv("MBG_PROF_OUT"));
	free(hdr.s);
	mb_idx_destroy(idx);
	return 0;
}
