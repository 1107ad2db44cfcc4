/*
 * mbg_sched.c -- coroutine worker pool with batch rounds (interface and rationale in mbg_sched.h).
 *
 * Each pool thread keeps its own suspended coroutines and its own stacks, so a coroutine always
 * resumes on the thread that started it (the reference's per-thread kalloc pools and mimalloc's
 * thread heaps stay consistent).  A round:
 *   1. every thread resumes the coroutines it suspended in the previous round, then starts new tasks
 *      until none are left; each coroutine runs until it yields (a request posted) or finishes;
 *   2. all threads meet at a barrier; one thread runs the round callback over every posted request;
 *   3. all threads meet again; if no coroutine was suspended the run is over.
 * Stacks come from mmap'd chunks (MAP_NORESERVE: only touched pages take memory).  Guard pages would
 * cost two memory mappings per stack, and a round can hold hundreds of thousands of suspended
 * coroutines, beyond the kernel's mapping limit; instead the lowest bytes of each stack hold a canary
 * that is checked whenever its coroutine yields or finishes, and a damaged canary aborts.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/mman.h>
#include <time.h>
#include <unistd.h>
#include "mbg_sched.h"

typedef struct thr_s thr_t;

typedef struct co_s {
	void *sp;                 /* saved stack pointer while suspended */
	char *top;                /* one past the highest stack byte (16-byte aligned) */
	uint64_t *canary;         /* the stack's lowest bytes */
	int64_t task;
	int done;
	thr_t *thr;
	struct co_s *next_free;
} co_t;

struct thr_s {
	mbs_pool_t *pool;
	int tid;
	pthread_t th;
	void *sched_sp;           /* the thread's scheduler context while a coroutine runs */
	co_t **susp, **susp2;     /* suspended coroutines: current round's and the one being resumed */
	int64_t n_susp, m_susp, m_susp2;
	co_t *free_co;
	int64_t n_co;             /* coroutine objects (each with a stack) owned by this thread */
	int64_t yields;
	double busy_s;            /* time spent running coroutines */
};

struct mbs_pool_s {
	int n;
	int64_t stack_bytes;
	thr_t *thr;
	pthread_barrier_t bar;
	pthread_mutex_t mu, run_mu;
	pthread_cond_t cv_start, cv_done;
	int64_t gen;
	int n_finished, stop;
	/* the current run */
	int64_t n_tasks, chunk;
	atomic_llong next;
	mbs_task_f task;
	mbs_round_f round;
	void *data;
	atomic_llong live;
	int finish;
	/* statistics */
	int64_t rounds, co_peak;
	double round_s, task_s, phase_s, phase_t0;
};

static __thread thr_t *tls_thr;
static __thread co_t *tls_co;

/* void mbs_switch(void **save_sp, void *load_sp): saves the callee-saved registers on the current
 * stack, stores the stack pointer at *save_sp, switches to load_sp and restores the registers saved
 * there.  mbs_trampoline is where a new coroutine starts: its co_t* arrives in rbx. */
void mbs_switch(void **save_sp, void *load_sp);
void mbs_trampoline(void);
__asm__(
	".text\n"
	".globl mbs_switch\n"
	".type mbs_switch,@function\n"
	"mbs_switch:\n"
	"\tpushq %rbp\n\tpushq %rbx\n\tpushq %r12\n\tpushq %r13\n\tpushq %r14\n\tpushq %r15\n"
	"\tmovq %rsp, (%rdi)\n"
	"\tmovq %rsi, %rsp\n"
	"\tpopq %r15\n\tpopq %r14\n\tpopq %r13\n\tpopq %r12\n\tpopq %rbx\n\tpopq %rbp\n"
	"\tret\n"
	".size mbs_switch,.-mbs_switch\n"
	".globl mbs_trampoline\n"
	".type mbs_trampoline,@function\n"
	"mbs_trampoline:\n"
	"\tmovq %rbx, %rdi\n"
	"\tcall mbs_co_main\n"
	"\tud2\n"
	".size mbs_trampoline,.-mbs_trampoline\n");

static double now_s(void)
{
	struct timespec ts;
	clock_gettime(CLOCK_MONOTONIC, &ts);
	return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static void die(const char *msg)
{
	fprintf(stderr, "[mbg_sched] %s\n", msg);
	abort();
}

void mbs_co_main(co_t *co);
void mbs_co_main(co_t *co)
{
	thr_t *t = co->thr;
	mbs_pool_t *p = t->pool;
	p->task(p->data, co->task, t->tid);
	co->done = 1;
	mbs_switch(&co->sp, t->sched_sp);
	die("finished coroutine resumed");
}

static void push_susp(thr_t *t, co_t *co)
{
	if (t->n_susp == t->m_susp) {
		t->m_susp = t->m_susp ? t->m_susp * 2 : 1024;
		t->susp = (co_t**)realloc(t->susp, t->m_susp * sizeof(co_t*));
		if (t->susp == 0) die("out of memory (suspended list)");
	}
	t->susp[t->n_susp++] = co;
}

#define STACKS_PER_CHUNK 64
#define CANARY_WORDS 8
static const uint64_t kCanary = 0x6d62675f73746b21ULL;

static void check_canary(const co_t *co)
{
	int k;
	for (k = 0; k < CANARY_WORDS; ++k)
		if (co->canary[k] != kCanary) die("coroutine stack overflow (canary overwritten)");
}

static co_t *get_co(thr_t *t)
{
	co_t *co = t->free_co;
	if (co) {
		t->free_co = co->next_free;
		return co;
	}
	{
		const size_t one = (size_t)t->pool->stack_bytes;
		char *base = (char*)mmap(0, one * STACKS_PER_CHUNK, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
		co_t *objs = (co_t*)calloc(STACKS_PER_CHUNK, sizeof(co_t));
		int k, w;
		if (base == MAP_FAILED || objs == 0) die("out of memory (coroutine stacks)");
		for (k = 0; k < STACKS_PER_CHUNK; ++k) {
			char *lo = base + one * k;
			objs[k].canary = (uint64_t*)lo;
			for (w = 0; w < CANARY_WORDS; ++w) objs[k].canary[w] = kCanary;
			objs[k].top = (char*)(((uintptr_t)(lo + one)) & ~(uintptr_t)15);
			objs[k].thr = t;
			objs[k].next_free = k + 1 < STACKS_PER_CHUNK ? &objs[k + 1] : 0;
		}
		t->n_co += STACKS_PER_CHUNK;
		t->free_co = objs[0].next_free;
		return &objs[0];
	}
}

static void put_co(thr_t *t, co_t *co)
{
	co->next_free = t->free_co;
	t->free_co = co;
}

static void resume(thr_t *t, co_t *co)
{
	tls_co = co;
	mbs_switch(&t->sched_sp, co->sp);
	tls_co = 0;
	check_canary(co);
}

static void start(thr_t *t, co_t *co, int64_t task)
{
	/* initial frame, read by mbs_switch's pops: r15 r14 r13 r12 rbx(=co) rbp, then the return address;
	 * after the ret the stack pointer is top-16, 16-byte aligned as the call in mbs_trampoline needs */
	void **sp = (void**)(co->top - 16 - 7 * sizeof(void*));
	sp[0] = sp[1] = sp[2] = sp[3] = 0;
	sp[4] = co;
	sp[5] = 0;
	sp[6] = (void*)mbs_trampoline;
	co->sp = sp;
	co->task = task;
	co->done = 0;
	resume(t, co);
}

static void run_job(thr_t *t)
{
	mbs_pool_t *p = t->pool;
	for (;;) {
		int64_t i, n_old, i0;
		co_t **old;
		double tb = now_s();
		/* 1. resume what the previous round suspended */
		old = t->susp, n_old = t->n_susp;
		t->susp = t->susp2, t->susp2 = old;
		{ int64_t m = t->m_susp; t->m_susp = t->m_susp2; t->m_susp2 = m; }
		t->n_susp = 0;
		for (i = 0; i < n_old; ++i) {
			co_t *co = old[i];
			resume(t, co);
			if (co->done) put_co(t, co);
			else push_susp(t, co);
		}
		/* 2. start new tasks */
		while ((i0 = atomic_fetch_add(&p->next, p->chunk)) < p->n_tasks) {
			int64_t e = i0 + p->chunk < p->n_tasks ? i0 + p->chunk : p->n_tasks;
			for (i = i0; i < e; ++i) {
				co_t *co = get_co(t);
				start(t, co, i);
				if (co->done) put_co(t, co);
				else push_susp(t, co);
			}
		}
		atomic_fetch_add(&p->live, t->n_susp);
		t->busy_s += now_s() - tb;
		/* 3. the round */
		if (pthread_barrier_wait(&p->bar) == PTHREAD_BARRIER_SERIAL_THREAD) {
			int64_t live = atomic_load(&p->live);
			p->phase_s += now_s() - p->phase_t0;
			int64_t n_co = 0;
			int k;
			for (k = 0; k < p->n; ++k) n_co += p->thr[k].n_susp;
			if (n_co > p->co_peak) p->co_peak = n_co;
			if (live > 0) {
				double t0 = now_s();
				if (p->round == 0) die("a task yielded but the run has no round callback");
				p->round(p->data, p->n);
				p->round_s += now_s() - t0;
				++p->rounds;
			}
			p->finish = live == 0;
			atomic_store(&p->live, 0);
			p->phase_t0 = now_s();
		}
		pthread_barrier_wait(&p->bar);
		if (p->finish) break;
	}
}

static void *worker(void *arg)
{
	thr_t *t = (thr_t*)arg;
	mbs_pool_t *p = t->pool;
	int64_t my_gen = 0;
	tls_thr = t;
	for (;;) {
		pthread_mutex_lock(&p->mu);
		while (p->gen == my_gen && !p->stop) pthread_cond_wait(&p->cv_start, &p->mu);
		if (p->stop) { pthread_mutex_unlock(&p->mu); break; }
		my_gen = p->gen;
		pthread_mutex_unlock(&p->mu);
		run_job(t);
		pthread_mutex_lock(&p->mu);
		if (++p->n_finished == p->n) pthread_cond_signal(&p->cv_done);
		pthread_mutex_unlock(&p->mu);
	}
	return 0;
}

mbs_pool_t *mbs_pool_init(int n_threads, int64_t stack_bytes)
{
	mbs_pool_t *p;
	int i;
	const long pg = sysconf(_SC_PAGESIZE);
	if (n_threads < 1) n_threads = 1;
	if (stack_bytes < 65536) stack_bytes = 65536;
	stack_bytes = (stack_bytes + pg - 1) / pg * pg;
	p = (mbs_pool_t*)calloc(1, sizeof(mbs_pool_t));
	if (p == 0) return 0;
	p->n = n_threads;
	p->stack_bytes = stack_bytes;
	p->thr = (thr_t*)calloc(n_threads, sizeof(thr_t));
	pthread_barrier_init(&p->bar, 0, n_threads);
	pthread_mutex_init(&p->mu, 0);
	pthread_mutex_init(&p->run_mu, 0);
	pthread_cond_init(&p->cv_start, 0);
	pthread_cond_init(&p->cv_done, 0);
	for (i = 0; i < n_threads; ++i) {
		p->thr[i].pool = p, p->thr[i].tid = i;
		if (pthread_create(&p->thr[i].th, 0, worker, &p->thr[i]) != 0) die("pthread_create failed");
	}
	return p;
}

void mbs_pool_destroy(mbs_pool_t *p)
{
	int i;
	if (p == 0) return;
	pthread_mutex_lock(&p->mu);
	p->stop = 1;
	pthread_cond_broadcast(&p->cv_start);
	pthread_mutex_unlock(&p->mu);
	for (i = 0; i < p->n; ++i) pthread_join(p->thr[i].th, 0);
	for (i = 0; i < p->n; ++i) {
		free(p->thr[i].susp);
		free(p->thr[i].susp2);
		/* stacks and coroutine objects stay mapped: the pool lives for the process */
	}
	pthread_barrier_destroy(&p->bar);
	free(p->thr);
	free(p);
}

int mbs_pool_size(const mbs_pool_t *p) { return p->n; }

void mbs_run(mbs_pool_t *p, int64_t n, int64_t chunk, mbs_task_f task, mbs_round_f round, void *data)
{
	double t0;
	if (n <= 0) return;
	pthread_mutex_lock(&p->run_mu);
	t0 = now_s();
	pthread_mutex_lock(&p->mu);
	p->n_tasks = n, p->chunk = chunk > 0 ? chunk : 1;
	atomic_store(&p->next, 0);
	atomic_store(&p->live, 0);
	p->task = task, p->round = round, p->data = data;
	p->finish = 0;
	p->n_finished = 0;
	p->phase_t0 = now_s();
	++p->gen;
	pthread_cond_broadcast(&p->cv_start);
	while (p->n_finished < p->n) pthread_cond_wait(&p->cv_done, &p->mu);
	pthread_mutex_unlock(&p->mu);
	p->task_s += now_s() - t0;
	pthread_mutex_unlock(&p->run_mu);
}

void mbs_yield(void)
{
	co_t *co = tls_co;
	thr_t *t = tls_thr;
	if (co == 0 || t == 0) die("mbs_yield outside a task");
	++t->yields;
	mbs_switch(&co->sp, t->sched_sp);
	tls_co = co;
}

int mbs_in_task(void) { return tls_co != 0; }
int mbs_tid(void) { return tls_co ? tls_thr->tid : -1; }

void mbs_get_stats(const mbs_pool_t *p, mbs_stats_t *st)
{
	int i;
	memset(st, 0, sizeof(*st));
	st->rounds = p->rounds, st->coroutines_peak = p->co_peak;
	st->round_s = p->round_s, st->task_s = p->task_s, st->phase_s = p->phase_s;
	for (i = 0; i < p->n; ++i) st->yields += p->thr[i].yields, st->stacks += p->thr[i].n_co, st->busy_s += p->thr[i].busy_s;
}
