/*
 * mbg_sched.c -- coroutine worker pool with an asynchronous batch dispatcher (interface and rationale
 * in mbg_sched.h).
 *
 * Each pool thread keeps its own coroutines and stacks, so a coroutine always resumes on the thread
 * that started it (the reference's per-thread kalloc pools and mimalloc's thread heaps stay
 * consistent).  A worker loops over: the coroutines whose requests have finished (handed back by the
 * dispatcher), then a chunk of new tasks; each coroutine runs until it posts a request (mbs_wait) or
 * finishes.  With nothing to run, a worker sleeps until one of its requests finishes.
 *
 * The dispatcher thread starts a batch when `flush` requests are pending, when half the workers are
 * idle and an eighth of that is pending, or when every worker is idle; it takes every posted request,
 * runs the batch callback, and hands each coroutine back to its thread.  Requests posted while a batch
 * is on the device wait for the next one, so host work and device work overlap.
 *
 * Termination: each thread counts its tasks claimed and not finished (`live`, raised before a claim).
 * When every task has been claimed and every thread's live count is zero, no request can be pending,
 * and the dispatcher ends the run.  The counters the workers touch per task or per request (live, the
 * posted count) are per thread, each thread's on cache lines of its own: pool-wide atomics there were
 * contended by every worker at once.
 *
 * Stacks come from mmap'd chunks (MAP_NORESERVE: only touched pages take memory).  Guard pages would
 * cost two memory mappings per stack, and a run can hold hundreds of thousands of suspended coroutines,
 * beyond the kernel's mapping limit; instead the lowest bytes of each stack hold a canary that each
 * thread checks over all of its stacks after every run, and a damaged canary aborts.
 *
 * An open run (mbs_open) takes tasks while it runs: mbs_add makes more indices runnable and wakes the
 * idle workers, and the run ends once it is closed and every added task has finished.  Chunks are
 * claimed with a compare-and-swap that never moves past the tasks added so far, so no index is lost
 * to a claim made before it was added.
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

/* a growable array of (request, coroutine) pairs */
typedef struct { void **req; co_t **co; int64_t n, m; } plist_t;

struct thr_s {
	_Alignas(128) atomic_llong live;   /* tasks this thread claimed and has not finished */
	atomic_llong n_posted;    /* posted.n, readable without the lock */
	mbs_pool_t *pool;
	int tid;
	pthread_t th;
	void *sched_sp;           /* the thread's scheduler context while a coroutine runs */
	pthread_mutex_t mu;       /* guards ready, posted and waiting */
	pthread_cond_t cv;
	co_t **ready;             /* coroutines whose requests have finished */
	int64_t n_ready, m_ready;
	co_t **rbuf;              /* the ready list is swapped into this one to be resumed */
	int64_t m_rbuf;
	plist_t posted;           /* posted requests not yet taken by the dispatcher */
	int waiting;
	co_t *free_co;
	int64_t n_co;             /* coroutine objects (each with a stack) owned by this thread */
	co_t **chunks;            /* ... allocated STACKS_PER_CHUNK at a time: their canaries are checked after each run */
	int64_t n_chunks, m_chunks;
	int64_t waits;
	double busy_s, idle_s;
	char pad_[128];           /* keeps the next thread's counters off this thread's last line */
};

struct mbs_pool_s {
	int n, id;
	int64_t stack_bytes;
	thr_t *thr;
	pthread_mutex_t mu, run_mu;
	pthread_cond_t cv_start, cv_done;
	int64_t gen;
	int n_finished, stop;
	pthread_t disp;
	/* the current run */
	atomic_llong n_tasks;     /* tasks added so far */
	int64_t chunk, flush;
	atomic_int open_run;      /* the run still takes tasks (mbs_add) */
	atomic_llong next;
	mbs_task_f task;
	mbs_batch_f batch;
	void *data;
	int64_t flush_thr;        /* a thread's posted count that pokes the dispatcher (flush / n) */
	atomic_int n_idle;
	atomic_int all_done;
	pthread_mutex_t dmu;      /* dispatcher wake-ups */
	pthread_cond_t dcv;
	plist_t batch_l;          /* the dispatcher's batch */
	int64_t *toff;            /* per thread: where its requests start in the batch */
	double run_t0;            /* start of the current run */
	/* statistics */
	int64_t batches, co_peak;
	double batch_s, task_s;
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

static void plist_reserve(plist_t *l, int64_t need)
{
	if (need <= l->m) return;
	l->m = need + (need >> 1) + 1024;
	l->req = (void**)realloc(l->req, (size_t)l->m * sizeof(void*));
	l->co = (co_t**)realloc(l->co, (size_t)l->m * sizeof(co_t*));
	if (l->req == 0 || l->co == 0) die("out of memory (request list)");
}

#define STACKS_PER_CHUNK 64
#define CANARY_WORDS 8
static const uint64_t kCanary = 0x6d62675f73746b21ULL;

static void check_canary(const co_t *co)
{
	if (co->canary[0] != kCanary || co->canary[CANARY_WORDS - 1] != kCanary) die("coroutine stack overflow (canary overwritten)");
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
		if (t->n_chunks == t->m_chunks) {
			t->m_chunks = t->m_chunks ? t->m_chunks * 2 : 64;
			t->chunks = (co_t**)realloc(t->chunks, (size_t)t->m_chunks * sizeof(co_t*));
			if (t->chunks == 0) die("out of memory (coroutine stacks)");
		}
		t->chunks[t->n_chunks++] = objs;
		t->free_co = objs[0].next_free;
		return &objs[0];
	}
}

static void poke(mbs_pool_t *p)
{
	pthread_mutex_lock(&p->dmu);
	pthread_cond_signal(&p->dcv);
	pthread_mutex_unlock(&p->dmu);
}

static int64_t live_sum(mbs_pool_t *p)
{
	int64_t s = 0;
	int k;
	for (k = 0; k < p->n; ++k) s += atomic_load(&p->thr[k].live);
	return s;
}

static int64_t posted_sum(mbs_pool_t *p)
{
	int64_t s = 0;
	int k;
	for (k = 0; k < p->n; ++k) s += atomic_load_explicit(&p->thr[k].n_posted, memory_order_relaxed);
	return s;
}

/* every stack of the thread: a task that overflowed its stack overwrote the words at its low end
 * (checked once a run rather than at each switch, where the cold low end of a stack would cost a
 * cache miss per coroutine) */
static void check_stacks(const thr_t *t)
{
	int64_t i;
	int k;
	for (i = 0; i < t->n_chunks; ++i)
		for (k = 0; k < STACKS_PER_CHUNK; ++k) check_canary(&t->chunks[i][k]);
}

static void finish(thr_t *t, co_t *co)
{
	mbs_pool_t *p = t->pool;
	co->next_free = t->free_co;
	t->free_co = co;
	/* this thread's last task while nothing is left to claim: possibly the run's last */
	if (atomic_fetch_sub(&t->live, 1) == 1 && atomic_load(&p->next) >= atomic_load(&p->n_tasks)) poke(p);
}

static void resume(thr_t *t, co_t *co)
{
	tls_co = co;
	mbs_switch(&t->sched_sp, co->sp);
	tls_co = 0;
	if (co->done) finish(t, co);
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
		double t0 = now_s();
		int64_t i, n, i0;
		int progressed;
		/* 1. coroutines whose requests have finished */
		pthread_mutex_lock(&t->mu);
		n = t->n_ready;
		if (n) {
			co_t **r = t->ready;
			int64_t m = t->m_ready;
			t->ready = t->rbuf, t->m_ready = t->m_rbuf;
			t->rbuf = r, t->m_rbuf = m;
			t->n_ready = 0;
		}
		pthread_mutex_unlock(&t->mu);
		for (i = 0; i < n; ++i) resume(t, t->rbuf[i]);
		progressed = n > 0;
		/* 2. a chunk of new tasks; live is raised before the claim, so that the dispatcher never sees
		 * every task claimed and none live while a claimed chunk has not started */
		i0 = atomic_load(&p->next);
		if (i0 < atomic_load(&p->n_tasks)) {
			int64_t nt = 0, e = 0;
			atomic_fetch_add(&t->live, p->chunk);
			for (;;) {                      /* [i0, e): at most a chunk, never past the tasks added */
				nt = atomic_load(&p->n_tasks);
				if (i0 >= nt) break;
				e = i0 + p->chunk < nt ? i0 + p->chunk : nt;
				if (atomic_compare_exchange_weak(&p->next, &i0, e)) break;
			}
			if (i0 < nt) {
				if (e - i0 < p->chunk) atomic_fetch_sub(&t->live, p->chunk - (e - i0));
				for (i = i0; i < e; ++i) start(t, get_co(t), i);
				progressed = 1;
			} else if (atomic_fetch_sub(&t->live, p->chunk) == p->chunk) poke(p);
		}
		t->busy_s += now_s() - t0;
		if (progressed) continue;
		/* 3. nothing to run: sleep until a request of this thread finishes, tasks are added, or the run
		 * ends (mbs_add signals each thread under its lock after adding, so no addition is missed) */
		if (atomic_load(&p->all_done)) break;
		t0 = now_s();
		pthread_mutex_lock(&t->mu);
		if (t->n_ready == 0 && !atomic_load(&p->all_done) && atomic_load(&p->next) >= atomic_load(&p->n_tasks)) {
			t->waiting = 1;
			atomic_fetch_add(&p->n_idle, 1);
			pthread_mutex_unlock(&t->mu);
			poke(p);
			pthread_mutex_lock(&t->mu);
			while (t->n_ready == 0 && !atomic_load(&p->all_done) && atomic_load(&p->next) >= atomic_load(&p->n_tasks))
				pthread_cond_wait(&t->cv, &t->mu);
			atomic_fetch_sub(&p->n_idle, 1);
			t->waiting = 0;
		}
		pthread_mutex_unlock(&t->mu);
		t->idle_s += now_s() - t0;
	}
}

/* the dispatcher's part of a run: batches until every task has finished */
static void dispatch_job(mbs_pool_t *p)
{
	plist_t *b = &p->batch_l;
	for (;;) {
		int done = 0, k;
		int64_t nb = 0, i;
		pthread_mutex_lock(&p->dmu);
		for (;;) {
			int64_t pend = posted_sum(p);
			int idle = atomic_load(&p->n_idle);
			if (!atomic_load(&p->open_run) && atomic_load(&p->next) >= atomic_load(&p->n_tasks) && live_sum(p) == 0) { done = 1; break; }
			if (pend > 0 && (pend >= p->flush || idle == p->n || (idle * 2 >= p->n && pend * 8 >= p->flush))) break;
			{
				struct timespec ts;
				clock_gettime(CLOCK_REALTIME, &ts);
				ts.tv_nsec += 500000;
				if (ts.tv_nsec >= 1000000000L) ts.tv_sec += 1, ts.tv_nsec -= 1000000000L;
				pthread_cond_timedwait(&p->dcv, &p->dmu, &ts);
			}
		}
		pthread_mutex_unlock(&p->dmu);
		if (done) {
			atomic_store(&p->all_done, 1);
			for (k = 0; k < p->n; ++k) {
				pthread_mutex_lock(&p->thr[k].mu);
				pthread_cond_signal(&p->thr[k].cv);
				pthread_mutex_unlock(&p->thr[k].mu);
			}
			return;
		}
		/* take every posted request */
		for (k = 0; k < p->n; ++k) {
			thr_t *t = &p->thr[k];
			pthread_mutex_lock(&t->mu);
			plist_reserve(b, nb + t->posted.n);
			p->toff[k] = nb;
			memcpy(b->req + nb, t->posted.req, (size_t)t->posted.n * sizeof(void*));
			memcpy(b->co + nb, t->posted.co, (size_t)t->posted.n * sizeof(co_t*));
			nb += t->posted.n;
			t->posted.n = 0;
			atomic_store_explicit(&t->n_posted, 0, memory_order_relaxed);
			pthread_mutex_unlock(&t->mu);
		}
		p->toff[p->n] = nb;
		if (nb == 0) continue;
		{
			int64_t live = live_sum(p);
			double t0 = now_s();
			if (live > p->co_peak) p->co_peak = live;
			p->batch(p->data, b->req, nb);
			p->batch_s += now_s() - t0;
			++p->batches;
		}
		/* hand each coroutine back to its thread */
		for (k = 0; k < p->n; ++k) {
			thr_t *t = &p->thr[k];
			int64_t a = p->toff[k], e = p->toff[k + 1];
			if (a == e) continue;
			pthread_mutex_lock(&t->mu);
			if (t->n_ready + (e - a) > t->m_ready) {
				t->m_ready = t->n_ready + (e - a) + 1024;
				t->ready = (co_t**)realloc(t->ready, (size_t)t->m_ready * sizeof(co_t*));
				if (t->ready == 0) die("out of memory (ready list)");
			}
			for (i = a; i < e; ++i) t->ready[t->n_ready++] = b->co[i];
			if (t->waiting) pthread_cond_signal(&t->cv);
			pthread_mutex_unlock(&t->mu);
		}
	}
}

static void wait_job(mbs_pool_t *p, int64_t *my_gen)
{
	pthread_mutex_lock(&p->mu);
	while (p->gen == *my_gen && !p->stop) pthread_cond_wait(&p->cv_start, &p->mu);
	*my_gen = p->gen;
	pthread_mutex_unlock(&p->mu);
}

static void job_finished(mbs_pool_t *p)
{
	pthread_mutex_lock(&p->mu);
	if (++p->n_finished == p->n + 1) pthread_cond_signal(&p->cv_done);
	pthread_mutex_unlock(&p->mu);
}

static atomic_int g_pool_ids;     /* thread names: mbs<pool>w<i> for workers, mbs<pool>d for the dispatcher */

static void *worker(void *arg)
{
	thr_t *t = (thr_t*)arg;
	mbs_pool_t *p = t->pool;
	int64_t my_gen = 0;
	char nm[16];
	snprintf(nm, sizeof(nm), "mbs%dw%d", p->id, t->tid);
	pthread_setname_np(pthread_self(), nm);
	tls_thr = t;
	for (;;) {
		wait_job(p, &my_gen);
		if (p->stop) break;
		run_job(t);
		check_stacks(t);
		job_finished(p);
	}
	return 0;
}

static void *dispatcher(void *arg)
{
	mbs_pool_t *p = (mbs_pool_t*)arg;
	int64_t my_gen = 0;
	char nm[16];
	snprintf(nm, sizeof(nm), "mbs%dd", p->id);
	pthread_setname_np(pthread_self(), nm);
	for (;;) {
		wait_job(p, &my_gen);
		if (p->stop) break;
		dispatch_job(p);
		job_finished(p);
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
	p->id = atomic_fetch_add(&g_pool_ids, 1);
	p->stack_bytes = stack_bytes;
	p->thr = (thr_t*)aligned_alloc(128, (size_t)n_threads * sizeof(thr_t));
	if (p->thr) memset(p->thr, 0, (size_t)n_threads * sizeof(thr_t));
	p->toff = (int64_t*)calloc(n_threads + 1, sizeof(int64_t));
	if (p->thr == 0 || p->toff == 0) die("out of memory (pool)");
	pthread_mutex_init(&p->mu, 0);
	pthread_mutex_init(&p->run_mu, 0);
	pthread_mutex_init(&p->dmu, 0);
	pthread_cond_init(&p->cv_start, 0);
	pthread_cond_init(&p->cv_done, 0);
	pthread_cond_init(&p->dcv, 0);
	for (i = 0; i < n_threads; ++i) {
		thr_t *t = &p->thr[i];
		t->pool = p, t->tid = i;
		pthread_mutex_init(&t->mu, 0);
		pthread_cond_init(&t->cv, 0);
		if (pthread_create(&t->th, 0, worker, t) != 0) die("pthread_create failed");
	}
	if (pthread_create(&p->disp, 0, dispatcher, p) != 0) die("pthread_create failed");
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
	pthread_join(p->disp, 0);
	for (i = 0; i < p->n; ++i) {
		thr_t *t = &p->thr[i];
		free(t->ready); free(t->rbuf); free(t->posted.req); free(t->posted.co); free(t->chunks);
		/* stacks and coroutine objects stay mapped: the pool lives for the process */
	}
	free(p->batch_l.req); free(p->batch_l.co); free(p->toff);
	free(p->thr);
	free(p);
}

int mbs_pool_size(const mbs_pool_t *p) { return p->n; }

/* starts a run of n tasks (open: more may be added) and returns; called with run_mu held */
static void run_start(mbs_pool_t *p, int64_t n, int open, int64_t chunk, mbs_task_f task, mbs_batch_f batch, void *data, int64_t flush)
{
	int k;
	p->run_t0 = now_s();
	pthread_mutex_lock(&p->mu);
	atomic_store(&p->n_tasks, n);
	atomic_store(&p->open_run, open);
	p->chunk = chunk > 0 ? chunk : 1, p->flush = flush > 0 ? flush : 16384;
	p->flush_thr = p->flush / p->n > 0 ? p->flush / p->n : 1;
	atomic_store(&p->next, 0);
	for (k = 0; k < p->n; ++k) atomic_store(&p->thr[k].live, 0), atomic_store(&p->thr[k].n_posted, 0);
	atomic_store(&p->n_idle, 0);
	atomic_store(&p->all_done, 0);
	p->task = task, p->batch = batch, p->data = data;
	p->n_finished = 0;
	++p->gen;
	pthread_cond_broadcast(&p->cv_start);
	pthread_mutex_unlock(&p->mu);
}

/* waits for the end of the run started by run_start, then releases run_mu */
static void run_end(mbs_pool_t *p)
{
	pthread_mutex_lock(&p->mu);
	while (p->n_finished < p->n + 1) pthread_cond_wait(&p->cv_done, &p->mu);
	pthread_mutex_unlock(&p->mu);
	p->task_s += now_s() - p->run_t0;
	pthread_mutex_unlock(&p->run_mu);
}

void mbs_run(mbs_pool_t *p, int64_t n, int64_t chunk, mbs_task_f task, mbs_batch_f batch, void *data, int64_t flush)
{
	if (n <= 0) return;
	pthread_mutex_lock(&p->run_mu);
	run_start(p, n, 0, chunk, task, batch, data, flush);
	run_end(p);
}

void mbs_open(mbs_pool_t *p, int64_t chunk, mbs_task_f task, mbs_batch_f batch, void *data, int64_t flush)
{
	pthread_mutex_lock(&p->run_mu);
	run_start(p, 0, 1, chunk, task, batch, data, flush);
}

int64_t mbs_add(mbs_pool_t *p, int64_t n)
{
	int k;
	int64_t first;
	if (n <= 0) return atomic_load(&p->n_tasks);
	first = atomic_fetch_add(&p->n_tasks, n);
	for (k = 0; k < p->n; ++k) {            /* under each thread's lock: a thread about to sleep sees the tasks */
		thr_t *t = &p->thr[k];
		pthread_mutex_lock(&t->mu);
		pthread_cond_signal(&t->cv);
		pthread_mutex_unlock(&t->mu);
	}
	return first;
}

void mbs_close(mbs_pool_t *p)
{
	atomic_store(&p->open_run, 0);
	poke(p);
	run_end(p);
}

void mbs_wait(void *req)
{
	co_t *co = tls_co;
	thr_t *t = tls_thr;
	mbs_pool_t *p;
	int64_t np;
	if (co == 0 || t == 0) die("mbs_wait outside a task");
	p = t->pool;
	if (p->batch == 0) die("a task waits but the run has no batch callback");
	pthread_mutex_lock(&t->mu);
	plist_reserve(&t->posted, t->posted.n + 1);
	t->posted.req[t->posted.n] = req, t->posted.co[t->posted.n] = co;
	np = ++t->posted.n;
	atomic_store_explicit(&t->n_posted, np, memory_order_relaxed);
	pthread_mutex_unlock(&t->mu);
	if (np == p->flush_thr) poke(p);   /* this thread's share of a full batch */
	++t->waits;
	mbs_switch(&co->sp, t->sched_sp);
	tls_co = co;
}

int mbs_in_task(void) { return tls_co != 0; }
int mbs_tid(void) { return tls_co ? tls_thr->tid : -1; }

void mbs_get_stats(const mbs_pool_t *p, mbs_stats_t *st)
{
	int i;
	memset(st, 0, sizeof(*st));
	st->batches = p->batches, st->coroutines_peak = p->co_peak;
	st->batch_s = p->batch_s, st->task_s = p->task_s;
	for (i = 0; i < p->n; ++i)
		st->waits += p->thr[i].waits, st->stacks += p->thr[i].n_co, st->busy_s += p->thr[i].busy_s, st->idle_s += p->thr[i].idle_s;
}
