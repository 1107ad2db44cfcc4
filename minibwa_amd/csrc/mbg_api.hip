// mbg_api.hip -- the C ABI of libmbg.so (declared in include/mbg.h).
//
// Host side of the library: device context, index residency (the reference's .mbw/.l2b files
// uploaded byte for byte), grow-only scratch pools and the batch entry points.  No torch, no
// allocation inside kernels; every HIP call is checked and mapped to an MBG_E* code.
#include <hip/hip_runtime.h>
#include <atomic>
#include <mutex>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include "mbg.h"
#include "mbg_internal.h"

using namespace mbg;

namespace {

thread_local std::string g_hip_err;

struct DBuf {
	void *p = nullptr;
	size_t n = 0;
};

} // namespace

// kinds of captured launch chains (an executable graph each per context)
enum { kGraphSeedPre, kGraphSeedPreSeg, kGraphSeedPost, kGraphStep, kGraphStepSeg, kGraphKinds };

struct mbg_ctx_s {
	int device = 0;
	int n_cu = 0;
	hipStream_t stream = nullptr;
	std::mutex mu;
	// resident index
	bool has_bwt = false, has_l2b = false;
	uint64_t *d_data = nullptr, *d_sa = nullptr, *d_sa_fine = nullptr;
	Sai *d_pre = nullptr;
	Idx ix{};
	uint64_t *d_pac = nullptr, *d_ambi = nullptr, *d_ctg_off = nullptr, *d_ctg_len = nullptr;
	std::vector<uint64_t> h_ctg_len;   // host copy of the contig lengths (argument checks)
	L2b lb{};
	uint64_t idx_bytes = 0;
	// scratch pools
	DBuf b_in[6], b_out[3], b_ctr, b_task, b_slab, b_gstate, b_defer;
	DBuf h_pinned;
	// seed -> locate -> extend step
	DBuf p_recs, p_rows, p_rowrec, p_first, p_best, p_jobs, p_meta, p_qb, p_tb, p_res, p_cig, p_ctr;
	DBuf s_slots1, s_slots2, s_cnt, s_off, s_spill, s_tmp, r_cnt, r_off, s_pack, s_qflag;
	// segmented pass 1 of long reads: segment counts / first tasks per read, segment tasks' reads,
	// long-read list, handoffs, chains, joins, flags, segment and fix-up slots, fallback list, counters
	DBuf g_nseg, g_base, g_rid, g_long, g_hand0, g_hand, g_chain, g_chn, g_conv, g_flag, g_slots, g_cnt, g_fix, g_fixcnt, g_fbl, g_ctr;
	// 512 unless MBG_SEED_SEG says otherwise (0: off).  On high-identity long reads the pass stays
	// bound by the longest exact stretch, which the lane holding it walks serially either way
	int32_t seg_len = 512, seg_slots = 16, seg_chain = 64, fix_slots = 16;
	int64_t cap_seg = 0;
	// launch shapes (MBG_GATHER_BPC, MBG_DP_WPC; the autotuner's knobs): 256-lane blocks per CU of the
	// seeding and SA kernels, and DP waves per CU; every such kernel takes its work from a queue, so
	// the shape changes only the speed
	int32_t gather_bpc = 8, dp_wpc = 32;
	DBuf l_buf[3];          // mb_align_pair batches: DP jobs, decision kinds, decided results
	DBuf m_buf[12];         // mate rescue
	DBuf a_buf[8];          // anchors
	DBuf c_buf[8];          // chains
	DBuf f_buf[20];         // the mapper's front pass
	DBuf h_front[10];       // its results in pinned host memory
	hipEvent_t fev[8] = {};
	// cumulative counters (mbg_counters)
	mbg_counters_t cnt{};
	bool count_work = false;
	DBuf f_stats;           // pass-1 work counters while count_work is on
	hipEvent_t dev_ev[2] = {};
	int64_t cap_task = 0, cap_spill = 0;
	int64_t cap_rec = 0, cap_rows = 0;
	hipEvent_t ev[8] = {};
	// ordering of calls on different streams (ctx_order / ctx_mark)
	hipEvent_t ord_ev = nullptr;
	hipStream_t ord_st = nullptr;
	bool ord_pending = false;
	// launch chains captured as graphs (run_graph): one executable graph per chain kind; off unless
	// MBG_GRAPH=1 or mbg_use_graphs (a capture was invalidated when another thread of the process
	// created a stream meanwhile; the mapper's other threads also allocate device memory while it runs)
	bool use_graphs = false;
	hipGraphExec_t graph_exec[kGraphKinds] = {};
	// long reads' segment lanes on a side stream beside pass 1 proper (MBG_SEED_SIDE=1 turns it on: it
	// shortens one front pass, but with several fronts and the DP already filling the device the
	// mapper ran slower with it)
	bool seg_side = false;
	hipStream_t side = nullptr;
	hipEvent_t fork_ev = nullptr, join_ev = nullptr;
};

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { g_hip_err = std::string(#x) + ": " + hipGetErrorString(e_); return MBG_EHIP; } } while (0)

// device bytes held by the grow-only work buffers of every context in the process (not the index)
static std::atomic<int64_t> g_work_bytes{0};

static int dgrow(DBuf &b, size_t bytes)
{
	if (bytes == 0) bytes = 16;
	if (b.n >= bytes) return MBG_OK;
	if (b.p) { hipFree(b.p); b.p = nullptr; g_work_bytes -= (int64_t)b.n; b.n = 0; }
	size_t n = bytes + bytes / 4 + 256;
	hipError_t e = hipMalloc(&b.p, n);
	if (e != hipSuccess) { g_hip_err = std::string("hipMalloc: ") + hipGetErrorString(e); b.p = nullptr; return MBG_ENOMEM; }
	b.n = n;
	g_work_bytes += (int64_t)n;
	return MBG_OK;
}

static void dfree(DBuf &b)
{
	if (b.p) hipFree(b.p), g_work_bytes -= (int64_t)b.n;
	b.p = nullptr, b.n = 0;
}

// grow-only pinned host buffer
static int hgrow(DBuf &b, size_t bytes)
{
	if (bytes == 0) bytes = 16;
	if (b.n >= bytes) return MBG_OK;
	if (b.p) { hipHostFree(b.p); b.p = nullptr; b.n = 0; }
	size_t n = bytes + bytes / 4 + 256;
	hipError_t e = hipHostMalloc(&b.p, n, 0);
	if (e != hipSuccess) { g_hip_err = std::string("hipHostMalloc: ") + hipGetErrorString(e); b.p = nullptr; return MBG_ENOMEM; }
	b.n = n;
	return MBG_OK;
}

#define TRY(x) do { int r_ = (x); if (r_ != MBG_OK) return r_; } while (0)

// Every entry point of a context uses the context's scratch (counters, slabs, work buffers).  A call
// that returns with its work still running on a caller's stream records ord_ev there (ctx_mark); the
// next call, if it runs on another stream (the context's own included), makes that stream wait for the
// event first (ctx_order), so that calls from threads on different streams never overlap on the
// scratch.  Calls on one stream are ordered by the stream itself.
static int ctx_order(mbg_ctx_t *c, hipStream_t st)
{
	if (c->ord_pending && c->ord_st != st) HIPCHK(hipStreamWaitEvent(st, c->ord_ev, 0));
	return MBG_OK;
}

static int ctx_mark(mbg_ctx_t *c, hipStream_t st)
{
	if (c->ord_ev == nullptr) HIPCHK(hipEventCreateWithFlags(&c->ord_ev, hipEventDisableTiming));
	HIPCHK(hipEventRecord(c->ord_ev, st));
	c->ord_st = st, c->ord_pending = true;
	return MBG_OK;
}

// the device current and the context's stream ordered after pending work on another stream
static int ctx_enter(mbg_ctx_t *c)
{
	HIPCHK(hipSetDevice(c->device));
	return ctx_order(c, c->stream);
}

// ---- the allocation-free SA locate (mbg_sa_batch_dev_noalloc) ----------------------------------
// sa_kernel's only scratch is its work-queue counter.  For callers that cannot allocate, it lives in
// this file-scope device array: one copy per GPU, held until the process ends, 256 bytes.  Host code
// never passes the array's name to a kernel (on the host it is only a stand-in for the symbol): each
// device's address comes from hipGetSymbolAddress, once per device, at the first mbg_init on it.
static constexpr int kNoAllocWords = 64;
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];
static constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static uint32_t *g_noalloc_ptr[kMaxDevices];

// with `device` current: resolve this device's copy of the scratch array
static int noalloc_setup(int device)
{
	if (device < 0 || device >= kMaxDevices) return MBG_EARG;
	std::lock_guard<std::mutex> lk(g_noalloc_mu);
	if (g_noalloc_ptr[device]) return MBG_OK;
	void *p = nullptr;
	hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	if (e != hipSuccess || p == nullptr) {
		g_hip_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
		return MBG_EHIP;
	}
	g_noalloc_ptr[device] = (uint32_t*)p;
	return MBG_OK;
}

// zeroes n words; a kernel rather than hipMemsetAsync, so that the call stays valid under graph capture
__global__ void noalloc_clear_kernel(uint32_t *p, int n)
{
	const int i = threadIdx.x;
	if (i < n) p[i] = 0;
}

extern "C" {

const char *mbg_strerror(int err)
{
	switch (err) {
	case MBG_OK: return "ok";
	case MBG_EARG: return "invalid argument";
	case MBG_EHIP: return "HIP runtime error";
	case MBG_ENOMEM: return "out of memory";
	case MBG_EIO: return "cannot read index file";
	case MBG_EFORMAT: return "bad index file format";
	case MBG_ENOIDX: return "no index resident in the context";
	case MBG_EUNSUPPORTED: return "job outside the supported range";
	case MBG_EOVERFLOW: return "output capacity exceeded";
	default: return "unknown error";
	}
}

const char *mbg_last_hip_error(void) { return g_hip_err.c_str(); }

int mbg_init(int device, mbg_ctx_t **out)
{
	if (out == nullptr) return MBG_EARG;
	*out = nullptr;
	int n_dev = 0;
	HIPCHK(hipGetDeviceCount(&n_dev));
	if (device < 0 || device >= n_dev) return MBG_EARG;
	HIPCHK(hipSetDevice(device));
	{   // how host threads wait for the device: MBG_HOST_WAIT=block sleeps in the driver (frees the core
	    // for other host work), spin / yield busy-wait; unset keeps HIP's automatic choice
		const char *w = getenv("MBG_HOST_WAIT");
		unsigned fl = 0;
		if (w && strcmp(w, "block") == 0) fl = hipDeviceScheduleBlockingSync;
		else if (w && strcmp(w, "spin") == 0) fl = hipDeviceScheduleSpin;
		else if (w && strcmp(w, "yield") == 0) fl = hipDeviceScheduleYield;
		if (fl && hipSetDeviceFlags(fl) != hipSuccess) (void)hipGetLastError();   // refused once the device is active: the default stays
	}
	mbg_ctx_t *c = new mbg_ctx_t();
	c->device = device;
	{   // MBG_SEED_SEG=len[,segment slots[,chain cap[,fix-up slots]]]: pass 1 splits reads of at least 2*len
	    // bases into segments searched in parallel (0: off); the capacities only decide how often a
	    // read falls back to a serial search, never the result
		const char *e = getenv("MBG_SEED_SEG");
		if (e) {
			int v[4] = {c->seg_len, c->seg_slots, c->seg_chain, c->fix_slots};
			sscanf(e, "%d,%d,%d,%d", &v[0], &v[1], &v[2], &v[3]);
			c->seg_len = v[0] >= 32 ? v[0] : 0;
			c->seg_slots = v[1] > 0 ? v[1] : 1, c->seg_chain = v[2] > 0 ? v[2] : 1, c->fix_slots = v[3] > 0 ? v[3] : 1;
		}
	}
	{
		const char *g = getenv("MBG_GATHER_BPC"), *d = getenv("MBG_DP_WPC");
		if (g && atoi(g) >= 1 && atoi(g) <= 64) c->gather_bpc = atoi(g);
		if (d && atoi(d) >= 1 && atoi(d) <= 128) c->dp_wpc = atoi(d);
		const char *gr = getenv("MBG_GRAPH");
		if (gr) c->use_graphs = atoi(gr) != 0;
		const char *sd = getenv("MBG_SEED_SIDE");
		if (sd) c->seg_side = atoi(sd) != 0;
	}
	hipDeviceProp_t prop;
	hipError_t e = hipGetDeviceProperties(&prop, device);
	if (e == hipSuccess) c->n_cu = prop.multiProcessorCount;
	if (c->n_cu <= 0) c->n_cu = 256;
	e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
	if (e != hipSuccess) { g_hip_err = hipGetErrorString(e); delete c; return MBG_EHIP; }
	if (c->seg_side) {   // created here, not at first use: no stream is created while another context captures
		if ((e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking)) == hipSuccess &&
		    (e = hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming)) == hipSuccess)
			e = hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming);
		if (e != hipSuccess) { g_hip_err = hipGetErrorString(e); mbg_destroy(c); return MBG_EHIP; }
	}
	int rc = noalloc_setup(device);
	if (rc != MBG_OK) { mbg_destroy(c); return rc; }
	*out = c;
	return MBG_OK;
}

static void free_bwt(mbg_ctx_t *c)
{
	if (c->d_data) hipFree(c->d_data);
	if (c->d_sa) hipFree(c->d_sa);
	if (c->d_sa_fine) hipFree(c->d_sa_fine);
	if (c->d_pre) hipFree(c->d_pre);
	c->d_data = c->d_sa = c->d_sa_fine = nullptr, c->d_pre = nullptr;
	c->has_bwt = false;
}

static void free_l2b(mbg_ctx_t *c)
{
	if (c->d_pac) hipFree(c->d_pac);
	if (c->d_ambi) hipFree(c->d_ambi);
	if (c->d_ctg_off) hipFree(c->d_ctg_off);
	if (c->d_ctg_len) hipFree(c->d_ctg_len);
	c->d_pac = c->d_ambi = c->d_ctg_off = c->d_ctg_len = nullptr;
	c->has_l2b = false;
}

void mbg_destroy(mbg_ctx_t *c)
{
	if (c == nullptr) return;
	hipSetDevice(c->device);
	if (c->ord_pending) hipEventSynchronize(c->ord_ev);   // work left on a caller's stream
	if (c->stream) hipStreamSynchronize(c->stream);
	if (c->ord_ev) hipEventDestroy(c->ord_ev);
	free_bwt(c);
	free_l2b(c);
	for (auto &b : c->b_in) dfree(b);
	for (auto &b : c->b_out) dfree(b);
	dfree(c->b_ctr); dfree(c->b_task); dfree(c->b_slab); dfree(c->b_gstate); dfree(c->b_defer);
	for (DBuf *b : {&c->p_recs, &c->p_rows, &c->p_rowrec, &c->p_first, &c->p_best, &c->p_jobs, &c->p_meta, &c->p_qb, &c->p_tb,
	                &c->p_res, &c->p_cig, &c->p_ctr, &c->s_slots1, &c->s_slots2, &c->s_cnt, &c->s_off, &c->s_spill, &c->s_tmp,
	                &c->r_cnt, &c->r_off, &c->s_pack, &c->s_qflag, &c->g_nseg, &c->g_base, &c->g_rid, &c->g_long, &c->g_hand0,
	                &c->g_hand, &c->g_chain, &c->g_chn, &c->g_conv, &c->g_flag, &c->g_slots, &c->g_cnt, &c->g_fix, &c->g_fixcnt,
	                &c->g_fbl, &c->g_ctr}) dfree(*b);
	for (auto &b : c->l_buf) dfree(b);
	for (auto &b : c->m_buf) dfree(b);
	for (auto &b : c->a_buf) dfree(b);
	for (auto &b : c->c_buf) dfree(b);
	for (auto &b : c->f_buf) dfree(b);
	for (auto &b : c->h_front) if (b.p) hipHostFree(b.p);
	for (hipEvent_t &e : c->ev) if (e) hipEventDestroy(e);
	for (hipEvent_t &e : c->fev) if (e) hipEventDestroy(e);
	for (hipEvent_t &e : c->dev_ev) if (e) hipEventDestroy(e);
	dfree(c->f_stats);
	if (c->h_pinned.p) hipHostFree(c->h_pinned.p);
	for (hipGraphExec_t &g : c->graph_exec) if (g) hipGraphExecDestroy(g);
	if (c->side) hipStreamSynchronize(c->side), hipStreamDestroy(c->side);
	if (c->fork_ev) hipEventDestroy(c->fork_ev);
	if (c->join_ev) hipEventDestroy(c->join_ev);
	if (c->stream) hipStreamDestroy(c->stream);
	delete c;
}

static uint64_t bwt_data_len(uint64_t len) /* bwt.c:52-58 */
{
	const uint64_t nb = (len + 127) / 128;
	return nb * 4 + (nb + 1) * 4;
}

static int upload(void **dst, const void *src, size_t bytes, hipStream_t st)
{
	hipError_t e = hipMalloc(dst, bytes ? bytes : 16);
	if (e != hipSuccess) { g_hip_err = hipGetErrorString(e); *dst = nullptr; return MBG_ENOMEM; }
	if (bytes) {
		e = hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, st);
		if (e != hipSuccess) { g_hip_err = hipGetErrorString(e); return MBG_EHIP; }
	}
	return MBG_OK;
}

static const uint32_t kFineBit = 2;     // rows between entries of the finer SA grid: 4

static int bwt_upload_locked(mbg_ctx_t *c, const mbg_bwt_desc_t *d, int32_t pre_len)
{
	if (d == nullptr || d->data == nullptr || pre_len < 0 || pre_len > 16) return MBG_EARG;
	if (d->L2[4] != d->seq_len || d->primary > d->seq_len) return MBG_EFORMAT;
	if (d->sa_bit > 30 || d->n_sa != ((d->seq_len + (1ULL << d->sa_bit)) >> d->sa_bit) || d->sa == nullptr) return MBG_EFORMAT;
	free_bwt(c);
	const uint64_t dl = bwt_data_len(d->seq_len);
	TRY(upload((void**)&c->d_data, d->data, dl * 8, c->stream));
	TRY(upload((void**)&c->d_sa, d->sa, d->n_sa * 8, c->stream));
	Idx &ix = c->ix;
	ix.data = c->d_data; ix.data_len = dl;
	ix.sa = c->d_sa; ix.sa_bit = d->sa_bit;
	ix.primary = d->primary; ix.seq_len = d->seq_len;
	for (int i = 0; i < 5; ++i) ix.L2[i] = d->L2[i];
	ix.pre = nullptr; ix.pre_len = 0;
	ix.sa_fine = nullptr; ix.fine_bit = 0;
	c->idx_bytes = dl * 8 + d->n_sa * 8;
	// SA of every 4th row (rows/4 x 8 bytes), so that a locate walks about 3 LF steps instead of
	// about 2^sa_bit - 1; skipped when the device cannot hold it
	if (d->sa_bit > kFineBit) {
		const uint64_t n_fine = (d->seq_len >> kFineBit) + 1;
		if (hipMalloc((void**)&c->d_sa_fine, n_fine * 8) == hipSuccess) {
			ix.fine_bit = kFineBit;
			HIPCHK(launch_sa_fine(ix, c->d_sa_fine, n_fine, c->stream));
			ix.sa_fine = c->d_sa_fine;
			c->idx_bytes += n_fine * 8;
		} else {
			(void)hipGetLastError();
			c->d_sa_fine = nullptr;
		}
	}
	if (pre_len > 0) {
		const size_t n = (size_t)1 << (2 * pre_len);
		hipError_t e = hipMalloc((void**)&c->d_pre, n * sizeof(Sai));
		if (e != hipSuccess) { g_hip_err = hipGetErrorString(e); return MBG_ENOMEM; }
		HIPCHK(launch_kmer_cache(ix, c->d_pre, pre_len, c->stream));
		ix.pre = c->d_pre; ix.pre_len = pre_len;
		c->idx_bytes += n * sizeof(Sai);
	}
	HIPCHK(hipStreamSynchronize(c->stream));
	c->has_bwt = true;
	return MBG_OK;
}

int mbg_bwt_upload(mbg_ctx_t *c, const mbg_bwt_desc_t *d, int32_t pre_len)
{
	if (c == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	TRY(ctx_enter(c));
	return bwt_upload_locked(c, d, pre_len);
}

static int l2b_upload_locked(mbg_ctx_t *c, const mbg_l2b_desc_t *d)
{
	if (d == nullptr || (d->n_ctg && d->ctg_len == nullptr) || (d->n_pac && d->pac == nullptr)) return MBG_EARG;
	std::vector<uint64_t> off(d->n_ctg + 1, 0);
	for (uint64_t i = 0; i < d->n_ctg; ++i) off[i + 1] = off[i] + d->ctg_len[i];
	if (off[d->n_ctg] != d->tot_len || d->n_pac * 32 < d->tot_len) return MBG_EFORMAT;
	free_l2b(c);
	TRY(upload((void**)&c->d_pac, d->pac, d->n_pac * 8, c->stream));
	TRY(upload((void**)&c->d_ambi, d->ambi, d->n_ambi * 16, c->stream));
	TRY(upload((void**)&c->d_ctg_off, off.data(), d->n_ctg * 8, c->stream));
	TRY(upload((void**)&c->d_ctg_len, d->ctg_len, d->n_ctg * 8, c->stream));
	HIPCHK(hipStreamSynchronize(c->stream));
	L2b &lb = c->lb;
	lb.pac = c->d_pac; lb.ambi = c->d_ambi; lb.ctg_off = c->d_ctg_off; lb.ctg_len = c->d_ctg_len;
	lb.n_ctg = d->n_ctg; lb.n_ambi = d->n_ambi; lb.tot_len = d->tot_len;
	c->h_ctg_len.assign(d->ctg_len, d->ctg_len + d->n_ctg);
	c->idx_bytes += d->n_pac * 8 + d->n_ambi * 16 + d->n_ctg * 16;
	c->has_l2b = true;
	return MBG_OK;
}

int mbg_l2b_upload(mbg_ctx_t *c, const mbg_l2b_desc_t *d)
{
	if (c == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	TRY(ctx_enter(c));
	return l2b_upload_locked(c, d);
}

static bool rd(FILE *fp, void *p, size_t sz, size_t n) { return n == 0 || fread(p, sz, n, fp) == n; }

int mbg_idx_load(mbg_ctx_t *c, const char *prefix, int32_t pre_len)
{
	if (c == nullptr || prefix == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	TRY(ctx_enter(c));
	std::string fn = std::string(prefix) + ".mbw";
	FILE *fp = fopen(fn.c_str(), "rb");
	if (fp == nullptr) return MBG_EIO;
	// .mbw: magic, sa_bit, primary, L2[1..4], data, n_sa, sa (bwt.c:631-646)
	char magic[4];
	uint32_t sa_bit = 0;
	uint64_t hdr[5], n_sa = 0;
	if (!rd(fp, magic, 1, 4) || memcmp(magic, "MBW\2", 4) != 0 || !rd(fp, &sa_bit, 4, 1) || !rd(fp, hdr, 8, 5)) { fclose(fp); return MBG_EFORMAT; }
	mbg_bwt_desc_t d;
	d.primary = hdr[0];
	d.L2[0] = 0;
	memcpy(&d.L2[1], &hdr[1], 32);
	d.seq_len = d.L2[4];
	d.sa_bit = sa_bit;
	const uint64_t dl = bwt_data_len(d.seq_len);
	std::vector<uint64_t> data(dl), sa;
	if (!rd(fp, data.data(), 8, dl) || !rd(fp, &n_sa, 8, 1) || sa_bit > 30 || n_sa != ((d.seq_len + (1ULL << sa_bit)) >> sa_bit)) { fclose(fp); return MBG_EFORMAT; }
	sa.resize(n_sa);
	if (!rd(fp, sa.data(), 8, n_sa)) { fclose(fp); return MBG_EFORMAT; }
	fclose(fp);
	d.data = data.data(); d.n_sa = n_sa; d.sa = sa.data();
	TRY(bwt_upload_locked(c, &d, pre_len));
	// .l2b (l2bit.c:243-330)
	fn = std::string(prefix) + ".l2b";
	fp = fopen(fn.c_str(), "rb");
	if (fp == nullptr) return MBG_EIO;
	uint32_t dummy;
	uint64_t h[7];
	if (!rd(fp, magic, 1, 4) || memcmp(magic, "L2B\1", 4) != 0 || !rd(fp, &dummy, 4, 1) || !rd(fp, h, 8, 7)) { fclose(fp); return MBG_EFORMAT; }
	mbg_l2b_desc_t l;
	l.n_ctg = h[0]; l.tot_len = h[1]; l.n_ambi = h[2];
	const uint64_t n_mask = h[3];
	l.n_pac = h[6];
	std::vector<uint64_t> clen(l.n_ctg), ambi(l.n_ambi * 2), mask(n_mask * 2), pac(l.n_pac);
	if (!rd(fp, clen.data(), 8, l.n_ctg) || !rd(fp, ambi.data(), 16, l.n_ambi) || !rd(fp, mask.data(), 16, n_mask) || !rd(fp, pac.data(), 8, l.n_pac)) { fclose(fp); return MBG_EFORMAT; }
	fclose(fp);
	l.ctg_len = clen.data(); l.ambi = ambi.data(); l.pac = pac.data();
	return l2b_upload_locked(c, &l);
}

int mbg_idx_info(mbg_ctx_t *c, mbg_idx_info_t *info)
{
	if (c == nullptr || info == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	info->seq_len = c->ix.seq_len; info->primary = c->ix.primary;
	info->tot_len = c->has_l2b ? c->lb.tot_len : 0; info->n_ctg = c->has_l2b ? c->lb.n_ctg : 0;
	info->pre_len = c->ix.pre_len; info->sa_bit = (int32_t)c->ix.sa_bit;
	info->device_bytes = c->idx_bytes;
	return MBG_OK;
}

int mbg_idx_attach(mbg_ctx_t *c, mbg_ctx_t *src)
{
	if (c == nullptr || src == nullptr || c == src) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	Idx ix;
	L2b lb;
	bool has_bwt, has_l2b;
	std::vector<uint64_t> clen;
	{
		std::lock_guard<std::mutex> ls(src->mu);
		if (!src->has_bwt) return MBG_ENOIDX;
		if (src->device != c->device) return MBG_EARG;
		ix = src->ix, lb = src->lb, has_bwt = src->has_bwt, has_l2b = src->has_l2b, clen = src->h_ctg_len;
	}
	TRY(ctx_enter(c));
	HIPCHK(hipStreamSynchronize(c->stream));
	free_bwt(c);
	free_l2b(c);
	// the device arrays stay owned by src (c->d_* remain null, so c never frees them)
	c->ix = ix, c->lb = lb, c->has_bwt = has_bwt, c->has_l2b = has_l2b;
	c->h_ctg_len.swap(clen);
	c->idx_bytes = 0;
	return MBG_OK;
}

int mbg_pre_copy(mbg_ctx_t *c, mbg_sai_t *out, uint64_t n)
{
	if (c == nullptr || out == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt || c->ix.pre_len == 0) return MBG_ENOIDX;
	if (n != (1ULL << (2 * c->ix.pre_len))) return MBG_EARG;
	TRY(ctx_enter(c));
	HIPCHK(hipMemcpy(out, c->d_pre, n * sizeof(Sai), hipMemcpyDeviceToHost));
	return MBG_OK;
}

int mbg_rank2a(mbg_ctx_t *c, int64_t n, const uint64_t *k, const uint64_t *l, uint64_t *ck, uint64_t *cl)
{
	if (c == nullptr || n < 0 || (n > 0 && (!k || !l || !ck || !cl))) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	if (n == 0) return MBG_OK;
	TRY(ctx_enter(c));
	for (int64_t i = 0; i < n; ++i)
		if (k[i] > l[i] || l[i] > c->ix.seq_len + 1) return MBG_EARG;
	TRY(dgrow(c->b_in[0], n * 8)); TRY(dgrow(c->b_in[1], n * 8)); TRY(dgrow(c->b_out[0], n * 32)); TRY(dgrow(c->b_out[1], n * 32));
	HIPCHK(hipMemcpyAsync(c->b_in[0].p, k, n * 8, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[1].p, l, n * 8, hipMemcpyHostToDevice, c->stream));
	HIPCHK(launch_rank2a(c->ix, n, (uint64_t*)c->b_in[0].p, (uint64_t*)c->b_in[1].p, (uint64_t*)c->b_out[0].p, (uint64_t*)c->b_out[1].p, c->stream));
	HIPCHK(hipMemcpyAsync(ck, c->b_out[0].p, n * 32, hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipMemcpyAsync(cl, c->b_out[1].p, n * 32, hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipStreamSynchronize(c->stream));
	return MBG_OK;
}

int mbg_extend(mbg_ctx_t *c, int64_t n, const mbg_sai_t *ik, const uint8_t *is_back, mbg_sai_t *ok)
{
	if (c == nullptr || n < 0 || (n > 0 && (!ik || !is_back || !ok))) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	if (n == 0) return MBG_OK;
	TRY(ctx_enter(c));
	for (int64_t i = 0; i < n; ++i) {
		const uint64_t xf = is_back[i] ? ik[i].x[0] : ik[i].x[1];
		if (xf > c->ix.seq_len + 1 || ik[i].size > c->ix.seq_len + 1 - xf) return MBG_EARG;
	}
	TRY(dgrow(c->b_in[0], n * 32)); TRY(dgrow(c->b_in[1], n)); TRY(dgrow(c->b_out[0], n * 128));
	HIPCHK(hipMemcpyAsync(c->b_in[0].p, ik, n * 32, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[1].p, is_back, n, hipMemcpyHostToDevice, c->stream));
	HIPCHK(launch_extend(c->ix, n, (const Sai*)c->b_in[0].p, (const uint8_t*)c->b_in[1].p, (Sai*)c->b_out[0].p, c->stream));
	HIPCHK(hipMemcpyAsync(ok, c->b_out[0].p, n * 128, hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipStreamSynchronize(c->stream));
	return MBG_OK;
}

// the grid of the seeding and SA kernels: n_items lanes of work in 256-lane blocks, at most
// gather_bpc blocks per CU (their queues hand out the rest)
static unsigned gather_grid(mbg_ctx_t *c, int64_t n_items)
{
	int64_t want = (n_items + 255) / 256;
	const int64_t cap = (int64_t)c->n_cu * c->gather_bpc;
	if (want > cap) want = cap;
	if (want < 1) want = 1;
	return (unsigned)want;
}

static unsigned persistent_grid(mbg_ctx_t *c, int64_t n_items, int per_block)
{
	int64_t want = (n_items + per_block - 1) / per_block;
	int64_t cap = (int64_t)c->n_cu * 8;          // 8 blocks of 256 lanes per CU keep the memory pipes busy
	if (want > cap) want = cap;
	if (want < 1) want = 1;
	return (unsigned)want;
}

static const int32_t kSlot1 = 16, kSlot2 = 8;

// ---- launch chains as HIP graphs ----------------------------------------------------------------
//
// A chain of launches with no allocation and no host synchronisation inside (a seeding pass; the
// whole seed -> locate -> extend step) is captured from the stream into a graph and launched as one
// unit.  Every call captures its chain again, since pointers and sizes change from call to call, and
// updates the context's executable graph of that kind in place (hipGraphExecUpdate), which is cheap
// when the chain has the same shape as the last one; a chain of another shape instantiates anew.
// Counters and flags are cleared by a kernel (zero_async), never by a memset node.  Capture is off by
// default (MBG_GRAPH=1 or mbg_use_graphs turns it on): on this runtime a capture was invalidated when
// another thread created a stream while it was open, and the graphs saved no time at any batch size
// measured.  Results are the same either way.

__global__ void zero_words_kernel(uint32_t *p, int64_t n)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

// zeroes `bytes` (a multiple of 4, 4-aligned) on the stream, as a kernel
static hipError_t zero_async(void *p, size_t bytes, hipStream_t st)
{
	const int64_t n = (int64_t)(bytes / 4);
	if (n <= 0) return hipSuccess;
	int64_t g = (n + 255) / 256;
	if (g > 2048) g = 2048;
	hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)g), dim3(256), 0, st, (uint32_t*)p, n);
	return hipGetLastError();
}

extern "C++" {
template <class F>
static int run_graph(mbg_ctx_t *c, hipStream_t st, int kind, F &&enqueue)
{
	if (!c->use_graphs) {
		++c->cnt.direct_chains;
		return enqueue(st);
	}
	HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
	const int rc = enqueue(st);
	hipGraph_t g = nullptr;
	const hipError_t ec = hipStreamEndCapture(st, &g);
	if (rc != MBG_OK) { if (g) hipGraphDestroy(g); return rc; }
	if (ec != hipSuccess) { g_hip_err = std::string("hipStreamEndCapture: ") + hipGetErrorString(ec); return MBG_EHIP; }
	hipGraphExec_t &ex = c->graph_exec[kind];
	if (ex) {
		hipGraphExecUpdateResult ur;
		hipGraphNode_t bad = nullptr;
		if (hipGraphExecUpdate(ex, g, &bad, &ur) != hipSuccess || ur != hipGraphExecUpdateSuccess) {
			(void)hipGetLastError();
			hipGraphExecDestroy(ex);
			ex = nullptr;
			++c->cnt.graph_instantiations;
		}
	} else ++c->cnt.graph_instantiations;
	hipError_t e = hipSuccess;
	if (!ex) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
	hipGraphDestroy(g);
	if (e != hipSuccess) { ex = nullptr; g_hip_err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e); return MBG_EHIP; }
	HIPCHK(hipGraphLaunch(ex, st));
	++c->cnt.graph_launches;
	return MBG_OK;
}
} // extern "C++"

// Segmented pass 1 (reads of at least 2*seg_len bases): the serial SMEM search is a chain of starts
// x0 = 0 -> x1 -> ..., one step each, a step emitting at most one SMEM (mb_bwt_smem, bwt.c:280-305).
// Segment k's lane starts at k*seg_len and records its chain up to its first value past the segment
// (its handoff); segment 0 is the read's own pass-1 task and is exact.  The fix-up walks the true
// chain from segment 0's handoff: where it meets a segment's chain, that lane's steps from there on are
// the true ones and the walk jumps to its handoff; where it does not, the fix-up takes the true steps
// itself.  Records of a segment before its join are dropped.  A segment or fix-up task that runs out
// of room flags its read, which then runs serially from segment 0's handoff.  The result is the serial
// search's set of SMEMs exactly, whatever the segment length and capacities.
// pools of the segmented pass (grown before any launch, so that the launches can be captured)
static int seg_alloc(mbg_ctx_t *c, int32_t n_reads, hipStream_t st)
{
	if (c->cap_seg < 1024 + n_reads / 16) c->cap_seg = 1024 + n_reads / 16;
	const int64_t cs = c->cap_seg, nr = n_reads;
	TRY(dgrow(c->g_nseg, (size_t)(nr + 1) * 4)); TRY(dgrow(c->g_base, (size_t)(nr + 1) * 4));
	TRY(dgrow(c->g_hand0, (size_t)nr * 4)); TRY(dgrow(c->g_flag, (size_t)nr * 4));
	TRY(dgrow(c->g_rid, (size_t)cs * 4)); TRY(dgrow(c->g_long, (size_t)nr * 4)); TRY(dgrow(c->g_hand, (size_t)cs * 4));
	TRY(dgrow(c->g_chain, (size_t)cs * c->seg_chain * 4)); TRY(dgrow(c->g_chn, (size_t)cs * 4)); TRY(dgrow(c->g_conv, (size_t)cs * 4));
	TRY(dgrow(c->g_slots, (size_t)cs * c->seg_slots * sizeof(SeedRec))); TRY(dgrow(c->g_cnt, (size_t)cs * 4));
	TRY(dgrow(c->g_fix, (size_t)cs * c->fix_slots * sizeof(SeedRec))); TRY(dgrow(c->g_fixcnt, (size_t)cs * 4));
	TRY(dgrow(c->g_fbl, (size_t)nr * 4)); TRY(dgrow(c->g_ctr, 64));
	size_t tb = 0;
	HIPCHK(scan_u32(nullptr, &tb, nullptr, nullptr, nr + 1, st));
	TRY(dgrow(c->s_tmp, tb > c->s_tmp.n ? tb : c->s_tmp.n));
	return MBG_OK;
}

static int seg_setup(mbg_ctx_t *c, hipStream_t st, int32_t n_reads, const int32_t *len, SeedArgs &a)
{
	const int64_t cs = c->cap_seg, nr = n_reads;
	uint32_t *nseg = (uint32_t*)c->g_nseg.p, *base = (uint32_t*)c->g_base.p, *gc = (uint32_t*)c->g_ctr.p;
	HIPCHK(zero_async(gc, 64, st));
	HIPCHK(zero_async(c->g_flag.p, (size_t)nr * 4, st));
	HIPCHK(zero_async(nseg + nr, 4, st));
	HIPCHK(launch_seg_count(len, n_reads, c->seg_len, nseg, st));
	size_t tb = c->s_tmp.n;
	HIPCHK(scan_u32(c->s_tmp.p, &tb, nseg, base, nr + 1, st));
	// the segment total (base[n_reads]) where seed_overflowed reads it
	HIPCHK(hipMemcpyAsync(gc + 5, base + nr, 4, hipMemcpyDeviceToDevice, st));
	HIPCHK(launch_seg_build(len, n_reads, c->seg_len, base, cs, (uint32_t*)c->g_rid.p, (uint32_t*)c->g_long.p, gc + 3,
	                        (uint32_t*)c->g_flag.p, st));
	a.seg_len = c->seg_len;
	a.segbase = base; a.seg_rid = (const uint32_t*)c->g_rid.p;
	a.hand0 = (uint32_t*)c->g_hand0.p; a.hand = (uint32_t*)c->g_hand.p;
	a.chain = (uint32_t*)c->g_chain.p; a.chain_n = (uint32_t*)c->g_chn.p; a.chain_cap = c->seg_chain;
	a.conv = (uint32_t*)c->g_conv.p; a.rflag = (uint32_t*)c->g_flag.p;
	return MBG_OK;
}

// callers that hold the read lengths on the host size the segment pools for them up front (otherwise
// the pools grow after a batch that outran them, whose reads then ran serially)
static void seg_hint(mbg_ctx_t *c, const int32_t *len_host, int32_t n)
{
	if (c->seg_len <= 0 || len_host == nullptr) return;
	int64_t need = 0;
	for (int32_t i = 0; i < n; ++i)
		if (len_host[i] >= 2 * c->seg_len) need += len_host[i] / c->seg_len - 1;
	if (need > 0 && need + 64 > c->cap_seg) c->cap_seg = need + 64;
}

// the segment launch, the fix-up, the fallback and the merge (after pass 1 over the reads)
static int seg_segments(mbg_ctx_t *c, hipStream_t st, int32_t n_reads, const SeedArgs &a)
{
	const int64_t cs = c->cap_seg;
	uint32_t *gc = (uint32_t*)c->g_ctr.p;
	SeedArgs s3 = a;                       // segments k >= 1
	s3.next = gc; s3.n_tasks = 0xffffffffu; s3.n_tasks_dev = (const uint32_t*)c->g_base.p + n_reads; s3.task_cap = cs;
	s3.slots = (SeedRec*)c->g_slots.p; s3.slot_cnt = (uint32_t*)c->g_cnt.p; s3.kslot = c->seg_slots;
	HIPCHK(launch_smem(c->ix, s3, 3, gather_grid(c, cs < n_reads ? cs : n_reads), st));
	return MBG_OK;
}

// segments == false: the segment launch already ran (seg_segments, on a side stream)
static int seg_run(mbg_ctx_t *c, hipStream_t st, int32_t n_reads, const int32_t *len, const SeedArgs &a, bool segments = true)
{
	const int64_t cs = c->cap_seg;
	uint32_t *gc = (uint32_t*)c->g_ctr.p;
	const unsigned grid = gather_grid(c, cs < n_reads ? cs : n_reads);
	if (segments) TRY(seg_segments(c, st, n_reads, a));
	SeedArgs s4 = a;                       // the fix-up of each long read
	s4.next = gc + 1; s4.n_tasks = 0xffffffffu; s4.n_tasks_dev = gc + 3; s4.task_cap = cs; s4.list = (const uint32_t*)c->g_long.p;
	s4.slots = (SeedRec*)c->g_fix.p; s4.slot_cnt = (uint32_t*)c->g_fixcnt.p; s4.kslot = c->fix_slots;
	HIPCHK(launch_smem(c->ix, s4, 4, grid, st));
	HIPCHK(launch_seg_flagged((const uint32_t*)c->g_long.p, gc + 3, (const uint32_t*)c->g_flag.p, (uint32_t*)c->g_fbl.p, gc + 4, n_reads,
	                          st));
	SeedArgs s5 = a;                       // reads whose segment or fix-up tasks overflowed
	s5.next = gc + 2; s5.n_tasks = 0xffffffffu; s5.n_tasks_dev = gc + 4; s5.task_cap = n_reads; s5.list = (const uint32_t*)c->g_fbl.p;
	s5.slots = nullptr; s5.slot_cnt = nullptr; s5.kslot = 0;
	HIPCHK(launch_smem(c->ix, s5, 5, grid, st));
	HIPCHK(launch_seg_merge(len, c->seg_len, (const uint32_t*)c->g_long.p, gc + 3, (const uint32_t*)c->g_flag.p,
	                        (const uint32_t*)c->g_base.p, (const SeedRec*)c->g_slots.p, (const uint32_t*)c->g_cnt.p, c->seg_slots,
	                        (const uint32_t*)c->g_conv.p, (const SeedRec*)c->g_fix.p, (const uint32_t*)c->g_fixcnt.p, c->fix_slots,
	                        a.spill, a.n_spill, a.spill_cap, cs, st));
	return MBG_OK;
}

// every pool of a seeding pass, grown before its launches.  Records and pass-2 tasks are sized from the
// reads and, when the caller knows them (bases >= 0), from their bases too: long reads give about two
// SMEMs per kilobase, so a pool sized by reads alone overflows on a context's first long-read batch and
// the pass runs again.
static int seed_alloc(mbg_ctx_t *c, hipStream_t st, int32_t n_reads, int64_t bases = -1)
{
	const int64_t by_bases = bases > 0 ? bases / 256 : 0;
	if (c->cap_rec < 8 * (int64_t)n_reads + by_bases + 1024) c->cap_rec = 8 * (int64_t)n_reads + by_bases + 1024;
	if (c->cap_task < 2 * (int64_t)n_reads + by_bases + 1024) c->cap_task = 2 * (int64_t)n_reads + by_bases + 1024;
	// a long read's records beyond its pass-1 slots spill: the spill list grows with the bases too
	if (c->cap_spill < (int64_t)n_reads + by_bases + 4096) c->cap_spill = (int64_t)n_reads + by_bases + 4096;
	const int64_t rc = c->cap_rec, tc = c->cap_task, sc = c->cap_spill, n_cnt = (int64_t)n_reads + tc;
	TRY(dgrow(c->b_ctr, 128));
	TRY(dgrow(c->p_recs, rc * sizeof(SeedRec)));
	TRY(dgrow(c->s_slots1, (size_t)n_reads * kSlot1 * sizeof(SeedRec) + 64));
	TRY(dgrow(c->s_slots2, (size_t)tc * kSlot2 * sizeof(SeedRec)));
	TRY(dgrow(c->s_cnt, (size_t)n_cnt * 4)); TRY(dgrow(c->s_off, (size_t)n_cnt * 4));
	TRY(dgrow(c->s_spill, (size_t)sc * sizeof(SeedRec)));
	TRY(dgrow(c->b_task, (size_t)tc * sizeof(SeedTask)));
	TRY(dgrow(c->s_pack, (size_t)n_reads * kPackWords * 4 + 64));
	TRY(dgrow(c->s_qflag, (size_t)n_reads * 4 + 64));
	size_t tmp_bytes = 0;
	HIPCHK(scan_u32(nullptr, &tmp_bytes, nullptr, nullptr, n_cnt, st));
	TRY(dgrow(c->s_tmp, tmp_bytes));
	if (c->seg_len > 0 && n_reads > 0) TRY(seg_alloc(c, n_reads, st));
	return MBG_OK;
}

// The launches of a seeding pass over pools seed_alloc sized, in three phases with no allocation and
// no host synchronisation, so that run_graph can capture them: before pass 1 (counters cleared, reads
// packed, long reads' segments laid out), pass 1 (with the segment lanes, fix-ups and fallbacks of
// long reads), and after it (pass-2 tasks, pass 2, compaction).
struct SeedPass {
	int32_t n_reads, min_len, max_sub_occ;
	const int64_t *off;
	const int32_t *len;
	const uint8_t *seq;
	unsigned long long *stats;
	SeedArgs a;
};

static int seed_phase_pre(mbg_ctx_t *c, hipStream_t st, SeedPass &sp)
{
	const int64_t tc = c->cap_task, sc = c->cap_spill, n_cnt = (int64_t)sp.n_reads + tc;
	uint32_t *ctr = (uint32_t*)c->b_ctr.p;
	uint32_t *cnt = (uint32_t*)c->s_cnt.p;
	HIPCHK(zero_async(ctr, 64, st));
	HIPCHK(zero_async(cnt, (size_t)n_cnt * 4, st));
	HIPCHK(zero_async(c->s_qflag.p, (size_t)sp.n_reads * 4 + 4, st));
	HIPCHK(launch_pack_reads(sp.seq, sp.off, sp.len, sp.n_reads, (uint32_t*)c->s_pack.p, (uint32_t*)c->s_qflag.p, st));
	SeedArgs &a = sp.a;
	a = SeedArgs{};
	a.seq = sp.seq; a.off = sp.off; a.len = sp.len; a.n_tasks = (uint32_t)sp.n_reads; a.min_len = sp.min_len; a.next = ctr;
	a.packed = (const uint32_t*)c->s_pack.p; a.qflag = (const uint32_t*)c->s_qflag.p;
	a.slots = (SeedRec*)c->s_slots1.p; a.slot_cnt = cnt; a.kslot = kSlot1;
	a.spill = (SeedRec*)c->s_spill.p; a.n_spill = (unsigned long long*)(ctr + 6); a.spill_cap = sc;
	a.stats = sp.stats;
	if (c->seg_len > 0 && sp.n_reads > 0) TRY(seg_setup(c, st, sp.n_reads, sp.len, a));
	return MBG_OK;
}

// side: long reads' segments k >= 1 run on the context's side stream beside pass 1 proper (the two
// launches share no data: pass 1 writes the reads' slots and segment 0's handoffs, the segment lanes
// their own slots and chains), so that the two latency-bound launches overlap; the fix-up waits for both
static int seed_phase_pass1(mbg_ctx_t *c, hipStream_t st, SeedPass &sp, bool side = false)
{
	const bool seg = c->seg_len > 0 && sp.n_reads > 0;
	if (seg && side && c->side) {
		HIPCHK(hipEventRecord(c->fork_ev, st));
		HIPCHK(hipStreamWaitEvent(c->side, c->fork_ev, 0));
		TRY(seg_segments(c, c->side, sp.n_reads, sp.a));
		HIPCHK(launch_smem(c->ix, sp.a, 1, gather_grid(c, sp.n_reads), st));
		HIPCHK(hipEventRecord(c->join_ev, c->side));
		HIPCHK(hipStreamWaitEvent(st, c->join_ev, 0));
		return seg_run(c, st, sp.n_reads, sp.len, sp.a, false);
	}
	if (sp.n_reads > 0) HIPCHK(launch_smem(c->ix, sp.a, 1, gather_grid(c, sp.n_reads), st));
	if (seg) TRY(seg_run(c, st, sp.n_reads, sp.len, sp.a));
	return MBG_OK;
}

static int seed_phase_post(mbg_ctx_t *c, hipStream_t st, SeedPass &sp)
{
	const int64_t rc = c->cap_rec, tc = c->cap_task, sc = c->cap_spill, n_cnt = (int64_t)sp.n_reads + tc;
	size_t tmp_bytes = c->s_tmp.n;
	uint32_t *ctr = (uint32_t*)c->b_ctr.p;
	int64_t *n_out = (int64_t*)(ctr + 2);
	unsigned long long *n_spill = (unsigned long long*)(ctr + 6);
	uint32_t *cnt = (uint32_t*)c->s_cnt.p;
	if (sp.n_reads > 0) {
		HIPCHK(launch_subtask((const SeedRec*)c->s_slots1.p, cnt, kSlot1, sp.n_reads, (const SeedRec*)c->s_spill.p, n_spill, sc,
		                      sp.min_len, sp.max_sub_occ, (SeedTask*)c->b_task.p, ctr + 1, tc, st));
		HIPCHK(zero_async(ctr, 4, st));
		SeedArgs a2 = sp.a;
		a2.tasks = (const SeedTask*)c->b_task.p; a2.n_tasks = 0xffffffffu; a2.n_tasks_dev = ctr + 1; a2.task_cap = tc;
		a2.slots = (SeedRec*)c->s_slots2.p; a2.slot_cnt = cnt + sp.n_reads; a2.kslot = kSlot2;
		a2.stats = sp.stats ? sp.stats + 3 : nullptr;
		HIPCHK(launch_smem(c->ix, a2, 2, gather_grid(c, sp.n_reads), st));
	}
	HIPCHK(scan_u32(c->s_tmp.p, &tmp_bytes, cnt, (uint32_t*)c->s_off.p, n_cnt, st));
	HIPCHK(launch_compact((const SeedRec*)c->s_slots1.p, (const SeedRec*)c->s_slots2.p, kSlot1, kSlot2, sp.n_reads, n_cnt, cnt,
	                      (const uint32_t*)c->s_off.p, (const SeedRec*)c->s_spill.p, n_spill, sc, (SeedRec*)c->p_recs.p, rc, n_out, st));
	return MBG_OK;
}

// Two-pass seeding of device-resident reads into c->p_recs (flat; c->cap_rec entries) with the
// record count at *n_out_dev.  Per-task output slots (kSlot1 per read, kSlot2 per pass-2 task) are
// compacted by an exclusive scan, so no output position needs an atomic except overflow spills.
// Counters in c->b_ctr: u32 [0] task counter, [1] pass-2 task count, i64 [1] record count,
// u64 [3] spill count (u32 [4] is the SA row counter, u64 [6] the gather checksum, u32 [14..18] the
// DP job counters).
//  * inner: the caller is capturing the stream itself; the three phases are only enqueued.
//  * otherwise, with graphs on, the phases before and after pass 1 run as captured graphs and pass 1
//    between them as plain launches, so that ev_pre (before pass 1), ev[0] (after it) and ev[1] (after
//    pass 2) time it; with graphs off every launch goes to the stream.
static int seed_core(mbg_ctx_t *c, hipStream_t st, int32_t n_reads, const int64_t *off, const int32_t *len, const uint8_t *seq,
                     int32_t min_len, int32_t max_sub_occ, unsigned long long *stats, int64_t **n_out_dev, hipEvent_t *ev,
                     hipEvent_t ev_pre = nullptr, bool inner = false, int64_t bases = -1)
{
	if (!inner) TRY(seed_alloc(c, st, n_reads, bases));
	*n_out_dev = (int64_t*)((uint32_t*)c->b_ctr.p + 2);
	SeedPass sp{n_reads, min_len, max_sub_occ, off, len, seq, stats, SeedArgs{}};
	const bool seg = c->seg_len > 0 && n_reads > 0;
	if (inner) {                // events only from a caller that is not capturing
		TRY(seed_phase_pre(c, st, sp));
		if (ev_pre) HIPCHK(hipEventRecord(ev_pre, st));
		TRY(seed_phase_pass1(c, st, sp));
		if (ev) HIPCHK(hipEventRecord(ev[0], st));
		TRY(seed_phase_post(c, st, sp));
		if (ev) HIPCHK(hipEventRecord(ev[1], st));
		return MBG_OK;
	}
	TRY(run_graph(c, st, seg ? kGraphSeedPreSeg : kGraphSeedPre, [&](hipStream_t s) { return seed_phase_pre(c, s, sp); }));
	if (ev_pre) HIPCHK(hipEventRecord(ev_pre, st));
	TRY(seed_phase_pass1(c, st, sp, c->seg_side));
	if (ev) HIPCHK(hipEventRecord(ev[0], st));
	TRY(run_graph(c, st, kGraphSeedPost, [&](hipStream_t s) { return seed_phase_post(c, s, sp); }));
	if (ev) HIPCHK(hipEventRecord(ev[1], st));
	return MBG_OK;
}

// after the stream has synchronised: grow any seeding pool that overflowed; true if one did
static bool seed_overflowed(mbg_ctx_t *c, int64_t *n_rec_host)
{
	uint32_t h[8];
	if (hipMemcpy(h, c->b_ctr.p, 32, hipMemcpyDeviceToHost) != hipSuccess) return false;
	int64_t n_out;
	uint64_t n_spill;
	memcpy(&n_out, h + 2, 8);
	memcpy(&n_spill, h + 6, 8);
	bool ov = false;
	if ((int64_t)h[1] > c->cap_task) c->cap_task = (int64_t)h[1] * 2, ov = true;
	if (c->seg_len > 0 && c->g_ctr.p) {     // segment tasks past the pool ran as fallbacks: exact, but slow
		uint32_t g[8];
		if (hipMemcpy(g, c->g_ctr.p, 32, hipMemcpyDeviceToHost) == hipSuccess && (int64_t)g[5] > c->cap_seg) c->cap_seg = (int64_t)g[5] * 2;
	}
	if ((int64_t)n_spill > c->cap_spill) c->cap_spill = (int64_t)n_spill * 2, ov = true;
	if (n_out > c->cap_rec) c->cap_rec = n_out * 2, ov = true;
	if (n_rec_host) *n_rec_host = n_out;
	return ov;
}

int mbg_seed_batch_dev_async(mbg_ctx_t *c, void *stream, int32_t n_seq, const int64_t *off, const int32_t *len, const uint8_t *seq,
                             int32_t min_len, int32_t max_sub_occ, mbg_seed_rec_t *out, int64_t cap, int64_t *n_out_dev)
{
	if (c == nullptr || n_seq < 0 || min_len < 1 || cap < 0 || (n_seq > 0 && (!off || !len || !seq || !out)) || n_out_dev == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	TRY(ctx_enter(c));
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	int64_t *nd = nullptr;
	TRY(seed_core(c, st, n_seq, off, len, seq, min_len, max_sub_occ, nullptr, &nd, nullptr));
	const int64_t m = cap < c->cap_rec ? cap : c->cap_rec;
	if (m > 0) HIPCHK(hipMemcpyAsync(out, c->p_recs.p, (size_t)m * sizeof(SeedRec), hipMemcpyDeviceToDevice, st));
	HIPCHK(hipMemcpyAsync(n_out_dev, nd, 8, hipMemcpyDeviceToDevice, st));
	return ctx_mark(c, st);
}

int mbg_seed_batch_dev(mbg_ctx_t *c, void *stream, int32_t n_seq, const int64_t *off, const int32_t *len, const uint8_t *seq,
                       int32_t min_len, int32_t max_sub_occ, mbg_seed_rec_t *out, int64_t cap, int64_t *n_out_host)
{
	if (c == nullptr || n_seq < 0 || min_len < 1 || cap < 0 || (n_seq > 0 && (!off || !len || !seq || !out)) || n_out_host == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	TRY(ctx_enter(c));
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	for (int attempt = 0; attempt < 4; ++attempt) {
		int64_t *nd = nullptr;
		TRY(seed_core(c, st, n_seq, off, len, seq, min_len, max_sub_occ, nullptr, &nd, nullptr));
		HIPCHK(hipStreamSynchronize(st));
		if (seed_overflowed(c, n_out_host)) continue;
		const int64_t m = cap < *n_out_host ? cap : *n_out_host;
		if (m > 0) HIPCHK(hipMemcpy(out, c->p_recs.p, (size_t)m * sizeof(SeedRec), hipMemcpyDeviceToDevice));
		return *n_out_host > cap ? MBG_EOVERFLOW : MBG_OK;
	}
	return MBG_EOVERFLOW;
}

int mbg_seed_batch(mbg_ctx_t *c, int32_t n_seq, const int32_t *len, const uint8_t *const *seq, int32_t min_len, int32_t max_sub_occ,
                   mbg_sai_t **out, int64_t *offs)
{
	if (c == nullptr || n_seq < 0 || min_len < 1 || out == nullptr || offs == nullptr || (n_seq > 0 && (!len || !seq))) return MBG_EARG;
	*out = nullptr;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	TRY(ctx_enter(c));
	std::vector<int64_t> off(n_seq + 1, 0);
	for (int32_t i = 0; i < n_seq; ++i) {
		if (len[i] < 0 || (len[i] > 0 && seq[i] == nullptr)) return MBG_EARG;
		off[i + 1] = off[i] + len[i];
	}
	const int64_t tot = off[n_seq];
	std::vector<uint8_t> buf(tot > 0 ? tot : 1);
	for (int32_t i = 0; i < n_seq; ++i) if (len[i] > 0) memcpy(buf.data() + off[i], seq[i], len[i]);
	TRY(dgrow(c->b_in[0], tot + 16)); TRY(dgrow(c->b_in[1], (n_seq + 1) * 8)); TRY(dgrow(c->b_in[2], (n_seq + 1) * 4));
	HIPCHK(hipMemcpyAsync(c->b_in[0].p, buf.data(), tot, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[1].p, off.data(), (n_seq + 1) * 8, hipMemcpyHostToDevice, c->stream));
	if (n_seq > 0) HIPCHK(hipMemcpyAsync(c->b_in[2].p, len, n_seq * 4, hipMemcpyHostToDevice, c->stream));
	int64_t n_out = 0;
	bool ok = false;
	seg_hint(c, len, n_seq);
	for (int attempt = 0; attempt < 4 && !ok; ++attempt) {
		int64_t *nd = nullptr;
		TRY(seed_core(c, c->stream, n_seq, (const int64_t*)c->b_in[1].p, (const int32_t*)c->b_in[2].p, (const uint8_t*)c->b_in[0].p,
		              min_len, max_sub_occ, nullptr, &nd, nullptr));
		HIPCHK(hipStreamSynchronize(c->stream));
		ok = !seed_overflowed(c, &n_out);
	}
	if (!ok) return MBG_EOVERFLOW;
	std::vector<SeedRec> recs(n_out);
	if (n_out) HIPCHK(hipMemcpy(recs.data(), c->p_recs.p, n_out * sizeof(SeedRec), hipMemcpyDeviceToHost));
	// group by read: pass-1 records first, each group in device order
	std::vector<int64_t> cnt(n_seq + 1, 0);
	for (const SeedRec &r : recs) ++cnt[r.rid + 1];
	for (int32_t i = 0; i < n_seq; ++i) cnt[i + 1] += cnt[i];
	for (int32_t i = 0; i <= n_seq; ++i) offs[i] = cnt[i];
	mbg_sai_t *o = (mbg_sai_t*)malloc((n_out > 0 ? n_out : 1) * sizeof(mbg_sai_t));
	if (o == nullptr) return MBG_ENOMEM;
	std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
	for (int pass = 1; pass <= 2; ++pass)
		for (const SeedRec &r : recs) {
			if ((int)r.pass != pass) continue;
			mbg_sai_t &s = o[pos[r.rid]++];
			s.x[0] = r.x0; s.x[1] = r.x1; s.size = r.size; s.info = r.info;
		}
	*out = o;
	return MBG_OK;
}

static int sa_dev_locked(mbg_ctx_t *c, hipStream_t st, int64_t n, uint64_t *x)
{
	if (n <= 0) return MBG_OK;
	if (n > 0xffffffffLL) return MBG_EUNSUPPORTED;
	TRY(dgrow(c->b_ctr, 128));
	uint32_t *next = (uint32_t*)c->b_ctr.p + 4;
	HIPCHK(hipMemsetAsync(next, 0, 4, st));
	HIPCHK(launch_sa(c->ix, n, nullptr, x, next, nullptr, gather_grid(c, n), st));
	return MBG_OK;
}

int mbg_sa_batch_dev(mbg_ctx_t *c, void *stream, int64_t n, uint64_t *x)
{
	if (c == nullptr || n < 0 || (n > 0 && x == nullptr)) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	TRY(ctx_enter(c));
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	TRY(sa_dev_locked(c, st, n, x));
	return ctx_mark(c, st);
}

int mbg_sa_batch_dev_noalloc(mbg_ctx_t *c, void *stream, int64_t n, uint64_t *x)
{
	if (c == nullptr || n < 0 || (n > 0 && x == nullptr)) return MBG_EARG;
	if (n > MBG_SA_NOALLOC_MAX_ROWS) return MBG_EUNSUPPORTED;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	if (n == 0) return MBG_OK;
	TRY(ctx_enter(c));
	uint32_t *ptr = nullptr;
	{
		std::lock_guard<std::mutex> g(g_noalloc_mu);
		ptr = g_noalloc_ptr[c->device];
	}
	if (ptr == nullptr) return MBG_EHIP;            // mbg_init resolves it; it cannot be missing here
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	hipLaunchKernelGGL(noalloc_clear_kernel, dim3(1), dim3(kNoAllocWords), 0, st, ptr, 1);
	HIPCHK(hipGetLastError());
	HIPCHK(launch_sa(c->ix, n, nullptr, x, ptr, nullptr, gather_grid(c, n), st));
	return ctx_mark(c, st);
}

int mbg_sa_batch(mbg_ctx_t *c, int64_t n, uint64_t *x)
{
	if (c == nullptr || n < 0 || (n > 0 && x == nullptr)) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	if (n == 0) return MBG_OK;
	TRY(ctx_enter(c));
	for (int64_t i = 0; i < n; ++i) if (x[i] > c->ix.seq_len) return MBG_EARG;
	TRY(dgrow(c->b_in[3], n * 8));
	HIPCHK(hipMemcpyAsync(c->b_in[3].p, x, n * 8, hipMemcpyHostToDevice, c->stream));
	TRY(sa_dev_locked(c, c->stream, n, (uint64_t*)c->b_in[3].p));
	HIPCHK(hipMemcpyAsync(x, c->b_in[3].p, n * 8, hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipStreamSynchronize(c->stream));
	return MBG_OK;
}

// ---- DP --------------------------------------------------------------------------------------

static const int kMaxCap = 4096;        // circular row buffer in LDS: 12*cap bytes per wave
static const int kMaxGcap = 1 << 22;    // circular row buffer in device memory, for wider bands

static const int kMaxQt = 16384;        // LDS bytes for staging a job's query + target

static int dp_plan(int32_t max_qlen, int32_t max_tlen, int32_t max_w, int *cap_out, int64_t *slab_out, int *qt_out, int *tcap_out)
{
	int w, n_col_;
	dp_geom(max_qlen, max_tlen, max_w, &w, &n_col_);
	const int nb = n_col_ * 16;
	int cap = 64;
	while (cap < nb + 17) cap <<= 1;
	if (cap > kMaxGcap) return MBG_EUNSUPPORTED;
	*cap_out = cap;                      // capped at kMaxCap by dp_state; wider jobs keep their rows in HBM
	int64_t slab = ((int64_t)max_qlen + max_tlen) * nb + 256;
	*slab_out = (slab + 255) & ~(int64_t)255;
	// jobs whose query and target fit are staged in LDS and run on the register schedule
	const int64_t qt = (((int64_t)max_qlen + 15) & ~15) + (((int64_t)max_tlen + 15) & ~15);
	*qt_out = qt <= kMaxQt ? (int)qt : 0;
	*tcap_out = dp_team_cap(nb);         // bands over the fast kernel's ring run on a team of waves
	return MBG_OK;
}

// ring lengths for a launch of `g` blocks whose widest band needs `cap`: jobs up to kMaxCap run on
// the LDS ring, wider ones on a per-block ring of 12*cap bytes in device memory
static int dp_state(mbg_ctx_t *c, int64_t g, int cap, int tcap, DpArgs &a)
{
	a.tcap = tcap;
	a.cap = cap < kMaxCap ? cap : kMaxCap;
	a.gcap = 0, a.gstate = nullptr;
	if (cap > kMaxCap) {
		TRY(dgrow(c->b_gstate, (size_t)g * 12 * cap));
		a.gcap = cap, a.gstate = (uint8_t*)c->b_gstate.p;
	}
	return MBG_OK;
}

static int dp_dev_locked(mbg_ctx_t *c, hipStream_t st, bool dual, int32_t n_jobs, const DpJob *jobs, const uint8_t *qseq, const uint8_t *tseq,
                         int8_t m, const int8_t *mat, DpRes *res, uint32_t *cigar, int cap, int64_t slab, int qt, int tcap)
{
	if (n_jobs <= 0) return MBG_OK;
	// resident blocks (one wave each): enough to fill the GPU, bounded by 2 GiB of backtrack slabs
	int64_t g = (int64_t)c->n_cu * c->dp_wpc;
	if (g > n_jobs) g = n_jobs;
	const int64_t max_by_mem = ((int64_t)2 << 30) / slab;
	if (g > max_by_mem) g = max_by_mem;
	if (g < 1) g = 1;
	TRY(dgrow(c->b_slab, (size_t)g * slab));
	DpArgs a{};
	a.jobs = jobs; a.n_jobs = n_jobs; a.qseq = qseq; a.tseq = tseq; a.res = res; a.cigar = cigar;
	a.pbuf = (uint8_t*)c->b_slab.p; a.slab = slab; a.qt_cap = qt; a.m = m;
	TRY(dp_state(c, g, cap, tcap, a));
	TRY(dgrow(c->b_ctr, 128));
	a.next = (unsigned int*)c->b_ctr.p + 14;
	HIPCHK(hipMemsetAsync(a.next, 0, 20, st));
	TRY(dgrow(c->b_defer, (size_t)n_jobs * 4));
	a.defer = (int32_t*)c->b_defer.p;
	dp_set_matrix(a, m, mat);
	HIPCHK(launch_dp(dual, a, (unsigned)g, st));
	return MBG_OK;
}

// results of a DP batch (device res / CIGAR slots at cigar_off) back to the caller, after dev_ev[0..1]
// bracketed its kernels on c->stream: only the CIGAR operations written cross PCIe, packed on the
// device at host-computed offsets.  The host jobs give each job's qlen, tlen and cigar_off: either DP
// jobs (dj; a CIGAR unless KSW_EZ_SCORE_ONLY) or mb_align_pair calls (aj; always a CIGAR slot, since
// the fast paths write one whatever the flag).
static int dp_readback(mbg_ctx_t *c, int32_t n_jobs, const DpJob *djobs, const mbg_dp_job_t *dj, const mbg_align_job_t *aj,
                       const DpRes *dres, mbg_dp_result_t *res, const uint32_t *dcigar, uint32_t *cigar)
{
	auto has_cigar = [&](int32_t i) { return aj != nullptr || !(dj[i].flag & 0x01); };
	auto bound = [&](int32_t i) { return aj ? aj[i].qlen + aj[i].tlen + 1 : dj[i].qlen + dj[i].tlen; };
	auto coff = [&](int32_t i) { return aj ? aj[i].cigar_off : dj[i].cigar_off; };
	HIPCHK(hipMemcpyAsync(res, dres, (size_t)n_jobs * sizeof(DpRes), hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipStreamSynchronize(c->stream));
	{
		float ms = 0;
		HIPCHK(hipEventElapsedTime(&ms, c->dev_ev[0], c->dev_ev[1]));
		c->cnt.ms_dp += ms, ++c->cnt.dp_launches, c->cnt.dp_jobs += n_jobs;
		for (int32_t i = 0; i < n_jobs; ++i) c->cnt.dp_cells += res[i].cells;
	}
	if (dcigar == nullptr) return MBG_OK;
	std::vector<int64_t> poff(n_jobs + 1);
	int64_t tot = 0;
	for (int32_t i = 0; i < n_jobs; ++i) {
		poff[i] = tot;
		if (has_cigar(i)) {
			if (res[i].n_cigar < 0 || res[i].n_cigar > bound(i)) return MBG_EHIP;
			tot += res[i].n_cigar;
		}
	}
	poff[n_jobs] = tot;
	if (tot > 0) {
		TRY(dgrow(c->b_in[5], (size_t)(n_jobs + 1) * 8)); TRY(dgrow(c->b_out[2], (size_t)tot * 4));
		TRY(hgrow(c->h_pinned, (size_t)tot * 4));
		HIPCHK(hipMemcpyAsync(c->b_in[5].p, poff.data(), (size_t)(n_jobs + 1) * 8, hipMemcpyHostToDevice, c->stream));
		HIPCHK(launch_cigar_pack(djobs, dres, n_jobs, (const int64_t*)c->b_in[5].p, dcigar, (uint32_t*)c->b_out[2].p, c->stream));
		HIPCHK(hipMemcpyAsync(c->h_pinned.p, c->b_out[2].p, (size_t)tot * 4, hipMemcpyDeviceToHost, c->stream));
		HIPCHK(hipStreamSynchronize(c->stream));
		const uint32_t *pk = (const uint32_t*)c->h_pinned.p;
		for (int32_t i = 0; i < n_jobs; ++i)
			if (has_cigar(i) && res[i].n_cigar > 0)
				memcpy(cigar + coff(i), pk + poff[i], (size_t)res[i].n_cigar * 4);
	}
	return MBG_OK;
}

static int dp_batch(mbg_ctx_t *c, bool dual, int32_t n_jobs, const mbg_dp_job_t *jobs, const uint8_t *qseq, int64_t qbytes,
                    const uint8_t *tseq, int64_t tbytes, int8_t m, const int8_t *mat, mbg_dp_result_t *res, uint32_t *cigar, int64_t cigar_cap)
{
	if (c == nullptr || n_jobs < 0 || m < 0 || m > 5 || (n_jobs > 0 && (!jobs || !res || !mat))) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (n_jobs == 0) return MBG_OK;
	TRY(ctx_enter(c));
	int32_t mq = 1, mt = 1, mw = 1;
	bool any_cigar = false;
	for (int32_t i = 0; i < n_jobs; ++i) {
		const mbg_dp_job_t &j = jobs[i];
		if (j.qlen < 0 || j.tlen < 0 || j.qoff < 0 || j.toff < 0 || j.qoff + j.qlen > qbytes || j.toff + j.tlen > tbytes) return MBG_EARG;
		if (!(j.flag & 0x01)) {
			any_cigar = true;
			if (cigar == nullptr || j.cigar_off < 0 || j.cigar_off + j.qlen + j.tlen > cigar_cap) return MBG_EARG;
		}
		int w, ncol;
		dp_geom(j.qlen, j.tlen, j.w, &w, &ncol);
		mq = j.qlen > mq ? j.qlen : mq;
		mt = j.tlen > mt ? j.tlen : mt;
		mw = w > mw ? w : mw;
	}
	int cap, qt, tcap;
	int64_t slab;
	TRY(dp_plan(mq, mt, mw, &cap, &slab, &qt, &tcap));
	if (!any_cigar) slab = 256;
	TRY(dgrow(c->b_in[0], qbytes + 16)); TRY(dgrow(c->b_in[1], tbytes + 16)); TRY(dgrow(c->b_in[4], (size_t)n_jobs * sizeof(DpJob)));
	TRY(dgrow(c->b_out[0], (size_t)n_jobs * sizeof(DpRes)));
	if (any_cigar) TRY(dgrow(c->b_out[1], (size_t)cigar_cap * 4));
	HIPCHK(hipMemcpyAsync(c->b_in[0].p, qseq, qbytes, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[1].p, tseq, tbytes, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[4].p, jobs, (size_t)n_jobs * sizeof(DpJob), hipMemcpyHostToDevice, c->stream));
	for (hipEvent_t &e : c->dev_ev) if (!e) HIPCHK(hipEventCreate(&e));
	HIPCHK(hipEventRecord(c->dev_ev[0], c->stream));
	TRY(dp_dev_locked(c, c->stream, dual, n_jobs, (const DpJob*)c->b_in[4].p, (const uint8_t*)c->b_in[0].p, (const uint8_t*)c->b_in[1].p,
	                  m, mat, (DpRes*)c->b_out[0].p, any_cigar ? (uint32_t*)c->b_out[1].p : nullptr, cap, slab, qt, tcap));
	HIPCHK(hipEventRecord(c->dev_ev[1], c->stream));
	return dp_readback(c, n_jobs, (const DpJob*)c->b_in[4].p, jobs, nullptr, (const DpRes*)c->b_out[0].p, res,
	                   any_cigar ? (const uint32_t*)c->b_out[1].p : nullptr, cigar);
}

int mbg_extd2_batch(mbg_ctx_t *c, int32_t n_jobs, const mbg_dp_job_t *jobs, const uint8_t *qseq, int64_t qbytes, const uint8_t *tseq, int64_t tbytes,
                    int8_t m, const int8_t *mat, mbg_dp_result_t *res, uint32_t *cigar, int64_t cigar_cap)
{
	return dp_batch(c, true, n_jobs, jobs, qseq, qbytes, tseq, tbytes, m, mat, res, cigar, cigar_cap);
}

int mbg_extz2_batch(mbg_ctx_t *c, int32_t n_jobs, const mbg_dp_job_t *jobs, const uint8_t *qseq, int64_t qbytes, const uint8_t *tseq, int64_t tbytes,
                    int8_t m, const int8_t *mat, mbg_dp_result_t *res, uint32_t *cigar, int64_t cigar_cap)
{
	return dp_batch(c, false, n_jobs, jobs, qseq, qbytes, tseq, tbytes, m, mat, res, cigar, cigar_cap);
}

int mbg_align_pair_batch(mbg_ctx_t *c, const mbg_align_opt_t *opt, int32_t n, const mbg_align_job_t *jobs, const uint8_t *qseq,
                         int64_t qbytes, const uint8_t *tseq, int64_t tbytes, const int8_t *mat, mbg_dp_result_t *res, uint32_t *cigar,
                         int64_t cigar_cap)
{
	if (c == nullptr || opt == nullptr || n < 0 || (n > 0 && (!jobs || !res || !mat || !cigar))) return MBG_EARG;
	if (opt->e <= 0 || opt->e2 <= 0) return MBG_EARG;            // max_bw_from_mm divides by them
	std::lock_guard<std::mutex> lk(c->mu);
	if (n == 0) return MBG_OK;
	TRY(ctx_enter(c));
	int32_t mq = 1, mt = 1, mw = 1;
	for (int32_t i = 0; i < n; ++i) {
		const mbg_align_job_t &j = jobs[i];
		if (j.qlen < 0 || j.tlen < 0 || j.qoff < 0 || j.toff < 0 || j.qoff + j.qlen > qbytes || j.toff + j.tlen > tbytes) return MBG_EARG;
		if (j.cigar_off < 0 || j.cigar_off + j.qlen + j.tlen + 1 > cigar_cap) return MBG_EARG;
		int w, ncol;
		dp_geom(j.qlen, j.tlen, j.w, &w, &ncol);                  // the band before any cap: an upper bound
		mq = j.qlen > mq ? j.qlen : mq;
		mt = j.tlen > mt ? j.tlen : mt;
		mw = w > mw ? w : mw;
	}
	int cap, qt, tcap;
	int64_t slab;
	TRY(dp_plan(mq, mt, mw, &cap, &slab, &qt, &tcap));
	const bool dual = !(opt->q == opt->q2 && opt->e == opt->e2); // align.c:373-378
	TRY(dgrow(c->b_in[0], qbytes + 16)); TRY(dgrow(c->b_in[1], tbytes + 16)); TRY(dgrow(c->b_in[2], (size_t)n * sizeof(AlignJob)));
	TRY(dgrow(c->b_in[3], 32)); TRY(dgrow(c->b_in[4], (size_t)n * sizeof(DpJob)));
	TRY(dgrow(c->l_buf[0], (size_t)n)); TRY(dgrow(c->l_buf[1], (size_t)n * sizeof(DpRes)));
	TRY(dgrow(c->b_out[0], (size_t)n * sizeof(DpRes))); TRY(dgrow(c->b_out[1], (size_t)cigar_cap * 4));
	HIPCHK(hipMemcpyAsync(c->b_in[0].p, qseq, qbytes, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[1].p, tseq, tbytes, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[2].p, jobs, (size_t)n * sizeof(AlignJob), hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[3].p, mat, 25, hipMemcpyHostToDevice, c->stream));
	for (hipEvent_t &e : c->dev_ev) if (!e) HIPCHK(hipEventCreate(&e));
	HIPCHK(hipEventRecord(c->dev_ev[0], c->stream));
	AlignArgs a{};
	a.jobs = (const AlignJob*)c->b_in[2].p; a.n = n; a.qseq = (const uint8_t*)c->b_in[0].p; a.tseq = (const uint8_t*)c->b_in[1].p;
	a.mat = (const int8_t*)c->b_in[3].p;
	a.sc_a = opt->a, a.sc_b = opt->b, a.q = opt->q, a.e = opt->e, a.q2 = opt->q2, a.e2 = opt->e2;
	a.generic = (opt->b_ts != 0 && opt->b != opt->b_ts) || opt->meth;   // align.c:334-335
	a.max_sw_mat = opt->max_sw_mat;
	a.dp = (DpJob*)c->b_in[4].p; a.kind = (uint8_t*)c->l_buf[0].p; a.res = (DpRes*)c->l_buf[1].p; a.cigar = (uint32_t*)c->b_out[1].p;
	HIPCHK(launch_align_plan(a, c->stream));
	TRY(dp_dev_locked(c, c->stream, dual, n, a.dp, a.qseq, a.tseq, 5, mat, (DpRes*)c->b_out[0].p, a.cigar, cap, slab, qt, tcap));
	HIPCHK(launch_align_merge(a.kind, a.res, (DpRes*)c->b_out[0].p, n, c->stream));
	HIPCHK(hipEventRecord(c->dev_ev[1], c->stream));
	return dp_readback(c, n, a.dp, nullptr, jobs, (const DpRes*)c->b_out[0].p, res, a.cigar, cigar);
}

// ksw_gen_nt4_mat (ksw2.h:199-219): 5x5 scores for A,C,G,T,N; mt 1 = C->T, 2 = G->A conversion
static void gen_nt4_mat(int8_t *mat, int a, int b, int b_ts, int b_ambi, int mt)
{
	a = a < 0 ? -a : a, b = b > 0 ? -b : b, b_ambi = b_ambi > 0 ? -b_ambi : b_ambi;
	for (int i = 0; i < 4; ++i) {
		for (int j = 0; j < 4; ++j) mat[i * 5 + j] = (int8_t)(i == j ? a : b);
		mat[i * 5 + 4] = (int8_t)b_ambi;
	}
	for (int j = 0; j < 5; ++j) mat[20 + j] = (int8_t)b_ambi;
	if (b_ts != 0 && b_ts != b) {
		b_ts = b_ts > 0 ? -b_ts : b_ts;
		mat[0 * 5 + 2] = mat[1 * 5 + 3] = mat[2 * 5 + 0] = mat[3 * 5 + 1] = (int8_t)b_ts;
	}
	if (mt == 1) mat[1 * 5 + 3] = (int8_t)a;
	else if (mt == 2) mat[2 * 5 + 0] = (int8_t)a;
}

// batch-wide ksw_ll parameters: the matrix's u8 bias and largest entry as ksw_ll_qinit computes them
// (ksw2_ll_sse.c:67-74, uint8 arithmetic), the score table, the SUBO/STOP thresholds
static int ll_params(int8_t m, const int8_t *mat, int32_t gapo, int32_t gape, int32_t xtra, LlArgs &a, int8_t sc[40], int *mat_min)
{
	if (m < 1 || m > 5 || mat == nullptr) return MBG_EARG;
	if (gapo < 1 || gape < 0 || gapo + gape > 255) return MBG_EUNSUPPORTED;
	uint8_t sh = 127, md = 0;
	for (int k = 0; k < m * m; ++k) {
		if (mat[k] < (int8_t)sh) sh = (uint8_t)mat[k];
		if (mat[k] > (int8_t)md) md = (uint8_t)mat[k];
	}
	a.shift = (uint8_t)(256 - sh), a.qmax = md, *mat_min = (int8_t)sh;
	if (a.qmax <= 0) return MBG_EUNSUPPORTED;      // the second-best window divides by it
	for (int t = 0; t < 5; ++t)                    // sc[t*8 + c]; c = 7 scores padding
		for (int q = 0; q < 8; ++q) sc[t * 8 + q] = t < m && q < m ? mat[t * m + q] : -1;
	a.gapo = gapo, a.gape = gape, a.m = m;
	a.minsc = (xtra & 0x40000) ? xtra & 0xffff : 0x10000;    // KSW_LL_SUBO, KSW_LL_STOP (ksw2.h:21-22)
	a.endsc = (xtra & 0x20000) ? xtra & 0xffff : 0x10000;
	return MBG_OK;
}

int mbg_update_extra_batch(mbg_ctx_t *c, int64_t n, mbg_extra_job_t *jobs, const uint8_t *qseq, int64_t qbytes, const uint8_t *tseq,
                           int64_t tbytes, const int8_t *mat, int8_t q, int8_t e, int32_t log_gap, uint32_t *cigar, int64_t cigar_cap)
{
	if (c == nullptr || n < 0 || (n > 0 && (!jobs || !qseq || !tseq || !mat || !cigar))) return MBG_EARG;
	int64_t cig_end = 0;
	for (int64_t i = 0; i < n; ++i) {
		const mbg_extra_job_t &j = jobs[i];
		int64_t ql = 0, tl = 0;
		if (j.n_cigar < 0 || j.cigar_off < 0 || j.cigar_off + j.n_cigar > cigar_cap || j.qoff < 0 || j.toff < 0) return MBG_EARG;
		for (int32_t k = 0; k < j.n_cigar; ++k) {      // the bases the CIGAR covers must lie in the buffers
			const uint32_t op = cigar[j.cigar_off + k] & 0xf, len = cigar[j.cigar_off + k] >> 4;
			if (op > 3) return MBG_EUNSUPPORTED;         // M, I, D and N only: the DP's operations
			if (op == 0 || op == 1) ql += len;
			if (op == 0 || op == 2 || op == 3) tl += len;
		}
		if (j.qoff + ql > qbytes || j.toff + tl > tbytes) return MBG_EARG;
		cig_end = j.cigar_off + j.n_cigar > cig_end ? j.cigar_off + j.n_cigar : cig_end;
	}
	std::lock_guard<std::mutex> lk(c->mu);
	if (n == 0) return MBG_OK;
	TRY(ctx_enter(c));
	TRY(dgrow(c->b_in[0], qbytes + 16)); TRY(dgrow(c->b_in[1], tbytes + 16)); TRY(dgrow(c->b_in[4], (size_t)n * sizeof(ExtraJob)));
	TRY(dgrow(c->b_out[1], (size_t)cig_end * 4 + 16));
	HIPCHK(hipMemcpyAsync(c->b_in[0].p, qseq, qbytes, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[1].p, tseq, tbytes, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[4].p, jobs, (size_t)n * sizeof(ExtraJob), hipMemcpyHostToDevice, c->stream));
	if (cig_end > 0) HIPCHK(hipMemcpyAsync(c->b_out[1].p, cigar, (size_t)cig_end * 4, hipMemcpyHostToDevice, c->stream));
	ExtraArgs a{};
	a.jobs = (ExtraJob*)c->b_in[4].p, a.n = n, a.qseq = (const uint8_t*)c->b_in[0].p, a.tseq = (const uint8_t*)c->b_in[1].p;
	a.cigar = (uint32_t*)c->b_out[1].p;
	for (int k = 0; k < 28; ++k) a.mat[k] = k < 25 ? mat[k] : 0;
	a.q = q, a.e = e, a.log_gap = log_gap != 0;
	HIPCHK(launch_extra(a, c->stream));
	HIPCHK(hipMemcpyAsync(jobs, c->b_in[4].p, (size_t)n * sizeof(ExtraJob), hipMemcpyDeviceToHost, c->stream));
	if (cig_end > 0) HIPCHK(hipMemcpyAsync(cigar, c->b_out[1].p, (size_t)cig_end * 4, hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipStreamSynchronize(c->stream));
	return MBG_OK;
}

// ---- anchors ----------------------------------------------------------------------------------

static int anchor_dev_locked(mbg_ctx_t *c, hipStream_t st, int32_t n_reads, int64_t n_slots, Sai *intv, const int64_t *uoff,
                             const int32_t *qlen, int32_t max_occ, uint64_t *rows, AncRec *anc, const int64_t *aoff, int64_t *nanc,
                             int64_t *nu)
{
	AncArgs a{};
	a.n_reads = n_reads, a.max_occ = max_occ, a.u = intv, a.uoff = uoff, a.qlen = qlen, a.aoff = aoff;
	a.rows = rows, a.anc = anc, a.nu = nu, a.nanc = nanc, a.lb = c->lb;
	HIPCHK(launch_anc_rows(a, st));
	TRY(sa_dev_locked(c, st, n_slots, rows));
	HIPCHK(launch_anc_build(a, st));
	return MBG_OK;
}

int mbg_anchor_batch_dev(mbg_ctx_t *c, void *stream, int32_t n_reads, int64_t n_slots, mbg_sai_t *intv, const int64_t *intv_off,
                         const int32_t *qlen, int32_t max_occ, uint64_t *rows, mbg_anchor_t *anc, const int64_t *anc_off, int64_t *n_anc,
                         int64_t *n_intv)
{
	if (c == nullptr || n_reads < 0 || n_slots < 0 || max_occ < 0) return MBG_EARG;
	if (n_reads > 0 && (!intv_off || !qlen || !anc_off || !n_anc || !n_intv || (n_slots > 0 && (!rows || !anc)))) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt || !c->has_l2b) return MBG_ENOIDX;
	if (n_reads == 0) return MBG_OK;
	TRY(ctx_enter(c));
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	TRY(anchor_dev_locked(c, st, n_reads, n_slots, (Sai*)intv, intv_off, qlen, max_occ, rows, (AncRec*)anc, anc_off, n_anc, n_intv));
	return ctx_mark(c, st);
}

int mbg_anchor_batch(mbg_ctx_t *c, int32_t n_reads, mbg_sai_t *intv, const int64_t *intv_off, const int32_t *qlen, int32_t max_occ,
                     mbg_anchor_t *anc, const int64_t *anc_off, int64_t *n_anc, int64_t *n_intv)
{
	if (c == nullptr || n_reads < 0 || max_occ < 0) return MBG_EARG;
	if (n_reads > 0 && (!intv_off || !qlen || !anc_off || !n_anc || intv_off[0] != 0 || anc_off[0] != 0)) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt || !c->has_l2b) return MBG_ENOIDX;
	if (n_reads == 0) return MBG_OK;
	const int64_t n_u = intv_off[n_reads], n_slots = anc_off[n_reads];
	if ((n_u > 0 && !intv) || (n_slots > 0 && !anc)) return MBG_EARG;
	for (int32_t r = 0; r < n_reads; ++r) {          // offsets, rows and the slots the reference preallocates
		if (intv_off[r + 1] < intv_off[r] || anc_off[r + 1] < anc_off[r] || qlen[r] < 0) return MBG_EARG;
		int64_t need = 0;
		for (int64_t i = intv_off[r]; i < intv_off[r + 1]; ++i) {
			const mbg_sai_t &u = intv[i];
			if (u.size > 0 && (u.x[0] > c->ix.seq_len || u.size - 1 > c->ix.seq_len - u.x[0])) return MBG_EARG;
			need += u.size < (uint64_t)max_occ ? (int64_t)u.size : max_occ;
		}
		if (need > anc_off[r + 1] - anc_off[r]) return MBG_EARG;
	}
	TRY(ctx_enter(c));
	hipStream_t st = c->stream;
	DBuf *b = c->a_buf;
	TRY(dgrow(b[0], (size_t)n_u * sizeof(Sai))); TRY(dgrow(b[1], (size_t)(n_reads + 1) * 8)); TRY(dgrow(b[2], (size_t)n_reads * 4));
	TRY(dgrow(b[3], (size_t)(n_reads + 1) * 8)); TRY(dgrow(b[4], (size_t)n_slots * 8)); TRY(dgrow(b[5], (size_t)n_slots * sizeof(AncRec)));
	TRY(dgrow(b[6], (size_t)n_reads * 8)); TRY(dgrow(b[7], (size_t)n_reads * 8));
	if (n_u > 0) HIPCHK(hipMemcpyAsync(b[0].p, intv, (size_t)n_u * sizeof(Sai), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b[1].p, intv_off, (size_t)(n_reads + 1) * 8, hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b[2].p, qlen, (size_t)n_reads * 4, hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b[3].p, anc_off, (size_t)(n_reads + 1) * 8, hipMemcpyHostToDevice, st));
	TRY(anchor_dev_locked(c, st, n_reads, n_slots, (Sai*)b[0].p, (const int64_t*)b[1].p, (const int32_t*)b[2].p, max_occ,
	                      (uint64_t*)b[4].p, (AncRec*)b[5].p, (const int64_t*)b[3].p, (int64_t*)b[6].p, (int64_t*)b[7].p));
	HIPCHK(hipMemcpyAsync(n_anc, b[6].p, (size_t)n_reads * 8, hipMemcpyDeviceToHost, st));
	if (n_intv) HIPCHK(hipMemcpyAsync(n_intv, b[7].p, (size_t)n_reads * 8, hipMemcpyDeviceToHost, st));
	if (n_u > 0) HIPCHK(hipMemcpyAsync(intv, b[0].p, (size_t)n_u * sizeof(Sai), hipMemcpyDeviceToHost, st));
	if (n_slots > 0) HIPCHK(hipMemcpyAsync(anc, b[5].p, (size_t)n_slots * sizeof(AncRec), hipMemcpyDeviceToHost, st));
	HIPCHK(hipStreamSynchronize(st));
	for (int32_t r = 0; r < n_reads; ++r)
		if (n_anc[r] < 0) return MBG_EOVERFLOW;
	return MBG_OK;
}

static int lchain_dev_locked(mbg_ctx_t *c, hipStream_t st, int32_t n_reads, int64_t n_slots, const AncRec *anc, const int64_t *aoff,
                             const int64_t *nanc, const mbg_chain_opt_t &o, void *work, size_t work_bytes, uint64_t *u, int32_t *n_u,
                             AncRec *out, int64_t *n_out, const ChainOpt *opts = nullptr, const int32_t *opt_idx = nullptr)
{
	if (work_bytes < MBG_LCHAIN_WORK_BYTES(n_slots)) return MBG_EARG;
	ChainArgs a{};
	a.opts = opts, a.opt_idx = opt_idx;
	a.n_reads = n_reads, a.anc = anc, a.aoff = aoff, a.nanc = nanc;
	a.max_dist_x = o.max_dist_x, a.max_dist_y = o.max_dist_y, a.bw = o.bw, a.max_skip = o.max_skip, a.max_iter = o.max_iter;
	a.min_sc = o.min_sc, a.pen_gap = o.pen_gap;
	a.ctg_off = c->lb.ctg_off, a.ctg_len = c->lb.ctg_len;
	char *w = (char*)work;                          // regions of n_slots elements, 256-byte aligned
	auto take = [&](size_t elem) { char *p = w; w += ((size_t)n_slots * elem + 255) / 256 * 256; return (void*)p; };
	w = (char*)(((uintptr_t)w + 255) / 256 * 256);
	a.p = (int64_t*)take(8), a.z = (U128*)take(16), a.ut = (uint64_t*)take(8), a.bb = (AncRec*)take(sizeof(AncRec));
	a.f = (int32_t*)take(4), a.t = (int32_t*)take(4), a.v = (int32_t*)take(4);
	a.u = u, a.n_u = n_u, a.b = out, a.n_b = n_out;
	HIPCHK(launch_lchain(a, st));
	return MBG_OK;
}

int mbg_lchain_batch_dev(mbg_ctx_t *c, void *stream, int32_t n_reads, int64_t n_slots, const mbg_anchor_t *anc, const int64_t *anc_off,
                         const int64_t *n_anc, const mbg_chain_opt_t *opt, void *work, size_t work_bytes, uint64_t *u, int32_t *n_u,
                         mbg_anchor_t *out, int64_t *n_out)
{
	if (c == nullptr || n_reads < 0 || n_slots < 0 || opt == nullptr) return MBG_EARG;
	if (n_reads > 0 && (!anc_off || !n_anc || !n_u || !n_out || (n_slots > 0 && (!anc || !work || !u || !out)))) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_l2b) return MBG_ENOIDX;
	if (n_reads == 0) return MBG_OK;
	TRY(ctx_enter(c));
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	TRY(lchain_dev_locked(c, st, n_reads, n_slots, (const AncRec*)anc, anc_off, n_anc, *opt, work, work_bytes, u, n_u, (AncRec*)out,
	                      n_out));
	return ctx_mark(c, st);
}

int mbg_lchain_batch(mbg_ctx_t *c, int32_t n_reads, const mbg_anchor_t *anc, const int64_t *anc_off, const int64_t *n_anc,
                     const mbg_chain_opt_t *opt, uint64_t *u, int32_t *n_u, mbg_anchor_t *out, int64_t *n_out)
{
	if (c == nullptr || n_reads < 0 || opt == nullptr) return MBG_EARG;
	if (n_reads > 0 && (!anc_off || !n_anc || !n_u || !n_out || anc_off[0] != 0)) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_l2b) return MBG_ENOIDX;
	if (n_reads == 0) return MBG_OK;
	const int64_t n_slots = anc_off[n_reads];
	if (n_slots > 0 && (!anc || !u || !out)) return MBG_EARG;
	for (int32_t r = 0; r < n_reads; ++r) {          // offsets, counts, and contigs the resident l2b holds
		if (anc_off[r + 1] < anc_off[r] || n_anc[r] < 0 || n_anc[r] > anc_off[r + 1] - anc_off[r] || n_anc[r] > INT32_MAX)
			return MBG_EARG;
		for (int64_t i = anc_off[r]; i < anc_off[r] + n_anc[r]; ++i)
			if (anc[i].sid < 0 || (uint64_t)(anc[i].sid >> 1) >= c->lb.n_ctg) return MBG_EARG;
	}
	TRY(ctx_enter(c));
	hipStream_t st = c->stream;
	DBuf *b = c->c_buf;
	const size_t wb = MBG_LCHAIN_WORK_BYTES(n_slots);
	TRY(dgrow(b[0], (size_t)n_slots * sizeof(AncRec))); TRY(dgrow(b[1], (size_t)(n_reads + 1) * 8)); TRY(dgrow(b[2], (size_t)n_reads * 8));
	TRY(dgrow(b[3], wb)); TRY(dgrow(b[4], (size_t)n_slots * 8)); TRY(dgrow(b[5], (size_t)n_reads * 4));
	TRY(dgrow(b[6], (size_t)n_slots * sizeof(AncRec))); TRY(dgrow(b[7], (size_t)n_reads * 8));
	if (n_slots > 0) HIPCHK(hipMemcpyAsync(b[0].p, anc, (size_t)n_slots * sizeof(AncRec), hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b[1].p, anc_off, (size_t)(n_reads + 1) * 8, hipMemcpyHostToDevice, st));
	HIPCHK(hipMemcpyAsync(b[2].p, n_anc, (size_t)n_reads * 8, hipMemcpyHostToDevice, st));
	TRY(lchain_dev_locked(c, st, n_reads, n_slots, (const AncRec*)b[0].p, (const int64_t*)b[1].p, (const int64_t*)b[2].p, *opt, b[3].p,
	                      b[3].n, (uint64_t*)b[4].p, (int32_t*)b[5].p, (AncRec*)b[6].p, (int64_t*)b[7].p));
	HIPCHK(hipMemcpyAsync(n_u, b[5].p, (size_t)n_reads * 4, hipMemcpyDeviceToHost, st));
	HIPCHK(hipMemcpyAsync(n_out, b[7].p, (size_t)n_reads * 8, hipMemcpyDeviceToHost, st));
	if (n_slots > 0) {
		HIPCHK(hipMemcpyAsync(u, b[4].p, (size_t)n_slots * 8, hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(out, b[6].p, (size_t)n_slots * sizeof(AncRec), hipMemcpyDeviceToHost, st));
	}
	HIPCHK(hipStreamSynchronize(st));
	return MBG_OK;
}

// ---- the mapper's front pass ---------------------------------------------------------------------

static_assert(sizeof(mbg_chain_opt_t) == sizeof(ChainOpt), "chain option layout");

int mbg_map_front(mbg_ctx_t *c, int32_t n, const uint8_t *seq, const int64_t *off, const int32_t *len, const mbg_front_opt_t *o,
                  mbg_front_t *out)
{
	if (c == nullptr || o == nullptr || out == nullptr || n < 0 || o->min_len < 1 || o->max_occ < 1 || o->n_copt < 0) return MBG_EARG;
	if (n > 0 && (!seq || !off || !len || !o->copt || !o->copt_idx || o->n_copt < 1)) return MBG_EARG;
	for (int32_t r = 0; r < n; ++r)
		if (len[r] < 0 || off[r] < 0 || off[r + 1] < off[r] + len[r] || o->copt_idx[r] < 0 || o->copt_idx[r] >= o->n_copt) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt || !c->has_l2b) return MBG_ENOIDX;
	memset(out, 0, sizeof(*out));
	TRY(ctx_enter(c));
	hipStream_t st = c->stream;
	for (hipEvent_t &e : c->fev) if (!e) HIPCHK(hipEventCreate(&e));
	DBuf *f = c->f_buf, *h = c->h_front;
	const int64_t tot = n > 0 ? off[n] : 0;
	// inputs
	TRY(dgrow(f[0], tot + 16)); TRY(dgrow(f[1], (size_t)(n + 1) * 8)); TRY(dgrow(f[2], (size_t)(n + 1) * 4));
	TRY(dgrow(f[3], (size_t)(n + 1) * 4)); TRY(dgrow(f[4], (size_t)o->n_copt * sizeof(ChainOpt) + 16));
	HIPCHK(hipEventRecord(c->fev[0], st));
	if (n > 0) {
		HIPCHK(hipMemcpyAsync(f[0].p, seq, tot, hipMemcpyHostToDevice, st));
		HIPCHK(hipMemcpyAsync(f[1].p, off, (size_t)(n + 1) * 8, hipMemcpyHostToDevice, st));
		HIPCHK(hipMemcpyAsync(f[2].p, len, (size_t)n * 4, hipMemcpyHostToDevice, st));
		HIPCHK(hipMemcpyAsync(f[3].p, o->copt_idx, (size_t)n * 4, hipMemcpyHostToDevice, st));
		HIPCHK(hipMemcpyAsync(f[4].p, o->copt, (size_t)o->n_copt * sizeof(ChainOpt), hipMemcpyHostToDevice, st));
	}
	// 1. seeding (grows its pools and repeats on overflow)
	int64_t n_rec = 0;
	int64_t *n_rec_dev = nullptr;
	bool ok = false;
	unsigned long long *wstats = nullptr;
	if (c->count_work) {
		TRY(dgrow(c->f_stats, 64));
		wstats = (unsigned long long*)c->f_stats.p;
	}
	hipEvent_t sev[2] = {c->fev[1], c->fev[7]};
	seg_hint(c, len, n);
	for (int attempt = 0; attempt < 4 && !ok; ++attempt) {
		if (wstats) HIPCHK(hipMemsetAsync(wstats, 0, 64, st));
		TRY(seed_core(c, st, n, (const int64_t*)f[1].p, (const int32_t*)f[2].p, (const uint8_t*)f[0].p, o->min_len, o->max_sub_occ,
		              wstats, &n_rec_dev, sev, c->fev[6], false, tot));
		HIPCHK(hipStreamSynchronize(st));
		ok = !seed_overflowed(c, &n_rec);
	}
	if (!ok) return MBG_EOVERFLOW;
	uint64_t work1[2] = {0, 0};
	if (wstats) {
		unsigned long long w[6];
		HIPCHK(hipMemcpy(w, wstats, sizeof(w), hipMemcpyDeviceToHost));
		work1[0] = w[0], work1[1] = w[1];
		c->cnt.seed1_occ_blocks += w[0], c->cnt.seed1_kmer_lookups += w[1];
	}
	// 2. group the records by read; anchor slots per read from the raw records
	TRY(dgrow(f[5], (size_t)(n + 1) * 4)); TRY(dgrow(f[6], (size_t)(n + 1) * 8)); TRY(dgrow(f[7], (size_t)(n + 1) * 8));
	TRY(dgrow(f[8], (size_t)(n + 1) * 8)); TRY(dgrow(f[9], (size_t)(n + 1) * 4)); TRY(dgrow(f[10], (size_t)n_rec * sizeof(Sai) + 16));
	uint32_t *cnt = (uint32_t*)f[5].p, *cursor = (uint32_t*)f[9].p;
	unsigned long long *slots = (unsigned long long*)f[6].p;
	int64_t *uoff = (int64_t*)f[7].p, *aoff = (int64_t*)f[8].p;
	Sai *intv = (Sai*)f[10].p;
	size_t t1 = 0, t2 = 0;
	HIPCHK(scan_u32_i64(nullptr, &t1, nullptr, nullptr, n + 1, st));
	HIPCHK(scan_u64_i64(nullptr, &t2, nullptr, nullptr, n + 1, st));
	TRY(dgrow(c->s_tmp, t1 > t2 ? t1 : t2));
	size_t tb = c->s_tmp.n;
	HIPCHK(hipMemsetAsync(cnt, 0, (size_t)(n + 1) * 4, st));
	HIPCHK(hipMemsetAsync(slots, 0, (size_t)(n + 1) * 8, st));
	HIPCHK(hipMemsetAsync(cursor, 0, (size_t)(n + 1) * 4, st));
	const unsigned gg = persistent_grid(c, n_rec, 256);
	if (n_rec > 0) HIPCHK(launch_grp_count((const SeedRec*)c->p_recs.p, n_rec_dev, c->cap_rec, o->max_occ, cnt, slots, gg, st));
	HIPCHK(scan_u32_i64(c->s_tmp.p, &tb, cnt, uoff, n + 1, st));
	tb = c->s_tmp.n;
	HIPCHK(scan_u64_i64(c->s_tmp.p, &tb, slots, aoff, n + 1, st));
	if (n_rec > 0) HIPCHK(launch_grp_scatter((const SeedRec*)c->p_recs.p, n_rec_dev, c->cap_rec, uoff, cursor, intv, gg, st));
	int64_t n_slots = 0;
	HIPCHK(hipEventRecord(c->fev[2], st));
	HIPCHK(hipMemcpyAsync(&n_slots, aoff + n, 8, hipMemcpyDeviceToHost, st));
	HIPCHK(hipStreamSynchronize(st));
	// 3. anchors (mb_anchor), 4. chains (mb_lchain_dp) with each read's options
	TRY(dgrow(f[11], (size_t)n * 8 + 8)); TRY(dgrow(f[12], (size_t)n * 8 + 8)); TRY(dgrow(f[13], (size_t)n_slots * 8 + 16));
	TRY(dgrow(f[14], (size_t)n_slots * sizeof(AncRec) + 16));
	int64_t *nu = (int64_t*)f[11].p, *nanc = (int64_t*)f[12].p;
	AncRec *anc = (AncRec*)f[14].p;
	if (n > 0)
		TRY(anchor_dev_locked(c, st, n, n_slots, intv, uoff, (const int32_t*)f[2].p, o->max_occ, (uint64_t*)f[13].p, anc, aoff, nanc, nu));
	HIPCHK(hipEventRecord(c->fev[3], st));
	const size_t wb = MBG_LCHAIN_WORK_BYTES(n_slots);
	TRY(dgrow(f[15], wb)); TRY(dgrow(f[16], (size_t)n_slots * 8 + 16)); TRY(dgrow(f[17], (size_t)n * 4 + 8));
	TRY(dgrow(f[18], (size_t)n_slots * sizeof(AncRec) + 16)); TRY(dgrow(f[19], (size_t)n * 8 + 8));
	uint64_t *u = (uint64_t*)f[16].p;
	int32_t *nuc = (int32_t*)f[17].p;
	AncRec *chn = (AncRec*)f[18].p;
	int64_t *nchn = (int64_t*)f[19].p;
	if (n > 0) {
		mbg_chain_opt_t none{};
		TRY(lchain_dev_locked(c, st, n, n_slots, anc, aoff, nanc, none, f[15].p, f[15].n, u, nuc, chn, nchn, (const ChainOpt*)f[4].p,
		                      (const int32_t*)f[3].p));
	}
	HIPCHK(hipEventRecord(c->fev[4], st));
	// 5. results to pinned host memory
	TRY(hgrow(h[0], (size_t)(n + 1) * 8)); TRY(hgrow(h[1], (size_t)n * 8 + 8)); TRY(hgrow(h[2], (size_t)n_rec * sizeof(Sai) + 16));
	TRY(hgrow(h[3], (size_t)(n + 1) * 8)); TRY(hgrow(h[4], (size_t)n * 8 + 8)); TRY(hgrow(h[5], (size_t)n_slots * sizeof(AncRec) + 16));
	TRY(hgrow(h[6], (size_t)n * 4 + 8)); TRY(hgrow(h[7], (size_t)n_slots * 8 + 16)); TRY(hgrow(h[8], (size_t)n * 8 + 8));
	TRY(hgrow(h[9], (size_t)n_slots * sizeof(AncRec) + 16));
	if (n > 0) {
		HIPCHK(hipMemcpyAsync(h[0].p, uoff, (size_t)(n + 1) * 8, hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(h[1].p, nu, (size_t)n * 8, hipMemcpyDeviceToHost, st));
		if (n_rec > 0) HIPCHK(hipMemcpyAsync(h[2].p, intv, (size_t)n_rec * sizeof(Sai), hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(h[3].p, aoff, (size_t)(n + 1) * 8, hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(h[4].p, nanc, (size_t)n * 8, hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(h[6].p, nuc, (size_t)n * 4, hipMemcpyDeviceToHost, st));
		HIPCHK(hipMemcpyAsync(h[8].p, nchn, (size_t)n * 8, hipMemcpyDeviceToHost, st));
		if (n_slots > 0) {
			HIPCHK(hipMemcpyAsync(h[5].p, anc, (size_t)n_slots * sizeof(AncRec), hipMemcpyDeviceToHost, st));
			HIPCHK(hipMemcpyAsync(h[7].p, u, (size_t)n_slots * 8, hipMemcpyDeviceToHost, st));
			HIPCHK(hipMemcpyAsync(h[9].p, chn, (size_t)n_slots * sizeof(AncRec), hipMemcpyDeviceToHost, st));
		}
	} else ((int64_t*)h[0].p)[0] = ((int64_t*)h[3].p)[0] = 0;
	HIPCHK(hipEventRecord(c->fev[5], st));
	HIPCHK(hipStreamSynchronize(st));
	int64_t *hn_anc = (int64_t*)h[4].p;
	for (int32_t r = 0; r < n; ++r) {
		if (hn_anc[r] < 0) return MBG_EOVERFLOW;         // slots from the raw records always suffice
		if (o->no_chain && o->no_chain[r]) hn_anc[r] = -1, ((int32_t*)h[6].p)[r] = 0, ((int64_t*)h[8].p)[r] = 0;
	}
	out->n_rec = n_rec, out->n_slots = n_slots;
	out->intv_off = (const int64_t*)h[0].p, out->n_intv = (const int64_t*)h[1].p, out->intv = (const mbg_sai_t*)h[2].p;
	out->anc_off = (const int64_t*)h[3].p, out->n_anc = hn_anc, out->anc = (const mbg_anchor_t*)h[5].p;
	out->n_u = (const int32_t*)h[6].p, out->u = (const uint64_t*)h[7].p;
	out->n_chn = (const int64_t*)h[8].p, out->chn = (const mbg_anchor_t*)h[9].p;
	float ms;
	HIPCHK(hipEventElapsedTime(&ms, c->fev[6], c->fev[1])); c->cnt.ms_seed1 += ms; out->ms_seed1 = ms;
	HIPCHK(hipEventElapsedTime(&ms, c->fev[1], c->fev[7])); c->cnt.ms_seed2 += ms; out->ms_seed2 = ms;
	out->seed1_occ_blocks = work1[0], out->seed1_kmer_lookups = work1[1];
	HIPCHK(hipEventElapsedTime(&ms, c->fev[0], c->fev[7])); out->ms_seed = ms;
	HIPCHK(hipEventElapsedTime(&ms, c->fev[7], c->fev[2])); out->ms_group = ms; c->cnt.ms_group += ms;
	HIPCHK(hipEventElapsedTime(&ms, c->fev[2], c->fev[3])); out->ms_anchor = ms; c->cnt.ms_anchor += ms;
	HIPCHK(hipEventElapsedTime(&ms, c->fev[3], c->fev[4])); out->ms_chain = ms; c->cnt.ms_chain += ms;
	HIPCHK(hipEventElapsedTime(&ms, c->fev[4], c->fev[5])); out->ms_copy = ms; c->cnt.ms_front_copy += ms;
	HIPCHK(hipEventElapsedTime(&ms, c->fev[0], c->fev[5])); out->ms_total = ms; c->cnt.ms_front += ms;
	++c->cnt.fronts, c->cnt.front_reads += n, c->cnt.front_records += n_rec, c->cnt.front_slots += n_slots;
	return MBG_OK;
}

int mbg_ll_batch(mbg_ctx_t *c, int32_t n_jobs, const mbg_ll_job_t *jobs, const uint8_t *qseq, int64_t qbytes, const uint8_t *tseq,
                 int64_t tbytes, int8_t m, const int8_t *mat, int32_t gapo, int32_t gape, int32_t xtra, mbg_ll_result_t *res)
{
	if (c == nullptr || n_jobs < 0 || (n_jobs > 0 && (!jobs || !res))) return MBG_EARG;
	LlArgs a{};
	int8_t sc[40];
	int mat_min;
	TRY(ll_params(m, mat, gapo, gape, xtra, a, sc, &mat_min));
	const int shift = a.shift, qmax = a.qmax;
	std::lock_guard<std::mutex> lk(c->mu);
	if (n_jobs == 0) return MBG_OK;
	TRY(ctx_enter(c));
	// jobs grouped by registers per lane; each gets its own span of the per-row maxima list
	std::vector<int32_t> order(n_jobs);
	std::vector<int64_t> boff(n_jobs);
	int64_t cnt[17] = {0}, nb = 0;
	std::vector<int8_t> slot(n_jobs);
	for (int32_t i = 0; i < n_jobs; ++i) {
		const mbg_ll_job_t &j = jobs[i];
		if (j.qlen < 0 || j.tlen < 0 || j.qoff < 0 || j.toff < 0 || j.qoff + j.qlen > qbytes || j.toff + j.tlen > tbytes) return MBG_EARG;
		const int k = ll_slots(j.qlen, j.size);
		if (k < 0) return MBG_EUNSUPPORTED;
		if (j.size <= 1) {
			// u8 scores must fit the biased byte (padding scores -1 + shift)
			if (mat_min >= 0 || qmax + shift > 255) return MBG_EUNSUPPORTED;
		} else if ((int64_t)qmax * (j.qlen < j.tlen ? j.qlen : j.tlen) > 32767) {
			return MBG_EUNSUPPORTED;               // the i16 core would saturate
		}
		slot[i] = (int8_t)k, ++cnt[k];
		boff[i] = nb, nb += j.tlen;
	}
	int64_t first[17], at[17];
	for (int k = 0, s = 0; k <= 16; ++k) first[k] = at[k] = s, s += (int)cnt[k];
	for (int32_t i = 0; i < n_jobs; ++i) order[at[slot[i]]++] = i;
	TRY(dgrow(c->b_in[0], qbytes + 16)); TRY(dgrow(c->b_in[1], tbytes + 16)); TRY(dgrow(c->b_in[4], (size_t)n_jobs * sizeof(LlJob)));
	TRY(dgrow(c->b_in[2], (size_t)n_jobs * 4)); TRY(dgrow(c->b_in[3], (size_t)n_jobs * 8 + 64));
	TRY(dgrow(c->b_out[0], (size_t)n_jobs * sizeof(LlRes))); TRY(dgrow(c->b_out[1], (size_t)(nb + 1) * 8));
	HIPCHK(hipMemcpyAsync(c->b_in[0].p, qseq, qbytes, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[1].p, tseq, tbytes, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[4].p, jobs, (size_t)n_jobs * sizeof(LlJob), hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[2].p, order.data(), (size_t)n_jobs * 4, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[3].p, boff.data(), (size_t)n_jobs * 8, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync((char*)c->b_in[3].p + (size_t)n_jobs * 8, sc, sizeof(sc), hipMemcpyHostToDevice, c->stream));
	a.jobs = (const LlJob*)c->b_in[4].p;
	a.qseq = (const uint8_t*)c->b_in[0].p, a.tseq = (const uint8_t*)c->b_in[1].p;
	a.res = (LlRes*)c->b_out[0].p;
	a.bbuf = (uint64_t*)c->b_out[1].p, a.boff = (const int64_t*)c->b_in[3].p;
	a.sc = (const int8_t*)c->b_in[3].p + (size_t)n_jobs * 8;
	a.rows_cap = nb, a.check = false;
	for (hipEvent_t &e : c->dev_ev) if (!e) HIPCHK(hipEventCreate(&e));
	HIPCHK(hipEventRecord(c->dev_ev[0], c->stream));
	for (int k = 1; k <= 16; ++k) {
		if (cnt[k] == 0) continue;
		a.order = (const int32_t*)c->b_in[2].p + first[k];
		a.n = (int32_t)cnt[k];
		HIPCHK(launch_ll(a, k, c->stream));
	}
	HIPCHK(hipEventRecord(c->dev_ev[1], c->stream));
	HIPCHK(hipMemcpyAsync(res, c->b_out[0].p, (size_t)n_jobs * sizeof(LlRes), hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipStreamSynchronize(c->stream));
	{
		float ms = 0;
		HIPCHK(hipEventElapsedTime(&ms, c->dev_ev[0], c->dev_ev[1]));
		c->cnt.ms_ll += ms, ++c->cnt.ll_launches, c->cnt.ll_jobs += n_jobs;
		for (int32_t i = 0; i < n_jobs; ++i) c->cnt.ll_cells += (int64_t)jobs[i].qlen * jobs[i].tlen;
	}
	return MBG_OK;
}

int mbg_matesw_batch(mbg_ctx_t *c, int32_t n, const mbg_matesw_job_t *jobs, const uint8_t *qseq, int64_t qbytes,
                     const mbg_matesw_opt_t *opt, mbg_matesw_result_t *res, uint32_t *cigar, int64_t cigar_cap,
                     mbg_matesw_stats_t *stats)
{
	if (c == nullptr || n < 0 || opt == nullptr || (n > 0 && (!jobs || !qseq || !res || !cigar))) return MBG_EARG;
	const mbg_matesw_opt_t &o = *opt;
	if (o.a <= 0 || o.mt < 0 || o.mt > 2) return MBG_EARG;
	// mb_matesw_align's ksw_ll parameters (pe.c:253-266): scores divided by a, rounded up
	const int b_mm = (o.b + o.a - 1) / o.a, gapo = (o.q + o.a - 1) / o.a, gape = (o.e + o.a - 1) / o.a;
	int8_t mat_ll[25], mat_dp[25];
	gen_nt4_mat(mat_ll, 1, b_mm, (o.b_ts + o.a - 1) / o.a, (o.b_ambi + o.a - 1) / o.a, o.mt);
	gen_nt4_mat(mat_dp, o.a, o.b, o.b_ts, o.b_ambi, o.mt);
	LlArgs la{};
	int8_t sc[40];
	int mat_min;
	TRY(ll_params(5, mat_ll, gapo, gape, 0x40000 | o.min_len, la, sc, &mat_min));
	if (mat_min >= 0 || la.qmax + la.shift > 255) return MBG_EUNSUPPORTED;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_l2b) return MBG_ENOIDX;
	if (n == 0) return MBG_OK;
	TRY(ctx_enter(c));
	int32_t max_q = 1, max_w = 1;
	std::vector<int64_t> tid(n), st(n), en(n), woff(n);
	int64_t wtot = 0;
	for (int32_t i = 0; i < n; ++i) {
		const mbg_matesw_job_t &j = jobs[i];
		const int64_t ts = j.ts < 0 ? 0 : j.ts, w = j.te > ts ? j.te - ts : 0;   // an empty window is no candidate
		if (j.qlen < 0 || j.qoff < 0 || j.qoff + j.qlen > qbytes || j.tid < 0 || (uint64_t)j.tid >= c->lb.n_ctg) return MBG_EARG;
		if (j.cigar_off < 0 || j.cigar_off + j.qlen + w > cigar_cap) return MBG_EARG;
		if (j.qlen > kLlMaxQlen || w > kMsMaxWindow) return MBG_EUNSUPPORTED;
		// ts + w, not te: an end at or before the start is an empty window here, where getseq_kernel
		// would clip a negative end to the contig end as l2b_getseq does
		tid[i] = j.tid, st[i] = ts, en[i] = ts + w, woff[i] = wtot, wtot += w;
		max_q = j.qlen > max_q ? j.qlen : max_q;
		max_w = w > max_w ? (int32_t)w : max_w;
	}
	int tab = 128;
	while (tab < 2 * max_q) tab <<= 1;
	// device buffers: 0 jobs, 1 reads, 2 windows, 3 tid/st/en/woff/wlen, 4 reversed reads, 5 reversed
	// windows, 6 ksw_ll jobs, 7 ksw_ll results, 8 extension jobs, 9 extension results, 10 results,
	// 11 CIGARs
	DBuf *m = c->m_buf;
	TRY(dgrow(m[0], (size_t)n * sizeof(MsJob))); TRY(dgrow(m[1], qbytes + 16)); TRY(dgrow(m[2], wtot + 16));
	TRY(dgrow(m[3], (size_t)n * 40)); TRY(dgrow(m[4], qbytes + 16)); TRY(dgrow(m[5], wtot + 16));
	TRY(dgrow(m[6], (size_t)n * sizeof(LlJob))); TRY(dgrow(m[7], (size_t)n * sizeof(LlRes)));
	TRY(dgrow(m[8], (size_t)n * sizeof(DpJob))); TRY(dgrow(m[9], (size_t)n * sizeof(DpRes)));
	TRY(dgrow(m[10], (size_t)n * sizeof(MsRes))); TRY(dgrow(m[11], (size_t)cigar_cap * 4));
	int64_t *d_tid = (int64_t*)m[3].p, *d_st = d_tid + n, *d_en = d_st + n, *d_woff = d_en + n, *d_wlen = d_woff + n;
	hipStream_t s = c->stream;
	HIPCHK(hipMemcpyAsync(m[0].p, jobs, (size_t)n * sizeof(MsJob), hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(m[1].p, qseq, qbytes, hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(d_tid, tid.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(d_st, st.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(d_en, en.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(d_woff, woff.data(), (size_t)n * 8, hipMemcpyHostToDevice, s));
	for (hipEvent_t &e : c->ev) if (!e) HIPCHK(hipEventCreate(&e));
	HIPCHK(hipEventRecord(c->ev[0], s));
	HIPCHK(launch_getseq(c->lb, n, d_tid, d_st, d_en, d_woff, (uint8_t*)m[2].p, d_wlen, s));
	HIPCHK(hipEventRecord(c->ev[1], s));
	MsArgs a{};
	a.jobs = (const MsJob*)m[0].p, a.n = n, a.qseq = (const uint8_t*)m[1].p, a.win = (const uint8_t*)m[2].p;
	a.woff = d_woff, a.wlen = d_wlen, a.rq = (uint8_t*)m[4].p, a.rt = (uint8_t*)m[5].p;
	a.ll_jobs = (LlJob*)m[6].p, a.ll_res = (const LlRes*)m[7].p, a.dp_jobs = (DpJob*)m[8].p, a.dp_res = (const DpRes*)m[9].p;
	a.res = (MsRes*)m[10].p, a.tab_size = tab;
	a.mt = o.mt, a.a = o.a, a.b_mm = b_mm, a.q = o.q, a.e = o.e, a.bw = o.bw, a.zdrop = o.zdrop, a.end_bonus = o.end_bonus;
	a.min_dp_max = o.min_dp_max;
	TRY(dgrow(c->r_cnt, (size_t)n * 4 + 16)); TRY(dgrow(c->r_off, (size_t)n * 4 + 16));
	TRY(dgrow(c->b_in[5], (size_t)n * sizeof(ExtraJob)));
	a.cigar = (const uint32_t*)m[11].p, a.ccnt = (uint32_t*)c->r_cnt.p, a.extra = (ExtraJob*)c->b_in[5].p;
	HIPCHK(launch_ungap(a, max_w, s));
	HIPCHK(hipEventRecord(c->ev[2], s));
	// ksw_ll over the (narrowed) windows: per-job spans of the per-row maxima list from a scan
	size_t tmp_bytes = 0;
	HIPCHK(scan_ll_rows(nullptr, &tmp_bytes, nullptr, nullptr, n, s));
	TRY(dgrow(c->s_tmp, tmp_bytes + 16));
	TRY(dgrow(c->b_in[3], (size_t)n * 8 + 64));
	TRY(dgrow(c->b_out[1], (size_t)(wtot + 1) * 8));
	HIPCHK(hipMemcpyAsync((char*)c->b_in[3].p + (size_t)n * 8, sc, sizeof(sc), hipMemcpyHostToDevice, s));
	HIPCHK(scan_ll_rows(c->s_tmp.p, &tmp_bytes, (const LlJob*)m[6].p, (int64_t*)c->b_in[3].p, n, s));
	la.jobs = (const LlJob*)m[6].p, la.qseq = (const uint8_t*)m[1].p, la.tseq = (const uint8_t*)m[2].p, la.res = (LlRes*)m[7].p;
	la.order = nullptr, la.n = n, la.bbuf = (uint64_t*)c->b_out[1].p, la.boff = (const int64_t*)c->b_in[3].p;
	la.sc = (const int8_t*)c->b_in[3].p + (size_t)n * 8, la.rows_cap = wtot, la.check = true;
	HIPCHK(launch_ll(la, ll_slots(max_q, 1), s));
	HIPCHK(hipEventRecord(c->ev[3], s));
	// the extension DP of the rescued candidates (pe.c:274-287)
	HIPCHK(launch_rescue_dp(a, s));
	int cap, qt, tcap;
	int64_t slab;
	TRY(dp_plan(max_q, max_w, o.bw, &cap, &slab, &qt, &tcap));
	TRY(dp_dev_locked(c, s, false, n, (const DpJob*)m[8].p, (const uint8_t*)m[4].p, (const uint8_t*)m[5].p, 5, mat_dp,
	                  (DpRes*)m[9].p, (uint32_t*)m[11].p, cap, slab, qt, tcap));
	HIPCHK(hipEventRecord(c->ev[4], s));
	HIPCHK(launch_rescue_finish(a, s));
	// mb_fix_cigar + mb_update_extra on the rescued hits (pe.c:309), CIGARs edited in their slots
	{
		ExtraArgs x{};
		x.jobs = a.extra, x.n = n, x.qseq = (const uint8_t*)m[1].p, x.tseq = (const uint8_t*)m[2].p, x.cigar = (uint32_t*)m[11].p;
		for (int k = 0; k < 28; ++k) x.mat[k] = k < 25 ? mat_dp[k] : 0;
		x.q = o.q, x.e = o.e, x.log_gap = 0;
		HIPCHK(launch_extra(x, s));
		HIPCHK(launch_rescue_final(a, s));
	}
	// CIGARs packed at the prefix sums of their lengths, so that only their operations cross PCIe
	size_t scan_bytes = 0;
	HIPCHK(scan_u32(nullptr, &scan_bytes, nullptr, nullptr, n, s));
	TRY(dgrow(c->s_tmp, scan_bytes + 16));
	HIPCHK(scan_u32(c->s_tmp.p, &scan_bytes, a.ccnt, (uint32_t*)c->r_off.p, n, s));
	TRY(dgrow(c->b_out[2], (size_t)cigar_cap * 4));
	HIPCHK(launch_rescue_cigar(a, (const uint32_t*)c->r_off.p, (uint32_t*)c->b_out[2].p, s));
	HIPCHK(hipEventRecord(c->ev[5], s));
	HIPCHK(hipMemcpyAsync(res, m[10].p, (size_t)n * sizeof(MsRes), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	// the CIGAR brought back: the fixed one of a rescued hit, else the extension DP's
	auto n_back = [&](int32_t i) -> int32_t {
		const mbg_matesw_result_t &r = res[i];
		return r.rescued ? r.n_cigar : r.ran_dp && r.ez.n_cigar > 0 ? r.ez.n_cigar : 0;
	};
	int64_t n_ops = 0;
	for (int32_t i = 0; i < n; ++i) n_ops += n_back(i);
	std::vector<uint32_t> packed((size_t)n_ops + 1);
	if (n_ops > 0) HIPCHK(hipMemcpy(packed.data(), c->b_out[2].p, (size_t)n_ops * 4, hipMemcpyDeviceToHost));
	int64_t at = 0;
	for (int32_t i = 0; i < n; ++i) {
		const int32_t k = n_back(i);
		if (k <= 0) continue;
		memcpy(cigar + jobs[i].cigar_off, packed.data() + at, (size_t)k * 4);
		at += k;
	}
	if (stats) {
		memset(stats, 0, sizeof(*stats));
		for (int32_t i = 0; i < n; ++i) {
			const mbg_matesw_result_t &r = res[i];
			stats->n_aligned += r.aligned, stats->n_rescued += r.rescued;
			if (r.ll.score >= 0) stats->ll_cells += (int64_t)jobs[i].qlen * (r.te2 - r.ts2);
			if (r.ran_dp) stats->dp_cells += r.ez.cells;
		}
		float *ms[] = {&stats->ms_window, &stats->ms_ungap, &stats->ms_ll, &stats->ms_dp, &stats->ms_finish};
		for (int k = 0; k < 5; ++k) HIPCHK(hipEventElapsedTime(ms[k], c->ev[k], c->ev[k + 1]));
		HIPCHK(hipEventElapsedTime(&stats->ms_total, c->ev[0], c->ev[5]));
	}
	return MBG_OK;
}

int mbg_ll_batch_dev(mbg_ctx_t *c, void *stream, int32_t n_jobs, const mbg_ll_job_t *jobs, const uint8_t *qseq, const uint8_t *tseq,
                     int8_t m, const int8_t *mat_host, int32_t gapo, int32_t gape, int32_t xtra, mbg_ll_result_t *res,
                     int32_t max_qlen, int64_t rows_total)
{
	if (c == nullptr || n_jobs < 0 || max_qlen < 0 || rows_total < 0 || (n_jobs > 0 && (!jobs || !qseq || !tseq || !res)))
		return MBG_EARG;
	LlArgs a{};
	int8_t sc[40];
	int mat_min;
	TRY(ll_params(m, mat_host, gapo, gape, xtra, a, sc, &mat_min));
	// the widest padding (16 positions a vector) sets the registers per lane for every job
	const int k = ll_slots(max_qlen, 1);
	if (k < 0) return MBG_EUNSUPPORTED;
	if (mat_min >= 0 || a.qmax + a.shift > 255) return MBG_EUNSUPPORTED;   // the u8 core's biased bytes
	std::lock_guard<std::mutex> lk(c->mu);
	if (n_jobs == 0) return MBG_OK;
	TRY(ctx_enter(c));
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	size_t tmp_bytes = 0;
	HIPCHK(scan_ll_rows(nullptr, &tmp_bytes, nullptr, nullptr, n_jobs, st));
	TRY(dgrow(c->s_tmp, tmp_bytes + 16));
	TRY(dgrow(c->b_in[3], (size_t)n_jobs * 8 + 64));
	TRY(dgrow(c->b_out[1], (size_t)(rows_total + 1) * 8));
	HIPCHK(hipMemcpyAsync((char*)c->b_in[3].p + (size_t)n_jobs * 8, sc, sizeof(sc), hipMemcpyHostToDevice, st));
	HIPCHK(scan_ll_rows(c->s_tmp.p, &tmp_bytes, (const LlJob*)jobs, (int64_t*)c->b_in[3].p, n_jobs, st));
	a.jobs = (const LlJob*)jobs, a.qseq = qseq, a.tseq = tseq, a.res = (LlRes*)res;
	a.order = nullptr, a.n = n_jobs;
	a.bbuf = (uint64_t*)c->b_out[1].p, a.boff = (const int64_t*)c->b_in[3].p;
	a.sc = (const int8_t*)c->b_in[3].p + (size_t)n_jobs * 8;
	a.rows_cap = rows_total, a.check = true;
	HIPCHK(launch_ll(a, k, st));
	HIPCHK(hipStreamSynchronize(st));
	return MBG_OK;
}

int mbg_extd2_batch_dev(mbg_ctx_t *c, void *stream, int32_t n_jobs, const mbg_dp_job_t *jobs, const uint8_t *qseq, const uint8_t *tseq,
                        int8_t m, const int8_t *mat_host, mbg_dp_result_t *res, uint32_t *cigar, int32_t max_qlen, int32_t max_tlen, int32_t max_w)
{
	if (c == nullptr || n_jobs < 0 || m < 0 || m > 5 || mat_host == nullptr || max_qlen < 0 || max_tlen < 0) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (n_jobs == 0) return MBG_OK;
	TRY(ctx_enter(c));
	int cap, qt, tcap;
	int64_t slab;
	TRY(dp_plan(max_qlen, max_tlen, max_w, &cap, &slab, &qt, &tcap));
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	TRY(dp_dev_locked(c, st, true, n_jobs, (const DpJob*)jobs, qseq, tseq, m, mat_host, (DpRes*)res, cigar, cap, slab, qt, tcap));
	return ctx_mark(c, st);
}

int mbg_getseq(mbg_ctx_t *c, int64_t n, const int64_t *tid, const int64_t *st, const int64_t *en, const int64_t *out_off, uint8_t *out,
               int64_t out_bytes, int64_t *len)
{
	if (c == nullptr || n < 0 || (n > 0 && (!tid || !st || !en || !out_off || !out || !len))) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_l2b) return MBG_ENOIDX;
	if (n == 0) return MBG_OK;
	TRY(ctx_enter(c));
	for (int64_t i = 0; i < n; ++i) {   // the clipped window must fit (the kernel clips the same way)
		if (tid[i] < 0 || (uint64_t)tid[i] >= c->lb.n_ctg) continue;
		const uint64_t clen = c->h_ctg_len[tid[i]];
		int64_t s = st[i] < 0 ? 0 : st[i], e = (uint64_t)en[i] > clen ? (int64_t)clen : en[i];
		if (e - s > 0 && (out_off[i] < 0 || out_off[i] + (e - s) > out_bytes)) return MBG_EARG;
	}
	TRY(dgrow(c->b_in[0], n * 8)); TRY(dgrow(c->b_in[1], n * 8)); TRY(dgrow(c->b_in[2], n * 8)); TRY(dgrow(c->b_in[3], n * 8));
	TRY(dgrow(c->b_out[0], out_bytes + 16)); TRY(dgrow(c->b_out[1], n * 8));
	HIPCHK(hipMemcpyAsync(c->b_in[0].p, tid, n * 8, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[1].p, st, n * 8, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[2].p, en, n * 8, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_in[3].p, out_off, n * 8, hipMemcpyHostToDevice, c->stream));
	HIPCHK(hipMemcpyAsync(c->b_out[0].p, out, out_bytes, hipMemcpyHostToDevice, c->stream));
	HIPCHK(launch_getseq(c->lb, n, (const int64_t*)c->b_in[0].p, (const int64_t*)c->b_in[1].p, (const int64_t*)c->b_in[2].p,
	                     (const int64_t*)c->b_in[3].p, (uint8_t*)c->b_out[0].p, (int64_t*)c->b_out[1].p, c->stream));
	HIPCHK(hipMemcpyAsync(out, c->b_out[0].p, out_bytes, hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipMemcpyAsync(len, c->b_out[1].p, n * 8, hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipStreamSynchronize(c->stream));
	return MBG_OK;
}

int mbg_work_bytes(int64_t *bytes, int64_t *dev_free, int64_t *dev_total)
{
	if (bytes) *bytes = g_work_bytes.load();
	if (dev_free || dev_total) {
		size_t f = 0, t = 0;
		HIPCHK(hipMemGetInfo(&f, &t));
		if (dev_free) *dev_free = (int64_t)f;
		if (dev_total) *dev_total = (int64_t)t;
	}
	return MBG_OK;
}

int mbg_dev_alloc(mbg_ctx_t *c, size_t bytes, void **ptr)
{
	if (c == nullptr || ptr == nullptr) return MBG_EARG;
	*ptr = nullptr;
	std::lock_guard<std::mutex> lk(c->mu);
	TRY(ctx_enter(c));
	hipError_t e = hipMalloc(ptr, bytes ? bytes : 16);
	if (e != hipSuccess) { g_hip_err = hipGetErrorString(e); *ptr = nullptr; return MBG_ENOMEM; }
	return MBG_OK;
}

int mbg_dev_free(mbg_ctx_t *c, void *ptr)
{
	if (c == nullptr) return MBG_EARG;
	if (ptr == nullptr) return MBG_OK;
	std::lock_guard<std::mutex> lk(c->mu);
	TRY(ctx_enter(c));
	HIPCHK(hipStreamSynchronize(c->stream));
	HIPCHK(hipFree(ptr));
	return MBG_OK;
}

int mbg_copy_to_device(mbg_ctx_t *c, void *dst, const void *src, size_t bytes)
{
	if (c == nullptr || (bytes > 0 && (dst == nullptr || src == nullptr))) return MBG_EARG;
	if (bytes == 0) return MBG_OK;
	std::lock_guard<std::mutex> lk(c->mu);
	TRY(ctx_enter(c));
	HIPCHK(hipStreamSynchronize(c->stream));     // ordered after the context stream's queued work
	HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
	return MBG_OK;
}

int mbg_copy_to_host(mbg_ctx_t *c, void *dst, const void *src, size_t bytes)
{
	if (c == nullptr || (bytes > 0 && (dst == nullptr || src == nullptr))) return MBG_EARG;
	if (bytes == 0) return MBG_OK;
	std::lock_guard<std::mutex> lk(c->mu);
	TRY(ctx_enter(c));
	HIPCHK(hipStreamSynchronize(c->stream));     // ordered after the context stream's queued work
	HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
	return MBG_OK;
}

int mbg_counters(mbg_ctx_t *c, mbg_counters_t *out, int reset)
{
	if (c == nullptr || out == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	*out = c->cnt;
	if (reset) memset(&c->cnt, 0, sizeof(c->cnt));
	return MBG_OK;
}

int mbg_use_graphs(mbg_ctx_t *c, int on)
{
	if (c == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	c->use_graphs = on != 0;
	return MBG_OK;
}

int mbg_count_work(mbg_ctx_t *c, int on)
{
	if (c == nullptr) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	c->count_work = on != 0;
	return MBG_OK;
}

int mbg_host_alloc(mbg_ctx_t *c, size_t bytes, void **ptr)
{
	if (c == nullptr || ptr == nullptr) return MBG_EARG;
	*ptr = nullptr;
	TRY(ctx_enter(c));
	hipError_t e = hipHostMalloc(ptr, bytes ? bytes : 16, 0);
	if (e != hipSuccess) { g_hip_err = std::string("hipHostMalloc: ") + hipGetErrorString(e); *ptr = nullptr; return MBG_ENOMEM; }
	return MBG_OK;
}

int mbg_host_free(mbg_ctx_t *c, void *ptr)
{
	if (c == nullptr) return MBG_EARG;
	if (ptr) HIPCHK(hipHostFree(ptr));
	return MBG_OK;
}

int mbg_stream_copy(mbg_ctx_t *c, void *stream, void *dst, const void *src, size_t bytes)
{
	if (c == nullptr || dst == nullptr || src == nullptr || (bytes & 15) != 0) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	TRY(ctx_enter(c));
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	HIPCHK(launch_stream_copy(dst, src, (int64_t)(bytes / 16), (unsigned)c->n_cu * 16, st));
	HIPCHK(hipStreamSynchronize(st));
	return MBG_OK;
}

int mbg_gather_bench(mbg_ctx_t *c, int64_t n_chains, int32_t steps, int32_t bs, float *ms, uint64_t *checksum)
{
	if (c == nullptr || n_chains <= 0 || steps <= 0 || ms == nullptr || (bs != 64 && bs != 128 && bs != 256 && bs != 512 && bs != 1024)) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt) return MBG_ENOIDX;
	TRY(ctx_enter(c));
	for (hipEvent_t &e : c->ev) if (!e) HIPCHK(hipEventCreate(&e));
	TRY(dgrow(c->b_ctr, 128));
	unsigned long long *sum = (unsigned long long*)c->b_ctr.p + 6;
	HIPCHK(hipMemsetAsync(sum, 0, 8, c->stream));
	HIPCHK(hipEventRecord(c->ev[0], c->stream));
	HIPCHK(launch_gather_chain(c->ix, n_chains, steps, bs, sum, c->stream));
	HIPCHK(hipEventRecord(c->ev[1], c->stream));
	unsigned long long h = 0;
	HIPCHK(hipMemcpyAsync(&h, sum, 8, hipMemcpyDeviceToHost, c->stream));
	HIPCHK(hipStreamSynchronize(c->stream));
	HIPCHK(hipEventElapsedTime(ms, c->ev[0], c->ev[1]));
	if (checksum) *checksum = h;
	return MBG_OK;
}

// ---- seed -> locate -> extend step --------------------------------------------------------------

static_assert(sizeof(mbg_job_meta_t) == sizeof(JobMeta), "job meta layout");
static_assert(sizeof(mbg_map_opt_t) == 12 * 4 + 28, "map opt layout");

int mbg_map_step_dev(mbg_ctx_t *c, void *stream, int32_t n_reads, const int64_t *off, const int32_t *len, const uint8_t *seq,
                     int32_t max_len, const mbg_map_opt_t *opt, int collect_stats, mbg_map_stats_t *stats, mbg_map_out_t *out)
{
	if (c == nullptr || opt == nullptr || stats == nullptr || n_reads < 0 || max_len < 0 || (n_reads > 0 && (!off || !len || !seq)))
		return MBG_EARG;
	if (opt->min_len < 1 || opt->max_occ < 1 || opt->e <= 0 || opt->e2 <= 0) return MBG_EARG;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->has_bwt || !c->has_l2b) return MBG_ENOIDX;
	TRY(ctx_enter(c));
	hipStream_t st = stream ? (hipStream_t)stream : c->stream;
	TRY(ctx_order(c, st));
	for (hipEvent_t &e : c->ev) if (!e) HIPCHK(hipEventCreate(&e));
	memset(stats, 0, sizeof(*stats));
	stats->n_reads = n_reads;
	MapOpt mo{};
	mo.a = opt->a; mo.b = opt->b; mo.q = opt->q; mo.e = opt->e; mo.q2 = opt->q2; mo.e2 = opt->e2; mo.end_bonus = opt->end_bonus;
	mo.ksw_w = opt->ksw_w; mo.zdrop = opt->zdrop;
	for (int i = 0; i < 25; ++i) mo.mat[i] = opt->mat[i];
	const bool dual = !(opt->q == opt->q2 && opt->e == opt->e2);
	// job sizes: query <= max_len, target <= l + extra (align.c:590-595)
	const int64_t l_extra = ((int64_t)max_len * opt->a + opt->end_bonus - opt->q) > 0 ? ((int64_t)max_len * opt->a + opt->end_bonus - opt->q) / opt->e : 0;
	const int32_t max_tlen = (int32_t)(max_len + l_extra + 1);
	int cap_dp, qt_dp, tcap_dp;
	int64_t slab;
	TRY(dp_plan(max_len > 0 ? max_len : 1, max_tlen, opt->ksw_w, &cap_dp, &slab, &qt_dp, &tcap_dp));
	const int64_t job_cap = 2 * (int64_t)n_reads + 16;
	if (c->cap_rows < 16 * (int64_t)n_reads + 1024) c->cap_rows = 16 * (int64_t)n_reads + 1024;
	for (int attempt = 0; attempt < 4; ++attempt) {
		// every pool first: the chain below allocates nothing, so that it can run as one graph
		TRY(dgrow(c->p_best, (size_t)n_reads * 8 + 8));
		TRY(dgrow(c->p_jobs, job_cap * sizeof(DpJob))); TRY(dgrow(c->p_meta, job_cap * sizeof(JobMeta)));
		TRY(dgrow(c->p_qb, (size_t)n_reads * max_len + 16)); TRY(dgrow(c->p_tb, (size_t)job_cap * max_tlen + 16));
		TRY(dgrow(c->p_res, job_cap * sizeof(DpRes)));
		TRY(dgrow(c->p_cig, (size_t)job_cap * (max_len + max_tlen) * 4 + 16));
		TRY(dgrow(c->p_ctr, 256));
		TRY(seed_alloc(c, st, n_reads));
		const int64_t rc = c->cap_rec, rw = c->cap_rows;
		TRY(dgrow(c->p_rows, rw * 8)); TRY(dgrow(c->p_rowrec, rw * 4)); TRY(dgrow(c->p_first, rc * 4));
		TRY(dgrow(c->r_cnt, rc * 4)); TRY(dgrow(c->r_off, rc * 4));
		size_t tmp_bytes = 0;
		HIPCHK(scan_u32(nullptr, &tmp_bytes, nullptr, nullptr, rc, st));
		TRY(dgrow(c->s_tmp, tmp_bytes > c->s_tmp.n ? tmp_bytes : c->s_tmp.n));
		int64_t g = (int64_t)c->n_cu * 32;
		if (g > job_cap) g = job_cap;
		const int64_t max_by_mem = ((int64_t)2 << 30) / slab;
		if (g > max_by_mem) g = max_by_mem;
		if (g < 1) g = 1;
		TRY(dgrow(c->b_slab, (size_t)g * slab));
		DpArgs d{};
		TRY(dp_state(c, g, cap_dp, tcap_dp, d));
		TRY(dgrow(c->b_defer, (size_t)job_cap * 4));
		unsigned long long *pc = (unsigned long long*)c->p_ctr.p;  // [0..5] job sink, [6] rows, [8..14] work stats, [15] jobs past the pools
		const bool timed = !c->use_graphs;                          // phase events only when the launches go one by one
		auto chain = [&](hipStream_t s) -> int {
			HIPCHK(zero_async(pc, 256, s));
			HIPCHK(zero_async(c->p_best.p, (size_t)n_reads * 8 + 8, s));
			// seeding: two launches of the persistent state machine, then compaction
			int64_t *n_rec = nullptr;
			TRY(seed_core(c, s, n_reads, off, len, seq, opt->min_len, opt->max_sub_occ, collect_stats ? pc + 8 : nullptr, &n_rec,
			              timed ? c->ev + 1 : nullptr, nullptr, true));
			size_t tb = c->s_tmp.n;
			// anchor rows at scanned offsets, SA locate
			const SeedRec *recs = (const SeedRec*)c->p_recs.p;
			HIPCHK(launch_row_counts(recs, n_rec, rc, opt->max_occ, (uint32_t*)c->r_cnt.p, s));
			HIPCHK(scan_u32(c->s_tmp.p, &tb, (const uint32_t*)c->r_cnt.p, (uint32_t*)c->r_off.p, rc, s));
			HIPCHK(launch_anchor_rows(recs, n_rec, rc, opt->max_occ, (const uint32_t*)c->r_cnt.p, (const uint32_t*)c->r_off.p,
			                          (uint64_t*)c->p_rows.p, (uint32_t*)c->p_rowrec.p, (uint32_t*)c->p_first.p, pc + 6, rw, s));
			HIPCHK(launch_best_seed(recs, n_rec, rc, (unsigned long long*)c->p_best.p, (unsigned)c->n_cu * 8, s));
			if (timed) HIPCHK(hipEventRecord(c->ev[3], s));
			uint32_t *ctr = (uint32_t*)c->b_ctr.p;
			HIPCHK(zero_async(ctr + 4, 4, s));
			HIPCHK(launch_sa(c->ix, rw, (const int64_t*)(pc + 6), (uint64_t*)c->p_rows.p, ctr + 4, collect_stats ? pc + 14 : nullptr,
			                 persistent_grid(c, 4 * (int64_t)n_reads + 1, 256), s));
			if (timed) HIPCHK(hipEventRecord(c->ev[4], s));
			// extension jobs and DP
			HIPCHK(launch_ext_jobs(c->lb, mo, n_reads, off, len, seq, recs, (const uint32_t*)c->p_first.p, (const uint64_t*)c->p_rows.p,
			                       (const unsigned long long*)c->p_best.p, (DpJob*)c->p_jobs.p, (JobMeta*)c->p_meta.p, (uint8_t*)c->p_qb.p,
			                       (uint8_t*)c->p_tb.p, job_cap, (int64_t)n_reads * max_len, job_cap * max_tlen, pc, s));
			if (timed) HIPCHK(hipEventRecord(c->ev[5], s));
			d.jobs = (const DpJob*)c->p_jobs.p; d.n_jobs = (int32_t)job_cap; d.n_jobs_dev = pc;
			d.qseq = (const uint8_t*)c->p_qb.p; d.tseq = (const uint8_t*)c->p_tb.p;
			d.res = (DpRes*)c->p_res.p; d.cigar = (uint32_t*)c->p_cig.p;
			d.pbuf = (uint8_t*)c->b_slab.p; d.slab = slab; d.qt_cap = qt_dp; d.m = 5;
			d.next = (unsigned int*)c->b_ctr.p + 14;
			HIPCHK(zero_async(d.next, 20, s));
			d.defer = (int32_t*)c->b_defer.p;
			dp_set_matrix(d, 5, opt->mat);
			HIPCHK(launch_dp(dual, d, (unsigned)g, s));
			return MBG_OK;
		};
		HIPCHK(hipEventRecord(c->ev[0], st));
		TRY(run_graph(c, st, c->seg_len > 0 && n_reads > 0 ? kGraphStepSeg : kGraphStep, chain));
		HIPCHK(hipEventRecord(c->ev[6], st));
		unsigned long long hc[16];
		HIPCHK(hipMemcpyAsync(hc, pc, sizeof(hc), hipMemcpyDeviceToHost, st));
		HIPCHK(hipStreamSynchronize(st));
		int64_t hn_rec = 0;
		bool ov = seed_overflowed(c, &hn_rec);
		if ((int64_t)hc[6] > rw) c->cap_rows = (int64_t)hc[6] * 2, ov = true;
		if (ov) continue;  // a pool overflowed: it has grown; run the step again
		stats->n_rec = hn_rec; stats->n_rows = (int64_t)hc[6]; stats->n_jobs = (int64_t)hc[0]; stats->n_fast = (int64_t)hc[3];
		stats->n_anchored = (int64_t)hc[4]; stats->q_bytes = (int64_t)hc[1]; stats->t_bytes = (int64_t)hc[2];
		stats->occ_blocks[0] = hc[8]; stats->kmer_lookups[0] = hc[9]; stats->extend_steps[0] = hc[10];
		stats->occ_blocks[1] = hc[11]; stats->kmer_lookups[1] = hc[12]; stats->extend_steps[1] = hc[13]; stats->lf_steps = hc[14];
		float ms;
		if (timed) {              // the phases: only when the chain ran launch by launch (graphs off, the default)
			HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[1])); stats->ms_seed1 = ms;
			HIPCHK(hipEventElapsedTime(&ms, c->ev[1], c->ev[2])); stats->ms_seed2 = ms;
			HIPCHK(hipEventElapsedTime(&ms, c->ev[2], c->ev[3])); stats->ms_rows = ms;
			HIPCHK(hipEventElapsedTime(&ms, c->ev[3], c->ev[4])); stats->ms_sa = ms;
			HIPCHK(hipEventElapsedTime(&ms, c->ev[4], c->ev[5])); stats->ms_jobs = ms;
			HIPCHK(hipEventElapsedTime(&ms, c->ev[5], c->ev[6])); stats->ms_dp = ms;
		}
		HIPCHK(hipEventElapsedTime(&ms, c->ev[0], c->ev[6])); stats->ms_total = ms;
		if (collect_stats && stats->n_jobs > 0) {
			std::vector<DpRes> res(stats->n_jobs);
			HIPCHK(hipMemcpy(res.data(), c->p_res.p, stats->n_jobs * sizeof(DpRes), hipMemcpyDeviceToHost));
			for (const DpRes &r : res) stats->cells += r.cells;
		}
		if (out) {
			out->recs = (const mbg_seed_rec_t*)c->p_recs.p; out->pos = (const uint64_t*)c->p_rows.p; out->pos_rec = (const uint32_t*)c->p_rowrec.p;
			out->jobs = (const mbg_dp_job_t*)c->p_jobs.p; out->meta = (const mbg_job_meta_t*)c->p_meta.p; out->res = (const mbg_dp_result_t*)c->p_res.p;
			out->cigar = (const uint32_t*)c->p_cig.p; out->qbuf = (const uint8_t*)c->p_qb.p; out->tbuf = (const uint8_t*)c->p_tb.p;
		}
		return hc[15] > 0 ? MBG_EOVERFLOW : MBG_OK;   // jobs of reads longer than max_len ran empty
	}
	return MBG_EOVERFLOW;
}

} // extern "C"
