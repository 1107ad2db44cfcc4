// mbg_io.cpp -- the FASTA/FASTQ reader in front of the device path: the records and chunks of
// mb_bseq_read / mb_bseq_read_frag (bseq.c:41-161) with kseq's parsing rules (kseq.h:102-238),
// written into flat arrays instead of one allocation per field, with the bases also encoded to nt4
// (kom_nt4_table, kommon.c:162-179) for upload.
//
//   read-ahead thread   gzread (plain files pass through zlib unchanged) into 1 MiB blocks, up to
//                       kAhead blocks ahead of the parser, so decompression overlaps parsing
//   parser              kseq's state machine over the blocks: a header starts at '>' or '@'; the name
//                       ends at any isspace() byte and the comment is the rest of that line; sequence
//                       lines follow (blank lines skipped) until a line starts with '>', '@' or '+';
//                       after '+', quality lines are read until the quality is as long as the
//                       sequence; a field read up to a line end loses one trailing '\r'
//   encoder             U/u -> T/t and nt4 codes over the chunk's records, split over threads
//
// Host code only; no GPU call.
#include <zlib.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "mbg.h"

namespace {

constexpr size_t kBlock = 1 << 20;
constexpr int kAhead = 4;
constexpr uint64_t kPairCheck = 1000000;      // frag mode pairs reads shorter than this (bseq.c:29)
constexpr uint64_t kKseqBuf = 16384;          // kseq's read size (kseq.h:242): it decides one end-of-input case

const uint8_t *nt4_table()
{
	static uint8_t t[256];
	static std::once_flag once;
	std::call_once(once, [] {
		for (int i = 0; i < 256; ++i) t[i] = i < 4 ? (uint8_t)i : 4;   // already-encoded bytes map to themselves
		const char *b = "ACGT";
		for (int k = 0; k < 4; ++k) t[(uint8_t)b[k]] = t[(uint8_t)(b[k] | 0x20)] = (uint8_t)k;
	});
	return t;
}

// blocks of the decompressed stream, produced by a reader thread
class Source {
public:
	explicit Source(gzFile f) : f_(f), th_([this] { run(); }) {}
	~Source()
	{
		{
			std::lock_guard<std::mutex> lk(mu_);
			stop_ = true;
		}
		cv_.notify_all();
		th_.join();
		gzclose(f_);
	}
	// the next block; empty at the end of the stream
	std::vector<char> next()
	{
		std::unique_lock<std::mutex> lk(mu_);
		cv_.wait(lk, [this] { return !q_.empty() || done_; });
		if (q_.empty()) return {};
		std::vector<char> b = std::move(q_.front());
		q_.pop_front();
		cv_.notify_all();
		return b;
	}
	bool failed() const { return err_; }

private:
	void run()
	{
		for (;;) {
			std::vector<char> b(kBlock);
			const int n = gzread(f_, b.data(), (unsigned)kBlock);
			std::unique_lock<std::mutex> lk(mu_);
			if (n < 0) err_ = true;
			if (n > 0) {
				b.resize((size_t)n);
				cv_.wait(lk, [this] { return q_.size() < (size_t)kAhead || stop_; });
				if (stop_) break;
				q_.push_back(std::move(b));
				cv_.notify_all();
			}
			if (n <= 0 || stop_) break;
		}
		std::lock_guard<std::mutex> lk(mu_);
		done_ = true;
		cv_.notify_all();
	}
	gzFile f_;
	std::mutex mu_;
	std::condition_variable cv_;
	std::deque<std::vector<char>> q_;
	bool done_ = false, stop_ = false, err_ = false;
	std::thread th_;
};

// kseq's byte stream over the blocks: getc, and reads up to a delimiter.  kseq notices the end of its
// input when a read returns fewer than 16 KiB; an input whose length is a multiple of that is noticed
// one read later, so the first delimiter read that starts exactly there returns the field rather than
// -1 (kseq.h:75-85,102-150).  end_seen() restates that.
class Stream {
public:
	explicit Stream(Source *s) : src_(s) {}
	int getc()
	{
		if (at_end()) { tried_ = true; return -1; }
		return (uint8_t)cur_[pos_++];
	}
	// bytes up to the delimiter (consumed; *dret = it, or 0 at the end) appended to the field
	// v[st ..); sep_line: '\n', and one trailing '\r' of the whole field dropped; else any isspace()
	// byte.  Returns -1 when the stream was already at its end (the field untouched apart from a reset
	// when !append), else the field's length.
	int64_t until(bool sep_line, std::vector<char> &v, size_t st, int *dret, bool append)
	{
		if (dret) *dret = 0;
		if (!append) v.resize(st);
		if (at_end()) {
			if (end_seen()) return -1;
			tried_ = true;
		}
		for (;;) {
			if (at_end()) { tried_ = true; break; }
			const char *b = cur_.data() + pos_, *e = cur_.data() + cur_.size(), *hit = nullptr;
			if (sep_line) hit = (const char*)memchr(b, '\n', (size_t)(e - b));
			else
				for (const char *p = b; p < e; ++p)
					if (*p == ' ' || (*p >= '\t' && *p <= '\r')) { hit = p; break; }
			const char *stop = hit ? hit : e;
			v.insert(v.end(), b, stop);
			pos_ = (size_t)(stop - cur_.data());
			if (hit) {
				if (dret) *dret = (uint8_t)*hit;
				++pos_;
				break;
			}
		}
		if (sep_line && v.size() - st > 1 && v.back() == '\r') v.pop_back();
		return (int64_t)(v.size() - st);
	}

	// the unread bytes of the current block (fetching the next block when it is used up)
	bool span(const char **b, const char **e)
	{
		if (at_end()) return false;
		*b = cur_.data() + pos_, *e = cur_.data() + cur_.size();
		return true;
	}
	void skip(size_t n) { pos_ += n; }

private:
	bool at_end()
	{
		if (pos_ < cur_.size()) return false;
		if (eof_) return true;
		cur_ = src_->next();
		pos_ = 0;
		if (cur_.empty()) eof_ = true;
		total_ += cur_.size();
		return eof_;
	}
	bool end_seen() const { return total_ % kKseqBuf != 0 || tried_; }
	Source *src_;
	std::vector<char> cur_;
	size_t pos_ = 0;
	uint64_t total_ = 0;                        // bytes of the input so far (all of it once eof_)
	bool eof_ = false, tried_ = false;
};

// the four fields of records, each appended at the end of its own array
struct Rec {
	std::vector<char> name, comment, seq, qual;
	void clear() { name.clear(), comment.clear(), seq.clear(), qual.clear(); }
};

// one record as kseq_read returns it, its fields appended to r's arrays (from their sizes at the
// call): >= 0 the sequence length, -1 end of input, -2 a quality string missing or of another length
class Parser {
public:
	explicit Parser(Stream *s) : s_(s) {}
	int64_t read(Rec &r)
	{
		int c;
		const size_t n0 = r.name.size(), c0 = r.comment.size(), s0 = r.seq.size(), q0 = r.qual.size();
		if (last_ == 0) {                                          // find the next header
			while ((c = s_->getc()) != -1 && c != '>' && c != '@') {}
			if (c == -1) return -1;
			last_ = c;
		}
		const int64_t fl = fast(r);
		if (fl >= 0) return fl;
		if (s_->until(false, r.name, n0, &c, false) < 0) return -1;
		if (c != '\n') s_->until(true, r.comment, c0, nullptr, false);
		while ((c = s_->getc()) != -1 && c != '>' && c != '+' && c != '@') {
			if (c == '\n') continue;
			r.seq.push_back((char)c);
			s_->until(true, r.seq, s0, nullptr, true);
		}
		const size_t sl = r.seq.size() - s0;
		if (c == '>' || c == '@') last_ = c;
		if (c != '+') return (int64_t)sl;
		while ((c = s_->getc()) != -1 && c != '\n') {}
		if (c == -1) return -2;
		while (s_->until(true, r.qual, q0, nullptr, true) >= 0 && r.qual.size() - q0 < sl) {}
		last_ = 0;
		if (sl != r.qual.size() - q0) return -2;
		return (int64_t)sl;
	}

private:
	// A FASTQ record whose four lines lie whole in the current block, with a one-line sequence and a
	// one-line quality of the same length and no '\r' at a line end, read at memchr speed.  For such a
	// record the general path below takes the same fields and leaves the stream at the same byte with
	// the same state; anything else (returns -1, nothing consumed) goes to the general path.
	int64_t fast(Rec &r)
	{
		const char *b, *e;
		if (!s_->span(&b, &e)) return -1;
		const char *l1 = (const char*)memchr(b, '\n', (size_t)(e - b));
		if (!l1) return -1;
		const char *sq = l1 + 1;
		if (sq >= e || *sq == '>' || *sq == '+' || *sq == '@' || *sq == '\n') return -1;
		const char *l2 = (const char*)memchr(sq, '\n', (size_t)(e - sq));
		if (!l2 || l2[-1] == '\r' || l2 + 1 >= e || l2[1] != '+') return -1;
		const char *l3 = (const char*)memchr(l2 + 1, '\n', (size_t)(e - l2 - 1));
		if (!l3) return -1;
		const char *qs = l3 + 1;
		const size_t sl = (size_t)(l2 - sq);
		if ((size_t)(e - qs) <= sl || qs[sl] != '\n' || (sl > 0 && qs[sl - 1] == '\r')) return -1;
		if (memchr(qs, '\n', sl)) return -1;                     // the quality is one line
		const char *nm = b;
		while (nm < l1 && !(*nm == ' ' || (*nm >= '\t' && *nm <= '\r'))) ++nm;
		if (nm < l1 && *nm != '\n' && l1[-1] == '\r') return -1;  // a comment ending in '\r'
		if (nm == l1 && l1 > b && l1[-1] == '\r') return -1;
		r.name.insert(r.name.end(), b, nm);
		if (nm < l1) r.comment.insert(r.comment.end(), nm + 1, l1);
		r.seq.insert(r.seq.end(), sq, l2);
		r.qual.insert(r.qual.end(), qs, qs + sl);
		s_->skip((size_t)(qs + sl + 1 - b));
		last_ = 0;
		return (int64_t)sl;
	}
	Stream *s_;
	int last_ = 0;
};

bool qname_same(const char *a, size_t la, const char *b, size_t lb)   // mb_qname_same (bseq.h:26-39)
{
	auto len = [](const char *s, size_t l) {
		return l >= 3 && s[l - 1] >= '0' && s[l - 1] <= '9' && s[l - 2] == '/' ? l - 2 : l;
	};
	la = len(a, la), lb = len(b, lb);
	return la == lb && memcmp(a, b, la) == 0;
}

} // namespace

struct mbg_fq_s {
	Source *src = nullptr;
	Stream *st = nullptr;
	Parser *ps = nullptr;
	int32_t n_threads = 1;
	bool have_pending = false;                  // the record frag mode read past the chunk (bseq.c:113)
	int pend_q = 0, pend_c = 0;                 // ... converted with the flags of the call that read it
	Rec pending;
};

// A chunk: committed records, then possibly one record parsed past them (its fields at the ends of
// the arrays).  qual runs parallel to seq: a record without a kept quality holds zeros there.
struct mbg_fq_buf_s {
	std::vector<int64_t> name_off, com_off, seq_off;
	Rec f;
	std::vector<uint8_t> has_qual, nt4;
	void clear()
	{
		name_off.assign(1, 0), com_off.assign(1, 0), seq_off.assign(1, 0);
		f.clear(), has_qual.clear(), nt4.clear();
	}
	int32_t n() const { return (int32_t)has_qual.size(); }
	// kseq2bseq (bseq.c:64-78) of the record parsed past the committed ones: the quality kept when
	// asked for and non-empty, the comment likewise
	void commit(int with_qual, int with_comment)
	{
		const size_t s0 = (size_t)seq_off.back(), sl = f.seq.size() - s0;
		if (!(with_comment && f.comment.size() > (size_t)com_off.back())) f.comment.resize((size_t)com_off.back());
		const bool q = with_qual && f.qual.size() > s0;
		if (!q) f.qual.assign(f.qual.begin(), f.qual.begin() + (ptrdiff_t)s0), f.qual.resize(s0 + sl, 0);
		has_qual.push_back(q ? 1 : 0);
		name_off.push_back((int64_t)f.name.size()), com_off.push_back((int64_t)f.comment.size());
		seq_off.push_back((int64_t)f.seq.size());
	}
	void drop()                                 // forget the record parsed past the committed ones
	{
		f.name.resize((size_t)name_off.back()), f.comment.resize((size_t)com_off.back());
		f.seq.resize((size_t)seq_off.back()), f.qual.resize((size_t)seq_off.back());
	}
	void take(Rec &r) const                     // a copy of the record parsed past the committed ones
	{
		r.name.assign(f.name.begin() + name_off.back(), f.name.end());
		r.comment.assign(f.comment.begin() + com_off.back(), f.comment.end());
		r.seq.assign(f.seq.begin() + seq_off.back(), f.seq.end());
		r.qual.assign(f.qual.begin() + seq_off.back(), f.qual.end());
	}
	void put(const Rec &r)                      // r as the record parsed past the committed ones
	{
		f.name.insert(f.name.end(), r.name.begin(), r.name.end());
		f.comment.insert(f.comment.end(), r.comment.begin(), r.comment.end());
		f.seq.insert(f.seq.end(), r.seq.begin(), r.seq.end());
		f.qual.insert(f.qual.end(), r.qual.begin(), r.qual.end());
	}
};

namespace {

// U -> T and the nt4 codes of every record, split over threads
void encode(mbg_fq_buf_s &b, int n_threads)
{
	const uint8_t *t = nt4_table();
	const int64_t n = (int64_t)b.f.seq.size();
	b.nt4.resize((size_t)n);
	auto work = [&](int64_t lo, int64_t hi) {
		for (int64_t i = lo; i < hi; ++i) {
			char &c = b.f.seq[(size_t)i];
			if (c == 'u' || c == 'U') --c;
			b.nt4[(size_t)i] = t[(uint8_t)c];
		}
	};
	const int nt = n < (1 << 20) ? 1 : std::max(1, n_threads);
	if (nt == 1) { work(0, n); return; }
	std::vector<std::thread> th;
	for (int k = 0; k < nt; ++k) th.emplace_back(work, n * k / nt, n * (k + 1) / nt);
	for (auto &x : th) x.join();
}

void view(const mbg_fq_buf_s &b, int32_t status, mbg_fq_chunk_t *out)
{
	out->n = b.n();
	out->status = status;
	out->name_off = b.name_off.data(), out->com_off = b.com_off.data(), out->seq_off = b.seq_off.data();
	out->name = b.f.name.data(), out->comment = b.f.comment.data(), out->seq = b.f.seq.data(), out->qual = b.f.qual.data();
	out->has_qual = b.has_qual.data(), out->nt4 = b.nt4.data();
}

} // namespace

extern "C" {

int mbg_fq_open(const char *path, int32_t n_threads, mbg_fq_t **out)
{
	if (out == nullptr) return MBG_EARG;
	*out = nullptr;
	gzFile f = path && strcmp(path, "-") != 0 ? gzopen(path, "r") : gzdopen(0, "r");
	if (f == nullptr) return MBG_EIO;
	gzbuffer(f, 1 << 20);
	mbg_fq_s *q = new mbg_fq_s;
	q->src = new Source(f);
	q->st = new Stream(q->src);
	q->ps = new Parser(q->st);
	q->n_threads = n_threads < 1 ? 1 : n_threads > 64 ? 64 : n_threads;
	*out = q;
	return MBG_OK;
}

void mbg_fq_close(mbg_fq_t *q)
{
	if (q == nullptr) return;
	delete q->ps;
	delete q->st;
	delete q->src;
	delete q;
}

int mbg_fq_buf_new(mbg_fq_buf_t **out)
{
	if (out == nullptr) return MBG_EARG;
	*out = new mbg_fq_buf_s;
	(*out)->clear();
	return MBG_OK;
}

void mbg_fq_buf_free(mbg_fq_buf_t *b) { delete b; }

int mbg_fq_read(mbg_fq_t *q, int64_t chunk_size, int32_t with_qual, int32_t with_comment, int32_t frag_mode, int32_t min_cnt,
                int64_t max_chunk_size, mbg_fq_buf_t *b, mbg_fq_chunk_t *out)
{
	if (q == nullptr || b == nullptr || out == nullptr) return MBG_EARG;
	b->clear();
	int64_t size = 0, ret;
	if (q->have_pending) {
		b->put(q->pending);
		b->commit(q->pend_q, q->pend_c);
		size = (int64_t)q->pending.seq.size();
		q->have_pending = false;
	}
	if (max_chunk_size < chunk_size) max_chunk_size = chunk_size;
	while ((ret = q->ps->read(b->f)) >= 0) {
		b->commit(with_qual, with_comment);
		size += ret;
		bool stop;
		if (chunk_size <= 0 || max_chunk_size <= 0) stop = true;
		else if (size >= max_chunk_size) stop = true;
		else stop = size >= chunk_size && b->n() >= min_cnt;
		if (stop) {
			if (frag_mode && (uint64_t)ret < kPairCheck) {
				while ((ret = q->ps->read(b->f)) >= 0) {
					size += ret;
					const int64_t *no = b->name_off.data();
					const int32_t k = b->n();
					const char *nm = b->f.name.data();
					if (qname_same(nm + no[k], b->f.name.size() - (size_t)no[k], nm + no[k - 1], (size_t)(no[k] - no[k - 1]))) {
						b->commit(with_qual, with_comment);
					} else {
						b->take(q->pending);
						b->drop();
						q->have_pending = true, q->pend_q = with_qual, q->pend_c = with_comment;
						break;
					}
				}
			}
			break;
		}
	}
	b->drop();                                                      // a record cut by an error or the end
	encode(*b, q->n_threads);
	view(*b, ret < -1 ? (int32_t)ret : (q->src->failed() ? MBG_EIO : 0), out);
	return MBG_OK;
}

int mbg_fq_read_frag(int32_t n_fp, mbg_fq_t **q, int64_t chunk_size, int32_t with_qual, int32_t with_comment, mbg_fq_buf_t *b,
                     mbg_fq_chunk_t *out)
{
	if (n_fp < 1 || q == nullptr || b == nullptr || out == nullptr) return MBG_EARG;
	for (int32_t i = 0; i < n_fp; ++i) if (q[i] == nullptr) return MBG_EARG;
	b->clear();
	int64_t size = 0;
	std::vector<Rec> recs((size_t)n_fp);
	for (;;) {                                                      // bseq.c:135-158: one record of each file per step
		int n_read = 0;
		for (int32_t i = 0; i < n_fp; ++i) {
			recs[(size_t)i].clear();
			if (q[i]->ps->read(recs[(size_t)i]) >= 0) ++n_read;
		}
		if (n_read < n_fp) break;
		for (int32_t i = 0; i < n_fp; ++i) {
			b->put(recs[(size_t)i]);
			b->commit(with_qual, with_comment);
			size += (int64_t)recs[(size_t)i].seq.size();
		}
		if (size >= chunk_size) break;
	}
	encode(*b, q[0]->n_threads);
	view(*b, 0, out);
	return MBG_OK;
}

} // extern "C"
