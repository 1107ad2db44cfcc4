// mbg_io.cpp -- the FASTA/FASTQ reader in front of the device path: the records and chunks of
// mb_bseq_read / mb_bseq_read_frag (bseq.c:41-161) with kseq's parsing rules (kseq.h:102-238),
// written into flat arrays instead of one allocation per field, with the bases also encoded to nt4
// (kom_nt4_table, kommon.c:162-179) for upload.
//
//   read-ahead thread   gzread into 1 MiB blocks, up to kAhead blocks ahead of the parser, so
//                       decompression overlaps parsing.  An uncompressed regular file is mapped
//                       instead and read in place (like any mapping, it must not shrink while read)
//   parse-ahead thread  kseq's state machine over the blocks, one thread per file, into batches of
//                       parsed records kept up to kAheadBytes ahead of the caller; a record whose
//                       four lines lie whole in one block is kept as pointers into that block (the
//                       batch holds the block), anything else is copied out: a header starts at
//                       '>' or '@'; the name ends at any isspace() byte and the comment is the rest of
//                       that line; sequence lines follow (blank lines skipped) until a line starts with
//                       '>', '@' or '+'; after '+', quality lines are read until the quality is as
//                       long as the sequence; a field read up to a line end loses one trailing '\r'
//   chunking            the caller's thread takes parsed records one at a time under bseq.c's rules
//                       (only lengths and names are looked at), so that what a chunk holds and where
//                       the next one starts are exactly the serial reader's
//   assembly            the chunk's records copied into its flat arrays, with U/u -> T/t and the nt4
//                       codes, split over threads by bytes
//
// Parsing a file never depends on the chunking parameters of the calls that consume it, which is what
// lets it run ahead: the records, in file order, are kseq_read's whatever the chunks are.
//
// Host code only; no GPU call.
#include <fcntl.h>
#include <pthread.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "mbg.h"

namespace {

constexpr size_t kBlock = 1 << 20;
constexpr int kAhead = 4;
constexpr size_t kBatchRecs = 16384;           // parsed records per batch
constexpr size_t kAheadBytes = (size_t)768 << 20;   // parsed bytes a file may run ahead of its caller
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

// an allocator whose resize() leaves new elements uninitialised: the chunk arrays are sized first and
// then written in parallel, so zero-filling them would be a serial pass over every byte
template <class T>
struct NoInit : std::allocator<T> {
	template <class U> struct rebind { using other = NoInit<U>; };
	NoInit() = default;
	template <class U> NoInit(const NoInit<U> &) {}
	template <class U> void construct(U *p) noexcept { ::new ((void*)p) U; }
	template <class U, class... A> void construct(U *p, A &&...a) { ::new ((void*)p) U(std::forward<A>(a)...); }
};
using Bytes = std::vector<char, NoInit<char>>;

// a block of the input: bytes p[0 .. n), kept alive by `keep` (the block's own buffer, or the
// mapping of a whole uncompressed file)
struct Block {
	std::shared_ptr<const char> keep;
	const char *p = nullptr;
	size_t n = 0;
	const char *data() const { return p; }
	size_t size() const { return n; }
	bool empty() const { return n == 0; }
	char operator[](size_t i) const { return p[i]; }
};

// blocks of the decompressed stream, produced by a reader thread; an uncompressed regular file is
// mapped instead, and its blocks are views of the mapping (no copy, no thread)
class Source {
public:
	explicit Source(gzFile f) : f_(f), th_([this] { run(); }) {}
	Source(std::shared_ptr<const char> map, size_t size) : map_(std::move(map)), map_size_(size) {}
	~Source()
	{
		if (!f_) return;
		{
			std::lock_guard<std::mutex> lk(mu_);
			stop_ = true;
		}
		cv_.notify_all();
		th_.join();
		gzclose(f_);
	}
	// the next block; empty at the end of the stream
	Block next()
	{
		if (!f_) {
			Block b;
			if (map_off_ < map_size_) {
				b.p = map_.get() + map_off_, b.n = std::min(kBlock, map_size_ - map_off_), b.keep = map_;
				map_off_ += b.n;
			}
			return b;
		}
		std::unique_lock<std::mutex> lk(mu_);
		cv_.wait(lk, [this] { return !q_.empty() || done_; });
		if (q_.empty()) return {};
		auto own = std::make_shared<Bytes>(std::move(q_.front()));
		q_.pop_front();
		cv_.notify_all();
		Block b;
		b.p = own->data(), b.n = own->size(), b.keep = std::shared_ptr<const char>(own, own->data());
		return b;
	}
	bool failed() const { return err_; }

private:
	void run()
	{
		pthread_setname_np(pthread_self(), "fq-read");
		for (;;) {
			Bytes b(kBlock);   // not zero-filled: gzread writes it
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
	gzFile f_ = nullptr;
	std::shared_ptr<const char> map_;
	size_t map_size_ = 0, map_off_ = 0;
	std::mutex mu_;
	std::condition_variable cv_;
	std::deque<Bytes> q_;
	bool done_ = false, stop_ = false, err_ = false;
	std::thread th_;
};

// an uncompressed regular file mapped whole, or null (then the file goes through zlib)
std::shared_ptr<const char> map_plain(const char *path, size_t *size)
{
	const int fd = open(path, O_RDONLY);
	if (fd < 0) return nullptr;
	struct stat sb;
	unsigned char magic[2] = {0, 0};
	void *m = MAP_FAILED;
	if (fstat(fd, &sb) == 0 && S_ISREG(sb.st_mode) && sb.st_size > 0 && pread(fd, magic, 2, 0) == 2 &&
	    !(magic[0] == 0x1f && magic[1] == 0x8b))
		m = mmap(nullptr, (size_t)sb.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
	close(fd);
	if (m == MAP_FAILED) return nullptr;
	madvise(m, (size_t)sb.st_size, MADV_SEQUENTIAL);
	*size = (size_t)sb.st_size;
	const size_t len = *size;
	return std::shared_ptr<const char>((const char*)m, [len](const char *q) { munmap((void*)q, len); });
}

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
		return (uint8_t)blk_[pos_++];
	}
	// bytes up to the delimiter (consumed; *dret = it, or 0 at the end) appended to the field
	// v[st ..); sep_line: '\n', and one trailing '\r' of the whole field dropped; else any isspace()
	// byte.  Returns -1 when the stream was already at its end (the field untouched apart from a reset
	// when !append), else the field's length.
	int64_t until(bool sep_line, Bytes &v, size_t st, int *dret, bool append)
	{
		if (dret) *dret = 0;
		if (!append) v.resize(st);
		if (at_end()) {
			if (end_seen()) return -1;
			tried_ = true;
		}
		for (;;) {
			if (at_end()) { tried_ = true; break; }
			const char *b = blk_.data() + pos_, *e = blk_.data() + blk_.size(), *hit = nullptr;
			if (sep_line) hit = (const char*)memchr(b, '\n', (size_t)(e - b));
			else
				for (const char *p = b; p < e; ++p)
					if (*p == ' ' || (*p >= '\t' && *p <= '\r')) { hit = p; break; }
			const char *stop = hit ? hit : e;
			v.insert(v.end(), b, stop);
			pos_ = (size_t)(stop - blk_.data());
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
		*b = blk_.data() + pos_, *e = blk_.data() + blk_.size();
		return true;
	}
	void skip(size_t n) { pos_ += n; }
	// what keeps the block span() points into alive (records of the fast path hold it)
	const std::shared_ptr<const char> &block() const { return blk_.keep; }

private:
	bool at_end()
	{
		if (pos_ < blk_.size()) return false;
		if (eof_) return true;
		blk_ = src_->next();
		pos_ = 0;
		if (blk_.empty()) eof_ = true;
		total_ += blk_.size();
		return eof_;
	}
	bool end_seen() const { return total_ % kKseqBuf != 0 || tried_; }
	Source *src_;
	Block blk_;
	size_t pos_ = 0;
	uint64_t total_ = 0;                        // bytes of the input so far (all of it once eof_)
	bool eof_ = false, tried_ = false;
};


// the four fields of records, each appended at the end of its own array
struct Rec {
	Bytes name, comment, seq, qual;
	void clear() { name.clear(), comment.clear(), seq.clear(), qual.clear(); }
};

// parsed records of one file, in file order.  Record i's four fields (name, comment, sequence,
// quality) are fld[i].p[k][0 .. fld[i].l[k]): inside one of `blocks` for a record the fast path found
// whole in a block (no copy), else in f, where the general path appended them (offsets into f until
// resolve() turns them into pointers, once the batch is complete and f no longer moves).  ret[i] is
// what kseq_read returned for the record (>= 0 its length, -1 the end, -2 a bad quality; the fields
// of a -2 record are what was read of it).
struct Fld {
	const char *p[4];
	int64_t l[4];
};
struct Parsed {
	Rec f;
	std::vector<std::shared_ptr<const char>> blocks;
	std::vector<Fld> fld;
	std::vector<int64_t> ret;
	std::vector<size_t> in_f;      // records whose fields are in f
	size_t field_bytes = 0;
	size_t n() const { return ret.size(); }
	size_t bytes() const { return field_bytes + 40 * ret.size(); }
	void resolve()
	{
		const char *base[4] = {f.name.data(), f.comment.data(), f.seq.data(), f.qual.data()};
		for (size_t i : in_f)
			for (int k = 0; k < 4; ++k) fld[i].p[k] = base[k] + (uintptr_t)fld[i].p[k];
	}
};

// one record as kseq_read returns it, appended to b: >= 0 the sequence length, -1 end of input, -2 a
// quality string missing or of another length
class Parser {
public:
	explicit Parser(Stream *s) : s_(s) {}
	int64_t read(Parsed &b)
	{
		Fld fd;
		Rec &r = b.f;
		const size_t o[4] = {r.name.size(), r.comment.size(), r.seq.size(), r.qual.size()};
		int64_t ret = -1;
		int c = 0;
		if (last_ == 0) {                                          // find the next header
			while ((c = s_->getc()) != -1 && c != '>' && c != '@') {}
			if (c != -1) last_ = c;
		}
		if (c != -1) {
			ret = fast(b, fd);
			if (ret >= 0) {
				b.fld.push_back(fd);
				return ret;
			}
			ret = general(r);
		}
		const size_t e[4] = {r.name.size(), r.comment.size(), r.seq.size(), r.qual.size()};
		for (int k = 0; k < 4; ++k) fd.p[k] = (const char*)(uintptr_t)o[k], fd.l[k] = (int64_t)(e[k] - o[k]), b.field_bytes += e[k] - o[k];
		b.in_f.push_back(b.fld.size());
		b.fld.push_back(fd);
		return ret;
	}

private:
	int64_t general(Rec &r)
	{
		int c;
		const size_t n0 = r.name.size(), c0 = r.comment.size(), s0 = r.seq.size(), q0 = r.qual.size();
		if (last_ == 0) {                                          // find the next header
			while ((c = s_->getc()) != -1 && c != '>' && c != '@') {}
			if (c == -1) return -1;
			last_ = c;
		}
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

	// A FASTQ record whose four lines lie whole in the current block, with a one-line sequence and a
	// one-line quality of the same length and no '\r' at a line end, read at memchr speed.  For such a
	// record the general path below takes the same fields and leaves the stream at the same byte with
	// the same state; anything else (returns -1, nothing consumed) goes to the general path.
	int64_t fast(Parsed &b_out, Fld &fd)
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
		fd.p[0] = b, fd.l[0] = nm - b;
		fd.p[1] = nm < l1 ? nm + 1 : l1, fd.l[1] = nm < l1 ? l1 - nm - 1 : 0;
		fd.p[2] = sq, fd.l[2] = (int64_t)sl;
		fd.p[3] = qs, fd.l[3] = (int64_t)sl;
		b_out.field_bytes += (size_t)(fd.l[0] + fd.l[1] + 2 * sl);
		if (b_out.blocks.empty() || b_out.blocks.back() != s_->block()) b_out.blocks.push_back(s_->block());
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

// the parse-ahead thread of one file, and the caller's position in what it parsed
class Ahead {
public:
	explicit Ahead(Parser *ps) : ps_(ps), th_([this] { run(); }) {}
	~Ahead()
	{
		{
			std::lock_guard<std::mutex> lk(mu_);
			stop_ = true;
		}
		cv_.notify_all();
		th_.join();
	}
	// the next record (after the end: the end again)
	void next(std::shared_ptr<Parsed> *b, size_t *i)
	{
		if (!cur_ || pos_ >= cur_->n()) {
			if (cur_ && cur_->ret.back() == -1) {           // the end record, every time from now on
				*b = cur_, *i = cur_->n() - 1;
				return;
			}
			std::unique_lock<std::mutex> lk(mu_);
			cv_.wait(lk, [this] { return !q_.empty(); });
			cur_ = std::move(q_.front());
			q_.pop_front();
			bytes_ -= cur_->bytes();
			cv_.notify_all();
			pos_ = 0;
		}
		*b = cur_, *i = pos_++;
	}

private:
	void run()
	{
		pthread_setname_np(pthread_self(), "fq-parse");
		for (bool end = false; !end;) {
			auto b = std::make_shared<Parsed>();
			while (b->n() < kBatchRecs) {
				const int64_t r = ps_->read(*b);
				b->ret.push_back(r);
				if (r == -1) { end = true; break; }
			}
			b->resolve();
			std::unique_lock<std::mutex> lk(mu_);
			cv_.wait(lk, [this] { return bytes_ < kAheadBytes || stop_; });
			if (stop_) return;
			bytes_ += b->bytes();
			q_.push_back(std::move(b));
			cv_.notify_all();
		}
	}
	Parser *ps_;
	std::mutex mu_;
	std::condition_variable cv_;
	std::deque<std::shared_ptr<Parsed>> q_;
	size_t bytes_ = 0;
	bool stop_ = false;
	std::shared_ptr<Parsed> cur_;
	size_t pos_ = 0;
	std::thread th_;
};

// one record of a chunk: where it was parsed, and the flags of the call that took it (kseq2bseq,
// bseq.c:64-78: the quality kept when asked for and non-empty, the comment likewise)
struct Take {
	std::shared_ptr<Parsed> b;
	size_t i;
	bool q, c;
	int64_t len() const { return b->ret[i]; }
	const char *name() const { return b->fld[i].p[0]; }
	size_t name_len() const { return (size_t)b->fld[i].l[0]; }
};

} // namespace

struct mbg_fq_s {
	Source *src = nullptr;
	Stream *st = nullptr;
	Parser *ps = nullptr;
	Ahead *ah = nullptr;
	int32_t n_threads = 1;
	bool have_pending = false;                  // the record frag mode took past the chunk (bseq.c:113)
	Take pending;                               // ... kept with the flags of the call that took it
};

// A chunk's flat arrays.  qual runs parallel to seq: a record without a kept quality holds zeros there.
struct mbg_fq_buf_s {
	std::vector<int64_t> name_off, com_off, seq_off;
	Rec f;
	std::vector<uint8_t> has_qual;
	std::vector<uint8_t, NoInit<uint8_t>> nt4;
};

namespace {

// the records of a chunk into b's arrays: offsets first (serial, a few adds a record), then the bytes,
// split over threads by bytes; U/u -> T/t and the nt4 codes on the way
void assemble(mbg_fq_buf_s &b, const std::vector<Take> &t, int n_threads)
{
	const size_t n = t.size();
	b.name_off.resize(n + 1), b.com_off.resize(n + 1), b.seq_off.resize(n + 1), b.has_qual.resize(n);
	b.name_off[0] = b.com_off[0] = b.seq_off[0] = 0;
	for (size_t k = 0; k < n; ++k) {
		const Fld &fd = t[k].b->fld[t[k].i];
		const int64_t sl = fd.l[2], cl = fd.l[1], ql = fd.l[3];
		b.name_off[k + 1] = b.name_off[k] + fd.l[0];
		b.com_off[k + 1] = b.com_off[k] + (t[k].c && cl > 0 ? cl : 0);
		b.seq_off[k + 1] = b.seq_off[k] + sl;
		b.has_qual[k] = t[k].q && ql > 0;
	}
	b.f.name.resize((size_t)b.name_off[n]), b.f.comment.resize((size_t)b.com_off[n]);
	b.f.seq.resize((size_t)b.seq_off[n]), b.f.qual.resize((size_t)b.seq_off[n]), b.nt4.resize((size_t)b.seq_off[n]);
	const uint8_t *tab = nt4_table();
	auto work = [&](size_t lo, size_t hi) {
		for (size_t k = lo; k < hi; ++k) {
			const Fld &fd = t[k].b->fld[t[k].i];
			const int64_t sl = b.seq_off[k + 1] - b.seq_off[k];
			if (b.name_off[k + 1] > b.name_off[k])
				memcpy(&b.f.name[(size_t)b.name_off[k]], fd.p[0], (size_t)(b.name_off[k + 1] - b.name_off[k]));
			if (b.com_off[k + 1] > b.com_off[k])
				memcpy(&b.f.comment[(size_t)b.com_off[k]], fd.p[1], (size_t)(b.com_off[k + 1] - b.com_off[k]));
			char *sq = sl > 0 ? &b.f.seq[(size_t)b.seq_off[k]] : nullptr;
			uint8_t *e = sl > 0 ? &b.nt4[(size_t)b.seq_off[k]] : nullptr;
			const char *src = fd.p[2];
			for (int64_t l = 0; l < sl; ++l) {
				char c = src[l];
				if (c == 'u' || c == 'U') --c;
				sq[l] = c, e[l] = tab[(uint8_t)c];
			}
			if (b.has_qual[k]) memcpy(&b.f.qual[(size_t)b.seq_off[k]], fd.p[3], (size_t)sl);
			else if (sl > 0) memset(&b.f.qual[(size_t)b.seq_off[k]], 0, (size_t)sl);
		}
	};
	const int64_t tot = b.seq_off[n] + b.name_off[n];
	const int nt = tot < (1 << 20) ? 1 : std::max(1, n_threads);
	if (nt == 1) { work(0, n); return; }
	std::vector<std::thread> th;
	size_t lo = 0;
	for (int k = 1; k <= nt && lo < n; ++k) {                      // ranges of about tot/nt bytes
		size_t hi = n;
		if (k < nt) {
			const int64_t goal = tot * k / nt;
			size_t a = lo, z = n;                                  // the first record past goal bytes
			while (a < z) {
				const size_t m = (a + z) / 2;
				if (b.seq_off[m + 1] + b.name_off[m + 1] > goal) z = m; else a = m + 1;
			}
			hi = std::max(lo, a);
		}
		th.emplace_back(work, lo, hi);
		lo = hi;
	}
	for (auto &x : th) x.join();
}

void view(const mbg_fq_buf_s &b, int32_t status, mbg_fq_chunk_t *out)
{
	out->n = (int32_t)b.has_qual.size();
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
	const bool named = path && strcmp(path, "-") != 0;
	size_t map_size = 0;
	std::shared_ptr<const char> map = named ? map_plain(path, &map_size) : nullptr;
	Source *src = nullptr;
	if (map) src = new Source(std::move(map), map_size);
	else {
		gzFile f = named ? gzopen(path, "r") : gzdopen(0, "r");
		if (f == nullptr) return MBG_EIO;
		gzbuffer(f, 1 << 20);
		src = new Source(f);
	}
	mbg_fq_s *q = new mbg_fq_s;
	q->src = src;
	q->st = new Stream(q->src);
	q->ps = new Parser(q->st);
	q->ah = new Ahead(q->ps);
	q->n_threads = n_threads < 1 ? 1 : n_threads > 64 ? 64 : n_threads;
	*out = q;
	return MBG_OK;
}

void mbg_fq_close(mbg_fq_t *q)
{
	if (q == nullptr) return;
	delete q->ah;                                                   // stops the parser before its stream goes
	delete q->ps;
	delete q->st;
	delete q->src;
	delete q;
}

int mbg_fq_buf_new(mbg_fq_buf_t **out)
{
	if (out == nullptr) return MBG_EARG;
	*out = new mbg_fq_buf_s;
	return MBG_OK;
}

void mbg_fq_buf_free(mbg_fq_buf_t *b) { delete b; }

int mbg_fq_read(mbg_fq_t *q, int64_t chunk_size, int32_t with_qual, int32_t with_comment, int32_t frag_mode, int32_t min_cnt,
                int64_t max_chunk_size, mbg_fq_buf_t *b, mbg_fq_chunk_t *out)
{
	if (q == nullptr || b == nullptr || out == nullptr) return MBG_EARG;
	std::vector<Take> t;
	int64_t size = 0, ret = 0;
	if (q->have_pending) {
		t.push_back(std::move(q->pending));
		size = t.back().len();
		q->have_pending = false;
	}
	if (max_chunk_size < chunk_size) max_chunk_size = chunk_size;
	for (;;) {
		Take r{nullptr, 0, with_qual != 0, with_comment != 0};
		q->ah->next(&r.b, &r.i);
		if ((ret = r.len()) < 0) break;
		t.push_back(std::move(r));
		size += ret;
		bool stop;
		if (chunk_size <= 0 || max_chunk_size <= 0) stop = true;
		else if (size >= max_chunk_size) stop = true;
		else stop = size >= chunk_size && (int64_t)t.size() >= min_cnt;
		if (stop) {
			if (frag_mode && (uint64_t)ret < kPairCheck) {
				for (;;) {
					Take r2{nullptr, 0, with_qual != 0, with_comment != 0};
					q->ah->next(&r2.b, &r2.i);
					if ((ret = r2.len()) < 0) break;
					size += ret;
					const Take &l = t.back();
					if (qname_same(r2.name(), r2.name_len(), l.name(), l.name_len())) {
						t.push_back(std::move(r2));
					} else {
						q->pending = std::move(r2);
						q->have_pending = true;
						break;
					}
				}
			}
			break;
		}
	}
	assemble(*b, t, q->n_threads);
	view(*b, ret < -1 ? (int32_t)ret : (q->src->failed() ? MBG_EIO : 0), out);
	return MBG_OK;
}

int mbg_fq_read_frag(int32_t n_fp, mbg_fq_t **q, int64_t chunk_size, int32_t with_qual, int32_t with_comment, mbg_fq_buf_t *b,
                     mbg_fq_chunk_t *out)
{
	if (n_fp < 1 || q == nullptr || b == nullptr || out == nullptr) return MBG_EARG;
	for (int32_t i = 0; i < n_fp; ++i) if (q[i] == nullptr) return MBG_EARG;
	std::vector<Take> t, step((size_t)n_fp);
	int64_t size = 0;
	for (;;) {                                                      // bseq.c:135-158: one record of each file per step
		int n_read = 0;
		for (int32_t i = 0; i < n_fp; ++i) {
			Take &r = step[(size_t)i];
			r.q = with_qual != 0, r.c = with_comment != 0;
			q[i]->ah->next(&r.b, &r.i);
			if (r.len() >= 0) ++n_read;
		}
		if (n_read < n_fp) break;
		for (int32_t i = 0; i < n_fp; ++i) {
			size += step[(size_t)i].len();
			t.push_back(step[(size_t)i]);
		}
		if (size >= chunk_size) break;
	}
	assemble(*b, t, q[0]->n_threads);
	view(*b, 0, out);
	return MBG_OK;
}

} // extern "C"
