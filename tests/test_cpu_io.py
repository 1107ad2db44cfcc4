"""The library's FASTA/FASTQ reader (mbg_fq_*, host code) against the reference's own mb_bseq_read /
mb_bseq_read_frag (bseq.c:41-161, kseq.h:102-238) compiled in oracle/_ref: every chunk's records (name,
bases with U turned into T, quality, comment) under the same chunking parameters, on seeded reads, on
hand-made inputs with kseq's corner cases and on fuzzed byte streams (sizes on and around kseq's 16 KiB
read size), plain and gzip.  The nt4 codes are checked against kom_nt4_table's definition."""
import gzip

import numpy as np
import pytest

from minibwa_amd import mbg

O = pytest.importorskip("oracle.oracle")

if not O.have_ref():
    pytest.skip("oracle/_ref not built", allow_module_level=True)


def _nt4(seq: bytes) -> np.ndarray:
    """kom_nt4_table (kommon.c:162-179): bytes 0-3 kept, A/C/G/T in either case 0-3, all else 4."""
    t = np.full(256, 4, np.uint8)
    t[:4] = np.arange(4)
    for k, b in enumerate(b"ACGT"):
        t[b] = t[b | 0x20] = k
    return t[np.frombuffer(seq, np.uint8)]


def _compare(path, params_seq, n_threads=3):
    ours, ref = mbg.FastxReader(path, n_threads), O.RefBseq(path)
    total = 0
    for it in range(100000):
        p = params_seq[it % len(params_seq)]
        a, nt4, _ = ours.read(**p)
        b = ref.read(**p)
        assert a == b, (path, it, p, a[:3], b[:3])
        seqs = b"".join(r[1] for r in a)
        assert (nt4 == _nt4(seqs)).all()
        total += len(a)
        if not a:
            break
    ours.close(), ref.close()
    return total


def _write(path, data: bytes, gz=False):
    if gz:
        with gzip.open(path, "wb") as f:
            f.write(data)
    else:
        with open(path, "wb") as f:
            f.write(data)


PARAMS = [dict(chunk_size=1000, min_cnt=0), dict(chunk_size=0), dict(chunk_size=300, min_cnt=5, max_chunk_size=2000),
          dict(chunk_size=450, frag_mode=True), dict(chunk_size=5000, with_qual=False, with_comment=False),
          dict(chunk_size=100, frag_mode=True, min_cnt=3, max_chunk_size=150)]


@pytest.mark.parametrize("gz", [False, True])
def test_seeded_reads(tmp_path, gz):
    from minibwa_amd import synth
    ref = synth.reference(20000, seed=3, rep_frac=0.0)
    r1, r2 = synth.read_pairs(ref, 2000, seed=5)
    f1, f2 = str(tmp_path / "a_1.fq"), str(tmp_path / "a_2.fq")
    synth.write_pairs_fastq(f1, f2, r1, r2, prefix="p")
    # interleaved copy (names end in /1 and /2), for frag mode
    l1, l2 = open(f1, "rb").read().split(b"\n"), open(f2, "rb").read().split(b"\n")
    inter = b"".join(b"\n".join(l1[i:i + 4]) + b"\n" + b"\n".join(l2[i:i + 4]) + b"\n" for i in range(0, len(l1) - 1, 4))
    fi = str(tmp_path / ("inter.fq" + (".gz" if gz else "")))
    _write(fi, inter, gz)
    assert _compare(fi, PARAMS) == 4000
    n = _compare(fi, [dict(chunk_size=10000, frag_mode=True, min_cnt=10)])
    assert n == 4000
    # two files read in step (mb_bseq_read_frag)
    ours = [mbg.FastxReader(f1), mbg.FastxReader(f2)]
    refs = [O.RefBseq(f1), O.RefBseq(f2)]
    got = 0
    while True:
        a, nt4, _ = mbg.FastxReader.read_frag(ours, 7000)
        b = O.RefBseq.read_frag(refs, 7000)
        assert a == b
        got += len(a)
        if not a:
            break
    assert got == 4000


EDGE = [
    b">r1 a comment\twith tab\nACGU\nacgu\n\nNNNN\n>r2\n\n>r3  \nAC\r\nGT\r\n",
    b"@q1/1 c\nACGTU\n+\nIIIII\n@q1/2\nTTTT\n+q1\n@III\n@q2\nA\nC\n+\nI\nI\n",
    b"@bad\nACGT\n+\nII\n@next\nAC\n+\nII\n",                                   # short quality, then recovery
    b"@trunc\nACGT\n+",                                                         # '+' line cut by the end
    b"@trunc2\nACGT\n+\nIII",                                                   # quality cut by the end
    b">crlf\r\nAC\r\n\r\nGT\r\n>x\r\n\r",
    b"junk before\n>a\nAC>notheader\nGG\n",
    b">",
    b"@",
    b">onlyname",
    b"\n\n>a b\n" + b"A" * 40000 + b"\n",
    b"@u\nuuUU\n+\n####\n>fa\nuU\n",
]


@pytest.mark.parametrize("i", range(len(EDGE)))
@pytest.mark.parametrize("gz", [False, True])
def test_edge_inputs(tmp_path, i, gz):
    f = str(tmp_path / f"e{i}.fx")
    _write(f, EDGE[i], gz)
    _compare(f, PARAMS)
    _compare(f, [dict(chunk_size=-1)])


def _fuzz_bytes(rng, n):
    alpha = np.frombuffer(b">@+\n\n\n\r \tACGTUNacgu!I#/12", np.uint8)
    w = np.ones(len(alpha))
    w[3:6] = 6                                     # newlines are common
    w[9:15] = 8
    return bytes(rng.choice(alpha, n, p=w / w.sum()))


@pytest.mark.parametrize("seed", range(6))
def test_fuzz(tmp_path, seed):
    """Random streams of header/quality/newline/CR bytes; sizes on kseq's 16 KiB read size and beside it."""
    rng = np.random.default_rng(seed)
    for k, n in enumerate([16384, 32768, 16383, 16385, int(rng.integers(100, 60000)), 49152]):
        data = _fuzz_bytes(rng, n)
        if k % 2:
            data = data[:-1] + b">"                # a header byte as the very last byte
        f = str(tmp_path / f"fz{seed}_{k}.fx")
        _write(f, data, gz=bool(k % 3 == 2))
        _compare(f, PARAMS)


def test_reader_errors(tmp_path):
    with pytest.raises(mbg.MbgError):
        mbg.FastxReader(str(tmp_path / "missing.fq"))


def _mixed_records(rng, n):
    """Standard 4-line FASTQ records mixed with ones the block-level path must leave to the general
    one: CRLF, multi-line sequence or quality, comments, '+name' lines, '@' leading a quality, empty
    sequences, blank lines, lowercase and U, FASTA records."""
    out = []
    for i in range(n):
        L = int(rng.integers(0, 300)) if rng.random() < 0.1 else int(rng.integers(20, 200))
        seq = bytes(rng.choice(np.frombuffer(b"ACGTNacgtU", np.uint8), L))
        qual = bytes(rng.choice(np.frombuffer(b"!#5?@IJ+>", np.uint8), L))
        kind = rng.integers(0, 10)
        name = b"m%d" % i + (b"/1" if i % 2 == 0 else b"/2")
        if kind <= 4:
            out.append(b"@" + name + b"\n" + seq + b"\n+\n" + qual + b"\n")
        elif kind == 5:
            out.append(b"@" + name + b" c=1\tx\r\n" + seq + b"\r\n+\r\n" + qual + b"\r\n")
        elif kind == 6 and L > 4:
            h = L // 2
            out.append(b"@" + name + b"\n" + seq[:h] + b"\n" + seq[h:] + b"\n+" + name + b"\n" + qual[:h] + b"\n" + qual[h:] + b"\n")
        elif kind == 7:
            out.append(b">" + name + b" fa\n" + seq + b"\n\n")
        elif kind == 8:
            out.append(b"@" + name + b"\n" + seq + b"\n+\n@" + qual[1:] + b"\n" if L else b"@" + name + b"\n\n+\n\n")
        else:
            out.append(b"@" + name + b"\n" + seq + b"\n+\n" + qual + b"\n\n")
    return b"".join(out)


@pytest.mark.parametrize("gz", [False, True])
def test_mixed_records_across_blocks(tmp_path, gz):
    """Several MB of mixed records, so the block-level path and the general path alternate and records
    straddle the reader's 1 MiB block boundaries."""
    rng = np.random.default_rng(17)
    data = _mixed_records(rng, 40000)
    assert len(data) > 3 << 20
    f = str(tmp_path / ("mixed.fq" + (".gz" if gz else "")))
    _write(f, data, gz)
    assert _compare(f, PARAMS) > 39000
    assert _compare(f, [dict(chunk_size=2_000_000, frag_mode=True, min_cnt=100)]) > 39000


@pytest.mark.parametrize("gz", [False, True])
def test_empty_and_block_sized_inputs(tmp_path, gz):
    """An empty file, and inputs whose length is exactly kseq's 16 KiB read size or the reader's 1 MiB
    block (uncompressed files are read in place, in views of that size), ending on a record boundary
    and one byte into a header."""
    rng = np.random.default_rng(5)
    cases = [b""]
    for size in (16384, 1 << 20, 2 << 20):
        body = _mixed_records(rng, size // 60)[:size - 64]
        body = body[:body.rfind(b"\n@") + 1] if b"\n@" in body else body
        pad = size - len(body)
        tail = b"@" + b"p" * (pad - 8) + b"\nA\n+\nI\n"    # a last record that ends the file exactly
        assert len(body + tail) == size
        cases.append(body + tail)
        cases.append((body + tail)[:-1] + b"@")          # the last byte starts a header
    for k, data in enumerate(cases):
        f = str(tmp_path / (f"s{k}.fq" + (".gz" if gz else "")))
        _write(f, data, gz)
        _compare(f, PARAMS)
