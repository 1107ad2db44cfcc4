import gzip
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

NT4 = np.full(256, 4, np.uint8)
for _i, _c in enumerate("ACGT"):
    NT4[ord(_c)] = _i
    NT4[ord(_c.lower())] = _i


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def read_fasta(path):
    """(names, sequences as str) of a FASTA/FASTQ file, gzip or plain."""
    op = gzip.open if path.endswith(".gz") else open
    names, seqs, cur = [], [], []
    with op(path, "rt") as f:
        lines = f.read().split("\n")
    i = 0
    if lines and lines[0].startswith("@"):
        while i + 1 < len(lines):
            if lines[i].startswith("@"):
                names.append(lines[i][1:].split()[0]); seqs.append(lines[i + 1]); i += 4
            else:
                i += 1
        return names, seqs
    for ln in lines:
        if ln.startswith(">"):
            if cur:
                seqs.append("".join(cur)); cur = []
            names.append(ln[1:].split()[0])
        elif ln:
            cur.append(ln.strip())
    if cur or len(seqs) < len(names):
        seqs.append("".join(cur))
    return names, seqs


def nt4(s: str) -> np.ndarray:
    return NT4[np.frombuffer(s.encode(), np.uint8)]


@pytest.fixture(scope="session")
def chrm_prefix():
    p = os.path.join(GOLDEN, "chrM")
    assert os.path.exists(p + ".mbw"), "tests/golden/chrM.mbw missing (tests/golden/make_golden.py)"
    return p


@pytest.fixture(scope="session")
def chrm_reads():
    _, s1 = read_fasta(os.path.join(GOLDEN, "chrM-read_1.fa.gz"))
    _, s2 = read_fasta(os.path.join(GOLDEN, "chrM-read_2.fa.gz"))
    return [nt4(s) for s in s1 + s2]


@pytest.fixture(scope="session")
def gpu_ctx():
    from minibwa_amd import mbg
    ctx = mbg.Context(0)
    yield ctx
    ctx.close()
