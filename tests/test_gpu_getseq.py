"""Device target extraction (mbg_getseq: getseq_kernel, l2b_getseq + l2b_getambi restated) against
the reference's own l2b_getseq (l2bit.c:52-77) on a reference with N runs (tests/nrun_ref.py):
windows around every ambiguous run (both ends, inside, spanning, at the 4- and 32-base packing
units), clipped and invalid windows (negative start, end past the contig, a negative end, which the
reference's unsigned comparison clips to the contig end, empty and reversed windows, bad contig
ids) and random windows.  The expected bytes come from oracle/_ref/l2b-check -d, which calls the
reference's function."""
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT
from tests.nrun_ref import REF_CLI, build, nrun_fasta

pytestmark = pytest.mark.gpu

CHECK = os.path.join(ROOT, "oracle", "_ref", "l2b-check")


def _ranges(ctgs, rng):
    out = []
    d = (0, 1, 2, 3, 4, 5, 31, 32, 33, 64, 200)
    for tid, (_, s) in enumerate(ctgs):
        L = len(s)
        for m in re.finditer(r"[^ACGTacgt]+", s):
            a, b = m.start(), m.end()
            for x in d:
                for y in d[::3]:
                    out += [(tid, a - x, b + y), (tid, a + x, b + y), (tid, a - x, a + y), (tid, b - x, b + y)]
        out += [(tid, 0, L), (tid, -10, L + 10), (tid, L - 5, L + 100), (tid, 10, 10), (tid, 20, 10), (tid, 5, -3),
                (tid, -7, -3), (tid, L, L + 5)]
        for _ in range(300):
            st = int(rng.integers(-50, L + 50))
            out.append((tid, st, st + int(rng.integers(0, 400 if rng.random() < 0.9 else 4000))))
    out += [(-1, 0, 10), (len(ctgs), 0, 10)]
    return out


def test_device_getseq_matches_reference_with_n_runs(tmp_path):
    from minibwa_amd import mbg
    if not (os.path.exists(REF_CLI) and os.path.exists(CHECK)):
        pytest.skip("oracle/_ref not built (make -C oracle ref l2b-check)")
    prefix = build(str(tmp_path))
    ctgs = nrun_fasta(str(tmp_path / "again.fa"))
    R = _ranges(ctgs, np.random.default_rng(17))
    with open(tmp_path / "ranges.txt", "w") as f:
        for t, s, e in R:
            f.write("%d %d %d\n" % (t, s, e))
    subprocess.run([CHECK, "-d", prefix + ".l2b", str(tmp_path / "ranges.txt"), str(tmp_path / "exp.bin")], check=True,
                   timeout=120)
    raw = open(tmp_path / "exp.bin", "rb").read()
    exp, p = [], 0
    for _ in R:
        r = int(np.frombuffer(raw, np.int64, 1, p)[0])
        p += 8
        exp.append((r, np.frombuffer(raw, np.uint8, max(r, 0), p)))
        p += max(r, 0)
    assert p == len(raw)
    n_ambi_bases = sum(int((e[1] == 4).sum()) for e in exp)
    assert n_ambi_bases > 10000
    print("[getseq] %d windows on %s, %d ambiguous bases expected" % (len(R), prefix, n_ambi_bases), flush=True)
    ctx = mbg.Context(0)
    try:
        ctx.idx_load(prefix, 10)
        t, s, e = (np.array([r[k] for r in R], np.int64) for k in range(3))
        got, ln = ctx.l2b_getseq(t, s, e, ctg_len=[len(x) for _, x in ctgs], with_len=True)
    finally:
        ctx.close()
    for i, ((r, b), g) in enumerate(zip(exp, got)):
        assert int(ln[i]) == r, (R[i], r, int(ln[i]))
        if r > 0:
            assert np.array_equal(g, b), (R[i], r)
