"""Host target extraction (minibwa_amd/integration/mbg_l2b.c, which the mapper binds in place of
l2b_getseq) against the reference's own l2b_getseq (l2bit.c:52-77) on a reference with N runs, by
oracle/_ref/l2b-check, which links both; then the mapper end to end on reads from that reference."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT
from tests.nrun_ref import REF_CLI, build, nrun_fasta

CHECK = os.path.join(ROOT, "oracle", "_ref", "l2b-check")
BATCHED = os.path.join(ROOT, "oracle", "_ref", "minibwa-batched-cpu")


@pytest.fixture(scope="module")
def nrun(tmp_path_factory):
    if not (os.path.exists(REF_CLI) and os.path.exists(CHECK)):
        pytest.skip("oracle/_ref not built (make -C oracle ref l2b-check)")
    return build(str(tmp_path_factory.mktemp("nrun")))


def test_getseq_matches_reference_with_n_runs(nrun):
    r = subprocess.run([CHECK, nrun + ".l2b", "50000", "7"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    n, n_ambi = map(int, r.stdout.split()[1:3])
    assert n_ambi >= 20 and n > 50000


def _reads(ctgs, n, seed, path1, path2):
    """Pairs drawn across the ambiguous runs (Ns in the reads too)."""
    rng = np.random.default_rng(seed)
    comp = str.maketrans("ACGTNacgtnRYKryk", "TGCANtgcanYRMyrm")
    with open(path1, "w") as f1, open(path2, "w") as f2:
        for i in range(n):
            name, s = ctgs[int(rng.integers(0, len(ctgs)))]
            p = int(rng.integers(0, len(s) - 400))
            frag = s[p:p + int(rng.integers(250, 400))].upper()
            a, b = frag[:150], frag[-150:].translate(comp)[::-1]
            for f, x in ((f1, a), (f2, b)):
                f.write("@p%d\n%s\n+\n%s\n" % (i, x, "I" * len(x)))


def test_mapper_on_n_runs_matches_reference(nrun, tmp_path):
    if not os.path.exists(BATCHED):
        pytest.skip("oracle/_ref/minibwa-batched-cpu not built (make -C oracle batched-cpu)")
    ctgs = nrun_fasta(str(tmp_path / "again.fa"))
    r1, r2 = str(tmp_path / "r1.fq"), str(tmp_path / "r2.fq")
    _reads(ctgs, 3000, 3, r1, r2)
    args = ["map", "-t4", nrun, r1, r2]
    ref = subprocess.run([REF_CLI] + args, capture_output=True, text=True, timeout=300)
    got = subprocess.run([BATCHED] + args, capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0 and got.returncode == 0, got.stderr[-2000:]
    body = lambda t: [l for l in t.split("\n") if l and not l.startswith("@PG")]
    a, b = body(ref.stdout), body(got.stdout)
    assert len(a) == len(b) >= 6000
    assert a == b
