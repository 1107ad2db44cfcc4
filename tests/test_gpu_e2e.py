"""Drop-in parity end to end: the reference CLI with its four hot-path hooks bound to libmbg
(oracle/_ref/minibwa-mbg, linked by ``make -C oracle hooked`` from the reference's unmodified
objects plus minibwa_amd/integration/mbg_wrap.c) must write the same SAM body as the reference
binary.  The expected md5s were recorded from the reference binary (tests/golden/make_golden.py).

Seeding (mb_seed_intv_batch), anchors (mb_anchor, whose SA locate runs inside it), chaining
(mb_lchain_dp), SA locate (mb_bwt_sa_batch), every banded DP call (ksw_extd2_sse / ksw_extz2_sse) and every mate-rescue local
SW (ksw_ll_qinit + ksw_ll_u8_core / ksw_ll_i16_core) run on the GPU; everything else is the
reference's own code.
"""
import hashlib
import json
import os
import subprocess

import pytest

from tests.conftest import GOLDEN, ROOT

pytestmark = pytest.mark.gpu

HOOKED = os.path.join(ROOT, "oracle", "_ref", "minibwa-mbg")
HOOKED_BATCH = os.path.join(ROOT, "oracle", "_ref", "mbmap-batch-mbg")   # api-test/ex-batch.c


def _sam_body_md5(text):
    body = "".join(l + "\n" for l in text.split("\n") if l and not l.startswith("@PG"))
    return hashlib.md5(body.encode()).hexdigest()


def _run(args, timeout=300, exe=HOOKED, stats=False):
    print("[e2e]", exe, " ".join(args), flush=True)
    env = dict(os.environ, MBG_WRAP_STATS="1") if stats else None
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=timeout, env=env)
    assert r.returncode == 0, r.stderr[-2000:]
    return (r.stdout, r.stderr) if stats else r.stdout


def _calls(err):
    """The hook call counts the wrapper prints at exit under MBG_WRAP_STATS=1."""
    lines = [l for l in err.split("\n") if l.startswith("[mbg_wrap] calls:")]
    assert lines, err[-500:]
    w = lines[-1].split()[2:]
    return {w[i]: int(w[i + 1]) for i in range(0, len(w), 2)}


@pytest.fixture(scope="module")
def gold():
    if not os.path.exists(HOOKED):
        pytest.skip("oracle/_ref/minibwa-mbg not built (make -C oracle hooked)")
    return json.load(open(os.path.join(GOLDEN, "golden.json")))


def test_chrm_pairs_sam_identical(gold):
    """The reference's own test data: 1,000 chrM read pairs (test/chrM-read_{1,2}.fa.gz)."""
    sam = _run(["map", os.path.join(GOLDEN, "chrM"), os.path.join(GOLDEN, "chrM-read_1.fa.gz"),
                os.path.join(GOLDEN, "chrM-read_2.fa.gz")])
    assert _sam_body_md5(sam) == gold["chrM.sam.body.md5"]


def test_batch_api_identical(gold, tmp_path):
    """The public batch API (mb_map_batch, minibwa.h:127-170) as the reference's own example drives
    it (api-test/ex-batch.c:43-77), hooks on the GPU: identical hit lines and CIGARs."""
    if not os.path.exists(HOOKED_BATCH):
        pytest.skip("oracle/_ref/mbmap-batch-mbg not built")
    import hashlib as H
    out = _run([os.path.join(GOLDEN, "chrM"), os.path.join(GOLDEN, "chrM-read_1.fa.gz")], exe=HOOKED_BATCH)
    assert H.md5(out.encode()).hexdigest() == gold["chrM.batch.md5"]
    from minibwa_amd import synth
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    r1, r2 = synth.read_pairs(ref, 10000, seed=11)
    f1, f2 = str(tmp_path / "r1.fq"), str(tmp_path / "r2.fq")
    synth.write_pairs_fastq(f1, f2, r1, r2, prefix="c0_")
    out = _run([os.path.join(GOLDEN, "cfg0"), f1], exe=HOOKED_BATCH)
    assert H.md5(out.encode()).hexdigest() == gold["cfg0.batch.md5"]


def test_cfg0_pairs_sam_identical(gold, tmp_path):
    """10,000 seeded 2x150 pairs on the 16,569 bp seeded reference."""
    import numpy as np
    from minibwa_amd import synth
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    r1, r2 = synth.read_pairs(ref, 10000, seed=11)
    f1, f2 = str(tmp_path / "r1.fq"), str(tmp_path / "r2.fq")
    synth.write_pairs_fastq(f1, f2, r1, r2, prefix="c0_")
    sam = _run(["map", os.path.join(GOLDEN, "cfg0"), f1, f2])
    n = sum(1 for l in sam.split("\n") if l and not l.startswith("@"))
    assert n == gold["cfg0.sam.n_lines"]
    assert _sam_body_md5(sam) == gold["cfg0.sam.body.md5"]


def test_rescue_pairs_sam_identical(gold, tmp_path):
    """A mate-rescue-heavy set (read 2 of each pair with 8% substitutions): hundreds of mates are
    found by mb_matesw (pe.c:319-378), whose ksw_ll and extension DP both run on the GPU through the
    hooks.  The SAM body and the number of ksw_ll calls equal the reference binary's."""
    from minibwa_amd import synth
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    r1, _ = synth.read_pairs(ref, 10000, seed=11)
    _, r2 = synth.read_pairs(ref, 10000, seed=11, sub=0.08)
    f1, f2 = str(tmp_path / "r1.fq"), str(tmp_path / "r2_rescue.fq")
    synth.write_pairs_fastq(f1, f2, r1, r2, prefix="c0_")
    sam, err = _run(["map", os.path.join(GOLDEN, "cfg0"), f1, f2], stats=True)
    n_ll = _calls(err)["ll"]
    assert n_ll == gold["rescue.n_ll_calls"] > 500
    assert _sam_body_md5(sam) == gold["rescue.sam.body.md5"]


@pytest.fixture(scope="module")
def repeat_prefix(tmp_path_factory, gold):
    """The repeat-rich 4-contig reference, indexed by minibwa_amd.index (byte-identical to the
    reference's indexer: the .mbw md5 is the reference binary's)."""
    from minibwa_amd import index, synth
    pre = str(tmp_path_factory.mktemp("rep") / "rep")
    ref = synth.repeat_reference()
    index.build(ref, pre, n_ctg=4, ctg_prefix="rc", device="cpu")
    with open(pre + ".mbw", "rb") as f:
        assert hashlib.md5(f.read()).hexdigest() == gold["repeat.mbw.md5"]
    return pre, ref


def test_repeat_pairs_sam_identical(repeat_prefix, gold, tmp_path):
    """4,000 2x150 pairs on the repeat-rich reference: intervals of hundreds of rows, sampled to
    max_occ, and anchor lists past 64 entries go through mb_anchor on the GPU."""
    from minibwa_amd import synth
    pre, ref = repeat_prefix
    r1, r2 = synth.read_pairs(ref, 4000, seed=9)
    f1, f2 = str(tmp_path / "rep1.fq"), str(tmp_path / "rep2.fq")
    synth.write_pairs_fastq(f1, f2, r1, r2, prefix="rp_")
    sam, err = _run(["map", pre, f1, f2], stats=True)
    c = _calls(err)
    assert c["anchor"] > 7000 and c["chain"] > 7000 and c["sa"] == 0
    assert _sam_body_md5(sam) == gold["repeat.sam.body.md5"]


def test_repeat_long_reads_sam_identical(repeat_prefix, gold, tmp_path):
    """40 long reads (2-8 kb, 4% substitutions) on the repeat-rich reference."""
    from minibwa_amd import synth
    pre, ref = repeat_prefix
    seq, off, ln = synth.single_reads(ref, 40, seed=4, lmin=2000, lmax=8000, sub=0.04)
    fl = str(tmp_path / "long.fq")
    synth.write_fastq(fl, seq, off, ln, prefix="lr_")
    sam, err = _run(["map", pre, fl], stats=True)
    c = _calls(err)
    assert c["anchor"] == 40 and c["chain"] >= 40
    assert _sam_body_md5(sam) == gold["repeat.long.sam.body.md5"]
