"""The end-to-end parity sets the batched mapper is checked on, with the md5s the reference binary
writes for them (tests/golden/golden.json, made by tests/golden/make_golden.py):

  chrM        the reference's own data: 1,000 chrM pairs (test/chrM-read_{1,2}.fa.gz)
  cfg0        10,000 seeded 2x150 pairs on the 16,569 bp seeded reference
  rescue      the same pairs with read 2 at 8% substitutions: hundreds of mate rescues (ksw_ll + extz2)
  batch_*     the public batch API (mb_map_batch) as api-test/ex-batch.c drives it
  repeat      4,000 pairs on a repeat-rich 4-contig reference (intervals over max_occ rows)
  repeat_long 40 reads of 2-8 kb, 4% substitutions, on the same reference (adaptive long-read options)
"""
import hashlib
import json
import os
import subprocess

from tests.conftest import GOLDEN

_cache = {}


def _body_md5(text):
    body = "".join(l + "\n" for l in text.split("\n") if l and not l.startswith("@PG"))
    return hashlib.md5(body.encode()).hexdigest()


def _files(tmp_path_factory):
    if "files" in _cache:
        return _cache["files"]
    from minibwa_amd import index, synth
    d = tmp_path_factory.mktemp("mapsets")
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    r1, r2 = synth.read_pairs(ref, 10000, seed=11)
    f = {k: str(d / v) for k, v in dict(c1="c0_1.fq", c2="c0_2.fq", r2="resc_2.fq", p1="rep1.fq", p2="rep2.fq", lr="long.fq").items()}
    synth.write_pairs_fastq(f["c1"], f["c2"], r1, r2, prefix="c0_")
    _, r2b = synth.read_pairs(ref, 10000, seed=11, sub=0.08)
    synth.write_pairs_fastq(str(d / "tmp1.fq"), f["r2"], r1, r2b, prefix="c0_")
    rr = synth.repeat_reference()
    f["rep"] = str(d / "rep")
    index.build(rr, f["rep"], n_ctg=4, ctg_prefix="rc", device="cpu")
    a1, a2 = synth.read_pairs(rr, 4000, seed=9)
    synth.write_pairs_fastq(f["p1"], f["p2"], a1, a2, prefix="rp_")
    seq, off, ln = synth.single_reads(rr, 40, seed=4, lmin=2000, lmax=8000, sub=0.04)
    synth.write_fastq(f["lr"], seq, off, ln, prefix="lr_")
    _cache["files"] = f
    return f


def _run(exe, args, threads, extra_env=None):
    env = dict(os.environ, MBG_MAP_STATS="1", MBG_MAP_THREADS=str(threads), **(extra_env or {}))
    print("[mapper]", exe, " ".join(args), flush=True)
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [l for l in r.stderr.split("\n") if l.startswith("[mbg_map] summary:")]
    assert lines, r.stderr[-2000:]
    w = lines[-1].replace("|", " ").split()[2:]
    st = {w[i]: w[i + 1] for i in range(0, len(w) - 1, 2)}
    return r.stdout, st


def check_sets(cli, api, name, tmp_path_factory, threads=4, extra_env=None):
    """Runs one set through the mapper programs and checks the reference's md5; returns the run's
    summary counters.  extra_env: settings added to the programs' environment."""
    gold = json.load(open(os.path.join(GOLDEN, "golden.json")))
    f = _files(tmp_path_factory)
    t = ["-t%d" % threads]
    if name == "chrM":
        out, st = _run(cli, ["map"] + t + [os.path.join(GOLDEN, "chrM"), os.path.join(GOLDEN, "chrM-read_1.fa.gz"),
                                            os.path.join(GOLDEN, "chrM-read_2.fa.gz")], threads, extra_env)
        assert _body_md5(out) == gold["chrM.sam.body.md5"]
    elif name == "cfg0":
        out, st = _run(cli, ["map"] + t + [os.path.join(GOLDEN, "cfg0"), f["c1"], f["c2"]], threads, extra_env)
        assert sum(1 for l in out.split("\n") if l and not l.startswith("@")) == gold["cfg0.sam.n_lines"]
        assert _body_md5(out) == gold["cfg0.sam.body.md5"]
    elif name == "rescue":
        out, st = _run(cli, ["map"] + t + [os.path.join(GOLDEN, "cfg0"), f["c1"], f["r2"]], threads, extra_env)
        assert _body_md5(out) == gold["rescue.sam.body.md5"]
        assert int(st["ll_jobs"]) == gold["rescue.n_ll_calls"]
    elif name == "batch_chrM":
        out, st = _run(api, [os.path.join(GOLDEN, "chrM"), os.path.join(GOLDEN, "chrM-read_1.fa.gz")], threads, extra_env)
        assert hashlib.md5(out.encode()).hexdigest() == gold["chrM.batch.md5"]
    elif name == "batch_cfg0":
        out, st = _run(api, [os.path.join(GOLDEN, "cfg0"), f["c1"]], threads, extra_env)
        assert hashlib.md5(out.encode()).hexdigest() == gold["cfg0.batch.md5"]
    elif name == "repeat":
        out, st = _run(cli, ["map"] + t + [f["rep"], f["p1"], f["p2"]], threads, extra_env)
        assert _body_md5(out) == gold["repeat.sam.body.md5"]
    elif name == "repeat_long":
        out, st = _run(cli, ["map"] + t + [f["rep"], f["lr"]], threads, extra_env)
        assert _body_md5(out) == gold["repeat.long.sam.body.md5"]
    else:
        raise KeyError(name)
    # every read went through one front pass; the DP calls went in device batches
    assert int(st["fronts"]) >= 1 and int(st["reads"]) > 0
    if name not in ("batch_chrM", "batch_cfg0"):
        assert int(st["batches"]) >= 1 and int(st["waits"]) >= int(st["batches"])
        assert int(st["dp_jobs"]) >= int(st["dp_calls"])
    return st
