"""Write the committed golden fixtures (run in a container where oracle/_ref is built).

Everything here comes from running the compiled reference (oracle/_ref/minibwa, built from the
reference's own sources by oracle/Makefile):

* ``chrM.mbw`` / ``chrM.l2b``   -- ``minibwa index`` of the reference's own test/chrM-human.fa.gz
* ``cfg0.fa`` / ``cfg0.mbw`` / ``cfg0.l2b`` -- the 16,569 bp seeded uniform reference (splitmix64
  seed 11) and its ``minibwa index``
* ``golden.json`` -- known answers: ``minibwa bench -b 2a|sa|msa`` checksums on chrM, md5 of the
  index files, and md5 of the SAM body (minus @PG) of ``minibwa map`` on the chrM read pairs and on
  the config-0 read pairs (10,000 pairs from minibwa_amd.synth with seed 11), and md5 of the
  output of the reference's batch-API example (api-test/ex-batch.c, mb_map_batch) on the chrM
  read-1 file and on the config-0 read-1 file; the SAM body md5 of a mate-rescue-heavy config-0
  set (read 2 with 8% substitutions) and its number of ksw_ll calls (``--dbg-aln-pe``); the SAM body
  md5 of ``minibwa map`` on the repeat-rich reference of minibwa_amd.synth.repeat_reference (4
  contigs, indexed in a temporary directory) for 4,000 2x150 pairs and for 40 noisy long reads.

Usage: python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "minibwa")
REF_BATCH = os.path.join(ROOT, "oracle", "_ref", "mbmap-batch")   # api-test/ex-batch.c


def md5(path):
    return hashlib.md5(open(path, "rb").read()).hexdigest()


def sam_body_md5(sam_text):
    body = "".join(l + "\n" for l in sam_text.split("\n") if l and not l.startswith("@PG"))
    return hashlib.md5(body.encode()).hexdigest()


def run(cmd, **kw):
    return subprocess.run(cmd, check=True, capture_output=True, text=True, **kw)


def main():
    from minibwa_amd import synth
    out = {}
    pre = os.path.join(HERE, "chrM")
    run([REF_CLI, "index", os.path.join(HERE, "chrM-human.fa.gz"), pre])
    out["chrM.mbw.md5"] = md5(pre + ".mbw")
    out["chrM.l2b.md5"] = md5(pre + ".l2b")
    for b in ("2a", "sa", "msa"):
        r = run([REF_CLI, "bench", "-b", b, "-n", "1000000", pre + ".mbw"])
        cs = [l for l in r.stderr.split("\n") if l.startswith("checksum")][0].split("=")[1].strip()
        out[f"bench.{b}.chrM"] = cs
    sam = run([REF_CLI, "map", pre, os.path.join(HERE, "chrM-read_1.fa.gz"), os.path.join(HERE, "chrM-read_2.fa.gz")]).stdout
    out["chrM.sam.body.md5"] = sam_body_md5(sam)
    r = run([REF_BATCH, pre, os.path.join(HERE, "chrM-read_1.fa.gz")]).stdout
    out["chrM.batch.md5"] = hashlib.md5(r.encode()).hexdigest()

    cfg0 = os.path.join(HERE, "cfg0")
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    synth.write_fasta(cfg0 + ".fa", ref, n_ctg=1, prefix="cfg0_")
    run([REF_CLI, "index", cfg0 + ".fa", cfg0])
    out["cfg0.mbw.md5"] = md5(cfg0 + ".mbw")
    with tempfile.TemporaryDirectory() as td:
        r1, r2 = synth.read_pairs(ref, 10000, seed=11)
        f1, f2 = os.path.join(td, "r1.fq"), os.path.join(td, "r2.fq")
        synth.write_pairs_fastq(f1, f2, r1, r2, prefix="c0_")
        sam = run([REF_CLI, "map", cfg0, f1, f2]).stdout
        out["cfg0.sam.body.md5"] = sam_body_md5(sam)
        out["cfg0.sam.n_lines"] = sum(1 for l in sam.split("\n") if l and not l.startswith("@"))
        r = run([REF_BATCH, cfg0, f1]).stdout
        out["cfg0.batch.md5"] = hashlib.md5(r.encode()).hexdigest()
        # a mate-rescue-heavy set: read 2 of every pair with 8% substitutions, so that hundreds of
        # mates are found by mb_matesw (pe.c:319-378) rather than by seeding
        _, r2 = synth.read_pairs(ref, 10000, seed=11, sub=0.08)
        f2 = os.path.join(td, "r2_rescue.fq")
        synth.write_pairs_fastq(f1, f2, r1, r2, prefix="c0_")
        p = run([REF_CLI, "map", "--dbg-aln-pe", cfg0, f1, f2])
        out["rescue.sam.body.md5"] = sam_body_md5(p.stdout)
        out["rescue.n_ll_calls"] = p.stderr.count("===>")
        # the repeat-rich reference: intervals of hundreds of rows, anchor lists past 64 entries
        rep = os.path.join(td, "rep")
        ref = synth.repeat_reference()
        synth.write_fasta(rep + ".fa", ref, n_ctg=4, prefix="rc")
        run([REF_CLI, "index", rep + ".fa", rep])
        out["repeat.mbw.md5"] = md5(rep + ".mbw")
        r1, r2 = synth.read_pairs(ref, 4000, seed=9)
        f1, f2 = os.path.join(td, "rep1.fq"), os.path.join(td, "rep2.fq")
        synth.write_pairs_fastq(f1, f2, r1, r2, prefix="rp_")
        out["repeat.sam.body.md5"] = sam_body_md5(run([REF_CLI, "map", rep, f1, f2]).stdout)
        seq, off, ln = synth.single_reads(ref, 40, seed=4, lmin=2000, lmax=8000, sub=0.04)
        fl = os.path.join(td, "long.fq")
        synth.write_fastq(fl, seq, off, ln, prefix="lr_")
        out["repeat.long.sam.body.md5"] = sam_body_md5(run([REF_CLI, "map", rep, fl]).stdout)
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1, sort_keys=True))


if __name__ == "__main__":
    main()
