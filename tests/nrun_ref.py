"""A small reference with ambiguous bases where target extraction must restore them: N runs at
contig starts and ends, of lengths around the 4- and 32-base packing units, several runs inside one
read-sized window, single IUPAC codes and lowercase bases; indexed by the reference CLI built under
oracle/_ref (its `index` command records the ambiguous runs in the .l2b)."""
import os
import subprocess

import numpy as np

from tests.conftest import ROOT

REF_CLI = os.path.join(ROOT, "oracle", "_ref", "minibwa")


def nrun_fasta(path, seed=5):
    rng = np.random.default_rng(seed)
    acgt = np.array(list("ACGT"))

    def bases(n):
        return "".join(acgt[rng.integers(0, 4, n)])

    ctgs = []
    s = "N" * 7 + bases(900)
    for run in (1, 2, 3, 4, 5, 31, 32, 33, 63, 64, 65, 100):
        s += "N" * run + bases(int(rng.integers(20, 180)))
    s += "NNNN" + bases(12) + "N" + bases(5) + "NNNNNNNNNNNN" + bases(9) + "N" * 33 + bases(700)   # clustered runs
    s += "R" + bases(50) + "y" + bases(50) + "K" + bases(60) + "acgtacgtnnnnacgt" + bases(400)  # IUPAC, lowercase
    s += "N" * 40
    ctgs.append(("nA", s))
    ctgs.append(("nB", "N" * 300 + bases(2500) + "N" * 29 + bases(1500)))
    ctgs.append(("nC", bases(3000)))
    ctgs.append(("nD", bases(64) + "N" * 64 + bases(64) + "N" + bases(2000)))
    with open(path, "w") as f:
        for name, seq in ctgs:
            f.write(">%s\n" % name)
            for i in range(0, len(seq), 70):
                f.write(seq[i:i + 70] + "\n")
    return ctgs


def build(d):
    """Writes d/nrun.fa and indexes it with the reference CLI; returns the index prefix."""
    fa, prefix = os.path.join(d, "nrun.fa"), os.path.join(d, "nrun")
    nrun_fasta(fa)
    subprocess.run([REF_CLI, "index", fa, prefix], check=True, capture_output=True, timeout=120)
    return prefix
