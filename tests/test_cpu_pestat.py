"""The batched mapper's insert-size model (minibwa_amd/integration/mbg_pe.c: mb_pestat with its pass
over the pairs split over threads) against the reference's own mb_pestat (pe.c:68-124), by
oracle/_ref/pestat-check, which links both: on seeded synthetic hit sets (FR and other orientations,
outliers, non-unique and unmapped mates, mates on other contigs, a max_pe_ins cut, sets too small to
estimate), every field of the four orientations' mb_pestat_t is compared bit for bit at 1, 3 and 16
threads.  The end-to-end mapper tests then check the SAM that pairing under these statistics writes."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

CHECK = os.path.join(ROOT, "oracle", "_ref", "pestat-check")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_pestat_matches_reference(seed):
    if not os.path.exists(CHECK):
        pytest.skip("oracle/_ref not built (make -C oracle ref pestat-check)")
    r = subprocess.run([CHECK, "50", str(seed)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    w = r.stdout.split()
    assert w[0] == "ok", r.stdout
    n_ok, n_failed = int(w[3]), int(w[4])
    assert n_ok >= 40 and n_failed >= 40, r.stdout      # both estimated and failed orientations were compared
