"""The serving loop on the GPU (minibwa_amd/bin/mbg-serve-bench: mbg_serve.c over the batched GPU
mapper): mixed continuous batches return the same records as each request alone, as a run whose
device-memory budget gates admission, and as the reference binary's `minibwa map -I 400,50`."""
import os

import pytest

from tests.conftest import ROOT
from tests.test_cpu_serve import check_against_reference, run_serve

pytestmark = pytest.mark.gpu

SERVE = os.path.join(ROOT, "minibwa_amd", "bin", "mbg-serve-bench")


def test_gpu_serving_batch_invariance_and_parity(tmp_path, tmp_path_factory):
    from minibwa_amd import synth
    from minibwa_amd import serve_workload as W
    from tests.mapper_sets import _files
    assert os.path.exists(SERVE), "minibwa_amd/bin not built (make -C minibwa_amd/integration)"
    f = _files(tmp_path_factory)
    rr = synth.repeat_reference()
    pools = W.write_pools(rr, str(tmp_path), 8000, 20, seed=5, long_len=4000)
    sched = W.schedule(60, 2000.0, seed=3, max_pairs=300, long_frac=0.1)
    sp = str(tmp_path / "sched.txt")
    W.write_schedule(sp, sched)
    mixed, solo, gated = (str(tmp_path / x) for x in ("mixed", "solo", "gated"))
    st_m = run_serve(SERVE, f["rep"], pools, sp, mixed)
    print("[serve] mixed:", st_m, flush=True)
    st_s = run_serve(SERVE, f["rep"], pools, sp, solo, ["-s"])
    st_g = run_serve(SERVE, f["rep"], pools, sp, gated, ["-b", "2"])
    print("[serve] gated:", st_g, flush=True)
    assert st_s["batches"] == 60 and st_g["gated_batches"] > 0
    assert st_m["work_bytes"] > 0 and st_m["bytes_per_base"] > 0
    for ext in (".pairs.sam", ".long.sam"):
        a = open(mixed + ext).read()
        assert a and a == open(solo + ext).read() == open(gated + ext).read()
    check_against_reference(tmp_path, f["rep"], pools, sched, mixed)
