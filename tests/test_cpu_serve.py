"""The serving loop (minibwa_amd/integration/mbg_serve.c) on the CPU executor
(oracle/_ref/serve-batched-cpu, the same loop and driver with the reference's own functions in place
of the device): requests mixed into continuous batches return the same records as each request
alone (batch invariance, with predefined insert-size statistics), and both equal the reference
binary's `minibwa map -I 400,50` on the same reads; a small device-memory budget closes batches
early (memory-gated admission) without changing any record."""
import json
import os
import subprocess

import pytest

from tests.conftest import ROOT

SERVE = os.path.join(ROOT, "oracle", "_ref", "serve-batched-cpu")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "minibwa")


def _body(text):
    return [l for l in text.split("\n") if l and not l.startswith("@")]


def run_serve(exe, prefix, pools, sched_path, out, extra=(), timeout=600):
    r = subprocess.run([exe, "-t4"] + list(extra) + ["-o", out, prefix] + list(pools) + [sched_path], capture_output=True,
                       text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-3000:]
    return json.loads(r.stdout.strip().split("\n")[-1])


def check_against_reference(tmp, prefix, pools, sched, mixed):
    from minibwa_amd import serve_workload as W
    n_p = sum(n for _, k, n in sched if k == "P")
    n_l = sum(n for _, k, n in sched if k == "L")
    u1, u2, ul = (str(tmp / x) for x in ("u1.fa", "u2.fa", "ul.fa"))
    W.pool_prefix(pools[0], n_p, u1), W.pool_prefix(pools[1], n_p, u2), W.pool_prefix(pools[2], n_l, ul)
    ref = subprocess.run([REF_CLI, "map", "-t4", "-I", "400,50", prefix, u1, u2], capture_output=True, text=True, timeout=600)
    assert ref.returncode == 0, ref.stderr[-2000:]
    assert _body(open(mixed + ".pairs.sam").read()) == _body(ref.stdout)
    if n_l:
        ref = subprocess.run([REF_CLI, "map", "-t4", prefix, ul], capture_output=True, text=True, timeout=600)
        assert ref.returncode == 0, ref.stderr[-2000:]
        assert _body(open(mixed + ".long.sam").read()) == _body(ref.stdout)
    return n_p, n_l


def test_serving_batch_invariance_and_parity(tmp_path, tmp_path_factory):
    from minibwa_amd import synth
    from minibwa_amd import serve_workload as W
    from tests.mapper_sets import _files
    if not (os.path.exists(SERVE) and os.path.exists(REF_CLI)):
        pytest.skip("oracle/_ref not built (make -C oracle ref batched-cpu)")
    f = _files(tmp_path_factory)
    rr = synth.repeat_reference()
    pools = W.write_pools(rr, str(tmp_path), 4000, 12, seed=5, long_len=3000)
    sched = W.schedule(40, 0.0, seed=3, max_pairs=180, long_frac=0.15)
    assert any(k == "L" for _, k, _ in sched) and sum(n for _, k, n in sched if k == "P") <= 4000
    sp = str(tmp_path / "sched.txt")
    W.write_schedule(sp, sched)
    mixed, solo, gated = (str(tmp_path / x) for x in ("mixed", "solo", "gated"))
    st_m = run_serve(SERVE, f["rep"], pools, sp, mixed)
    st_s = run_serve(SERVE, f["rep"], pools, sp, solo, ["-s"])
    st_g = run_serve(SERVE, f["rep"], pools, sp, gated, ["-b", "1"])
    assert st_m["requests"] == 40 and st_s["batches"] == 40
    assert st_m["batches"] < 40 and st_m["reqs_per_batch"] > 1      # requests really shared batches
    assert st_g["gated_batches"] > 0 and st_g["batches"] > st_m["batches"]
    for ext in (".pairs.sam", ".long.sam"):
        a = open(mixed + ext).read()
        assert a and a == open(solo + ext).read() == open(gated + ext).read()
    n_p, n_l = check_against_reference(tmp_path, f["rep"], pools, sched, mixed)
    assert n_p > 500 and n_l >= 1
