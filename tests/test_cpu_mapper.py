"""The batched mapper's driver checked on the CPU: minibwa_amd/integration/mbg_map.c (coroutine
rounds, hooks, front pass, the replaced pipeline step) linked with oracle/map_exec_cpu.c, which answers
its device requests with the reference's own functions (oracle/_ref/minibwa-batched-cpu, built by
``make -C oracle batched-cpu``).  Its SAM and hit output must equal the reference binary's (md5s
recorded by tests/golden/make_golden.py), so any difference the GPU mapper shows on the same sets
comes from the device side, not from the batching."""
import os

import pytest

from tests.conftest import GOLDEN, ROOT
from tests.mapper_sets import check_sets

CLI = os.path.join(ROOT, "oracle", "_ref", "minibwa-batched-cpu")
API = os.path.join(ROOT, "oracle", "_ref", "mbmap-batch-batched-cpu")


@pytest.fixture(scope="module")
def exes():
    if not (os.path.exists(CLI) and os.path.exists(API)):
        pytest.skip("oracle/_ref/minibwa-batched-cpu not built (make -C oracle batched-cpu)")
    return CLI, API


@pytest.mark.parametrize("name", ["chrM", "cfg0", "rescue", "batch_chrM", "batch_cfg0", "repeat", "repeat_long"])
def test_batched_driver_matches_reference(exes, tmp_path_factory, name):
    st = check_sets(exes[0], exes[1], name, tmp_path_factory, threads=4)
    assert st["executor"] == "cpu-reference"


REF_CLI = os.path.join(ROOT, "oracle", "_ref", "minibwa")

# output options that change what mb_fmt_sam / mb_fmt_paf write (map-main.c option letters)
OUT_OPTS = {
    "rg": ["-R", r"@RG\tID:grp7\tSM:s1"],
    "comment": ["-y"],
    "supp_soft": ["-Y"],
    "xa": ["--outn", "5", "--xa", "12", "--outs", "0.5"],
    "no_unmap_eqx": ["-u", "--eqx"],
    "no_pairing": ["-P"],
    "paf": ["-f"],
}


def _with_comments(src, dst):
    """The same FASTQ with a comment after each name (for -y)."""
    with open(src) as f, open(dst, "w") as g:
        for i, line in enumerate(f):
            if i % 4 == 0:
                line = line.rstrip("\n") + " BC:Z:AC%d\tXX:i:%d\n" % (i % 7, i)
            g.write(line)


@pytest.mark.parametrize("opt", sorted(OUT_OPTS))
def test_output_options_match_reference(exes, tmp_path_factory, tmp_path, opt):
    """The driver's native SAM/PAF formatter (mbg_fmt.c) on output options the golden sets leave at
    their defaults: every record is also formatted by the reference's mb_format and compared
    (MBG_FMT_CHECK=1 aborts on the first difference), and the whole output equals the reference
    binary's on the same inputs."""
    import subprocess
    from tests.mapper_sets import _body_md5, _files
    if not os.path.exists(REF_CLI):
        pytest.skip("oracle/_ref/minibwa not built (make -C oracle)")
    f = _files(tmp_path_factory)
    r1, r2 = f["p1"], f["p2"]
    if opt == "comment":
        r1, r2 = str(tmp_path / "c1.fq"), str(tmp_path / "c2.fq")
        _with_comments(f["p1"], r1)
        _with_comments(f["p2"], r2)
    args = ["map", "-t4"] + OUT_OPTS[opt] + [f["rep"], r1, r2]
    ref = subprocess.run([REF_CLI] + args, capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0, ref.stderr[-2000:]
    env = dict(os.environ, MBG_MAP_STATS="1", MBG_FMT_CHECK="1")
    got = subprocess.run([exes[0]] + args, capture_output=True, text=True, timeout=300, env=env)
    assert got.returncode == 0, got.stderr[-3000:]
    assert "formatter native" in got.stderr
    n = sum(1 for l in got.stdout.split("\n") if l and not l.startswith("@"))
    assert n >= 4000
    assert _body_md5(got.stdout) == _body_md5(ref.stdout)
    if opt == "rg":
        assert "\tRG:Z:grp7" in got.stdout
    if opt == "xa":
        assert "\tXA:Z:" in got.stdout or "\tn2:i:" in got.stdout
    if opt == "comment":
        assert "\tBC:Z:AC" in got.stdout


def test_long_reads_paf_match_reference(exes, tmp_path_factory):
    """Long reads in PAF (-f), the long-read workload's output, through the native formatter."""
    import subprocess
    from tests.mapper_sets import _body_md5, _files
    if not os.path.exists(REF_CLI):
        pytest.skip("oracle/_ref/minibwa not built (make -C oracle)")
    f = _files(tmp_path_factory)
    args = ["map", "-t4", "-f", f["rep"], f["lr"]]
    ref = subprocess.run([REF_CLI] + args, capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0, ref.stderr[-2000:]
    env = dict(os.environ, MBG_MAP_STATS="1", MBG_FMT_CHECK="1")
    got = subprocess.run([exes[0]] + args, capture_output=True, text=True, timeout=300, env=env)
    assert got.returncode == 0, got.stderr[-3000:]
    assert "formatter native" in got.stderr
    assert "\tcg:Z:" in got.stdout
    assert _body_md5(got.stdout) == _body_md5(ref.stdout)


def test_tuned_settings_file_changes_nothing(exes, tmp_path_factory, tmp_path):
    """A tuned.env file (read at start-up through MBG_TUNE) with extreme batching settings leaves the
    SAM as the reference's: the settings change how requests are batched, never results."""
    tune = tmp_path / "tuned.env"
    tune.write_text("# extreme batching\nMBG_FLUSH=7\n")
    st0 = check_sets(exes[0], exes[1], "rescue", tmp_path_factory, threads=4, extra_env={"MBG_TUNE": "none"})
    st = check_sets(exes[0], exes[1], "rescue", tmp_path_factory, threads=4, extra_env={"MBG_TUNE": str(tune)})
    assert st["executor"] == "cpu-reference"
    assert int(st["batches"]) > int(st0["batches"])      # the file was read: smaller device batches, more of them


def test_many_mini_batches_match_reference(exes, tmp_path_factory):
    """Mini-batches of 300 kbp (-K 300k) over the rescue-heavy pairs: ten mini-batches, each with its
    own insert-size estimate (map-main.c:180-190), the pipeline's read, map and write steps of
    neighbouring mini-batches overlapping; the SAM must equal the reference binary's with the same -K."""
    import subprocess
    from tests.mapper_sets import _body_md5, _files
    if not os.path.exists(REF_CLI):
        pytest.skip("oracle/_ref/minibwa not built (make -C oracle)")
    f = _files(tmp_path_factory)
    args = ["map", "-t4", "-K", "300k", os.path.join(GOLDEN, "cfg0"), f["c1"], f["r2"]]
    ref = subprocess.run([REF_CLI] + args, capture_output=True, text=True, timeout=300)
    assert ref.returncode == 0, ref.stderr[-2000:]
    env = dict(os.environ, MBG_MAP_STATS="1")
    got = subprocess.run([exes[0]] + args, capture_output=True, text=True, timeout=300, env=env)
    assert got.returncode == 0, got.stderr[-3000:]
    pe_r = [l for l in ref.stderr.split("\n") if "mb_pestat::FR] (25, 50, 75)" in l]
    pe_g = [l for l in got.stderr.split("\n") if "mb_pestat::FR] (25, 50, 75)" in l]
    assert len(pe_r) >= 8 and pe_g == pe_r
    assert _body_md5(got.stdout) == _body_md5(ref.stdout)
