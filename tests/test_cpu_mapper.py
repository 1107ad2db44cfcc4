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
