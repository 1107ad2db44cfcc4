"""The batched GPU mapper end to end (minibwa_amd/bin/minibwa-gpu and mbmap-batch-gpu, built by
``make -C minibwa_amd/integration``): the reference's CLI and batch API with one device front pass
(seeding -> anchors -> chains) per mini-batch and every DP and mate-rescue local-SW call executed in
device rounds.  SAM and hit output must equal the reference binary's on every parity set
(tests/mapper_sets.py), and the run summary must show that the work went through the device."""
import os

import pytest

from tests.conftest import ROOT
from tests.mapper_sets import check_sets

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "minibwa_amd", "bin", "minibwa-gpu")
API = os.path.join(ROOT, "minibwa_amd", "bin", "mbmap-batch-gpu")


@pytest.fixture(scope="module")
def exes():
    assert os.path.exists(CLI) and os.path.exists(API), "minibwa_amd/bin not built (make -C minibwa_amd/integration)"
    return CLI, API


@pytest.mark.parametrize("name", ["chrM", "cfg0", "rescue", "batch_chrM", "batch_cfg0", "repeat", "repeat_long"])
def test_gpu_mapper_matches_reference(exes, tmp_path_factory, name):
    st = check_sets(exes[0], exes[1], name, tmp_path_factory, threads=8)
    assert st["executor"] == "gpu"


@pytest.mark.parametrize("name", ["cfg0", "rescue", "repeat_long"])
def test_gpu_mapper_tuned_settings_change_nothing(exes, tmp_path_factory, tmp_path, name):
    """The autotuner's settings (launch shapes of the seeding, SA and DP kernels, the device batch size,
    the long-read segment length), read from a tuned.env file at start-up, change speed only: with
    extreme values the SAM still equals the reference's."""
    tune = tmp_path / "tuned.env"
    tune.write_text("# extreme launch shapes\nMBG_GATHER_BPC=1\nMBG_DP_WPC=3\nMBG_FLUSH=100\nMBG_SEED_SEG=64\n")
    st = check_sets(exes[0], exes[1], name, tmp_path_factory, threads=8, extra_env={"MBG_TUNE": str(tune)})
    assert st["executor"] == "gpu"
