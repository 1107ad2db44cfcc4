"""mbg_align_pair_batch (mb_align_pair, align.c:329-390, on the device) against the reference's own
static function, exported by oracle/ref_align_shim.c: every ksw_extz_t field and the CIGAR, on the
seeded calls of tests/align_cases.py, which reach both ungapped fast paths (taken and refused), the
band cap, the max_sw_mat skip, generic scoring and both DP kernels; batched per option set."""
import collections

import numpy as np
import pytest

from oracle import oracle as O
from tests.align_cases import cases

pytestmark = pytest.mark.gpu

FIELDS = ("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score", "reach_end")


def _run(ctx, group):
    p, msm = group[0][0], group[0][1]
    qs = [c[2] for c in group]
    ts = [c[3] for c in group]
    params = {"w": np.array([c[5] for c in group]), "end_bonus": np.array([c[6] for c in group]),
              "zdrop": np.array([c[7] for c in group]), "flag": np.array([c[8] for c in group])}
    mats = {c[4].tobytes() for c in group}
    assert len(mats) == 1
    return ctx.align_pair(p, msm, qs, ts, params, group[0][4])


@pytest.mark.parametrize("seed", [5, 6])
def test_align_pair_batch_vs_reference(gpu_ctx, seed):
    groups = collections.defaultdict(list)
    for c in cases(seed, 2500):
        groups[(tuple(c[0].tolist()), c[1])].append(c)
    n_fast = n_dp = 0
    for key, group in sorted(groups.items()):
        print(f"[align_pair] seed {seed} opts {key} calls {len(group)}", flush=True)
        out = _run(gpu_ctx, group)
        for i, c in enumerate(group):
            want = O.ref_align_pair(*c)
            got = {f: int(out.res[f][i]) for f in FIELDS}
            got["cigar"] = out.cigar_of(i)
            assert got == {f: want[f] for f in FIELDS + ("cigar",)}, \
                f"seed {seed} opts {key} call {i}: qlen={len(c[2])} tlen={len(c[3])} w={c[5]} eb={c[6]} zd={c[7]} flag={c[8]:#x}"
            if int(out.res["cells"][i]) > 0:
                n_dp += 1
            else:
                n_fast += 1
    assert n_fast > 300 and n_dp > 300


def test_align_pair_batch_edges(gpu_ctx):
    """An empty batch; a call with an empty query (the fast path pushes a zero-length M); a lone
    max_sw_mat skip; and the entry point's argument checks."""
    from minibwa_amd import mbg
    p = O.align_params()
    mat = O.gen_nt4_mat(2, 8, 0, 1)
    z = np.zeros(0, np.uint8)
    t = np.array([0, 1, 2, 3], np.uint8)
    assert len(gpu_ctx.align_pair(p, 0, [], [], {"w": 10, "end_bonus": 5, "zdrop": 100, "flag": 0}, mat).res) == 0
    out = gpu_ctx.align_pair(p, 0, [z, z], [t, z], {"w": 10, "end_bonus": 5, "zdrop": 100, "flag": np.array([0x40, 0])}, mat)
    for i, (q_, t_, fl) in enumerate([(z, t, 0x40), (z, z, 0)]):
        want = O.ref_align_pair(p, 0, q_, t_, mat, 10, 5, 100, fl)
        assert {f: int(out.res[f][i]) for f in FIELDS} == {f: want[f] for f in FIELDS} and out.cigar_of(i) == want["cigar"]
    q = np.arange(300, dtype=np.uint8) % 4
    out = gpu_ctx.align_pair(p, 1000, [q], [q.copy()], {"w": 10, "end_bonus": 5, "zdrop": 100, "flag": 0x08}, mat)
    want = O.ref_align_pair(p, 0, q, q, mat, 10, 5, 100, 0x08)   # equal: the global fast path comes first
    assert int(out.res["score"][0]) == want["score"] and out.cigar_of(0) == want["cigar"]
    q2 = q.copy()
    q2[::7] = (q2[::7] + 1) % 4
    out = gpu_ctx.align_pair(p, 1000, [q], [q2], {"w": 10, "end_bonus": 5, "zdrop": 100, "flag": 0x08}, mat)
    want = O.ref_align_pair(p, 1000, q, q2, mat, 10, 5, 100, 0x08)
    assert int(out.res["zdropped"][0]) == 1 == want["zdropped"] and out.cigar_of(0) == [] == want["cigar"]
    with pytest.raises(mbg.MbgError):
        gpu_ctx.align_pair(O.align_params(e=0), 0, [q], [q], {"w": 10, "end_bonus": 5, "zdrop": 100, "flag": 0}, mat)
