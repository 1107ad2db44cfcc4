"""Host-side checks that need no GPU: the C ABI library and its header, the Python mirror's struct
layouts, option derivation against the reference's own mb_opt_init/mb_opt_adap, the index builder
against the reference indexer's bytes, and the synthetic-input generators."""
import ctypes as C
import hashlib
import json
import os
import re
import subprocess

import numpy as np
import pytest

from tests.conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "mbg.h")
LIBMBG = os.path.join(ROOT, "minibwa_amd", "libmbg.so")


def _declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mbg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    """libmbg.so loads on a host without a GPU and exports every function include/mbg.h declares
    (no compute call is made)."""
    assert os.path.exists(LIBMBG), "libmbg.so not built (python -m minibwa_amd.build)"
    lib = C.CDLL(LIBMBG)
    names = _declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", LIBMBG], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(names) <= exported
    from minibwa_amd import mbg
    assert set(mbg.EXPORTED) <= set(names)


def test_strerror_without_device():
    lib = C.CDLL(LIBMBG)
    lib.mbg_strerror.restype = C.c_char_p
    msgs = {rc: lib.mbg_strerror(rc).decode() for rc in range(0, -9, -1)}
    assert msgs[0] and len(set(msgs.values())) == len(msgs)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_context_fails_loudly_without_gpu():
    """No CPU fallback: without a device the context refuses to start."""
    from minibwa_amd import mbg
    with pytest.raises(mbg.MbgError):
        mbg.Context(0)


def test_header_layouts_match_python(tmp_path):
    """sizeof/offsetof of every struct crossing the ABI, compiled from include/mbg.h by the host C
    compiler, equal the ctypes/numpy layouts the Python mirror uses."""
    from minibwa_amd import mbg
    checks = {
        "mbg_sai_t": ["size", "info"],
        "mbg_seed_rec_t": ["rid", "pass"],
        "mbg_dp_job_t": ["qoff", "toff", "qlen", "tlen", "q", "e", "q2", "e2", "w", "zdrop", "end_bonus", "flag",
                         "cigar_off"],
        "mbg_align_job_t": ["qoff", "toff", "qlen", "tlen", "w", "end_bonus", "zdrop", "flag", "cigar_off"],
        "mbg_align_opt_t": ["a", "b", "b_ts", "q", "e", "q2", "e2", "meth", "max_sw_mat"],
        "mbg_dp_result_t": ["max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score",
                            "reach_end", "n_cigar", "cells"],
        "mbg_job_meta_t": ["rid", "side", "rev", "tid", "tpos"],
        "mbg_map_opt_t": ["min_len", "max_sub_occ", "max_occ", "a", "b", "q", "e", "q2", "e2", "end_bonus", "ksw_w",
                          "zdrop", "mat"],
        "mbg_map_stats_t": ["n_reads", "n_rec", "n_rows", "n_anchored", "n_fast", "n_jobs", "q_bytes", "t_bytes",
                            "cells", "occ_blocks", "kmer_lookups", "extend_steps", "lf_steps", "ms_seed1", "ms_seed2",
                            "ms_rows", "ms_sa", "ms_jobs", "ms_dp", "ms_total"],
        "mbg_map_out_t": ["recs", "pos", "pos_rec", "jobs", "meta", "res", "cigar", "qbuf", "tbuf"],
        "mbg_bwt_desc_t": ["primary", "L2", "seq_len", "data", "sa_bit", "n_sa", "sa"],
        "mbg_l2b_desc_t": ["tot_len", "n_ctg", "ctg_len", "n_pac", "pac", "n_ambi", "ambi"],
        "mbg_idx_info_t": ["seq_len", "primary", "tot_len", "n_ctg", "pre_len", "sa_bit", "device_bytes"],
        "mbg_ll_job_t": ["qoff", "toff", "qlen", "tlen", "size", "pad"],
        "mbg_ll_result_t": ["score", "te", "qe", "score2", "te2"],
        "mbg_matesw_job_t": ["qoff", "qlen", "tid", "ts", "te", "min_sc", "pad", "cigar_off"],
        "mbg_matesw_opt_t": ["a", "mt"],
        "mbg_anchor_t": ["sid", "len", "qpos", "flag", "tpos"],
        "mbg_chain_opt_t": ["max_dist_x", "max_dist_y", "bw", "max_skip", "max_iter", "min_sc", "pen_gap", "pad"],
        "mbg_extra_job_t": ["qoff", "toff", "qs", "qe", "ts", "te", "rev", "n_cigar", "cigar_off", "blen", "mlen", "n_ambi",
                            "dp_max"],
        "mbg_matesw_result_t": ["max_ug", "max_i", "n_good", "n_kmer", "ts2", "te2", "aligned", "rescued", "ll", "ez", "qs",
                                "qe", "ts", "te", "dp_max2", "ran_dp", "n_cigar", "blen", "mlen", "n_ambi", "dp_max"],
    }
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "mbg.h"', "int main(void) {"]
    for s, fields in checks.items():
        src.append(f'printf("{s} sizeof %zu\\n", sizeof({s}));')
        for f in fields:
            src.append(f'printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    src.append("return 0; }")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(c), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)

    def np_layout(s, dt, rename=None):
        assert got[(s, "sizeof")] == dt.itemsize, s
        for f in checks[s]:
            if f in ("size", "info") and s == "mbg_sai_t":
                continue
            name = (rename or {}).get(f, f)
            if name in dt.fields:
                assert got[(s, f)] == dt.fields[name][1], (s, f)

    np_layout("mbg_dp_job_t", mbg.DP_JOB_DTYPE)
    np_layout("mbg_dp_result_t", mbg.DP_RES_DTYPE)
    np_layout("mbg_align_job_t", mbg.ALIGN_JOB_DTYPE)
    np_layout("mbg_job_meta_t", mbg.JOB_META_DTYPE)
    np_layout("mbg_ll_job_t", mbg.LL_JOB_DTYPE)
    np_layout("mbg_ll_result_t", mbg.LL_RES_DTYPE)
    np_layout("mbg_matesw_job_t", mbg.MATESW_JOB_DTYPE)
    np_layout("mbg_extra_job_t", mbg.EXTRA_JOB_DTYPE)
    np_layout("mbg_matesw_result_t", mbg.MATESW_RES_DTYPE)
    np_layout("mbg_anchor_t", mbg.ANCHOR_DTYPE)
    assert got[("mbg_chain_opt_t", "sizeof")] == C.sizeof(mbg.ChainOpt)
    for f, _ in mbg.ChainOpt._fields_:
        assert got[("mbg_chain_opt_t", f)] == getattr(mbg.ChainOpt, f).offset, f
    assert got[("mbg_matesw_opt_t", "sizeof")] == C.sizeof(mbg.MateswOpt)
    sr = mbg.SEED_REC_DTYPE
    assert got[("mbg_seed_rec_t", "sizeof")] == sr.itemsize
    assert got[("mbg_seed_rec_t", "rid")] == sr.fields["rid"][1] and got[("mbg_seed_rec_t", "pass")] == sr.fields["pass"][1]

    def ct_layout(s, cls):
        assert got[(s, "sizeof")] == C.sizeof(cls), s
        for f in checks[s]:
            x = "x" if f == "x" else f
            if hasattr(cls, x):
                assert got[(s, f)] == getattr(cls, x).offset, (s, f)

    ct_layout("mbg_map_opt_t", mbg.MapOpt)
    ct_layout("mbg_align_opt_t", mbg.AlignOpt)
    ct_layout("mbg_map_stats_t", mbg.MapStats)
    ct_layout("mbg_map_out_t", mbg.MapOut)
    ct_layout("mbg_bwt_desc_t", mbg.BwtDesc)
    ct_layout("mbg_l2b_desc_t", mbg.L2bDesc)
    ct_layout("mbg_idx_info_t", mbg.IdxInfo)
    assert got[("mbg_sai_t", "sizeof")] == C.sizeof(mbg.Sai) == 32


def test_nt4_mat_definition():
    """ksw_gen_nt4_mat (ksw2.h:199-219): a on the diagonal, -b off it, -b_ambi on the N row/column."""
    from minibwa_amd import mbg
    for a, b, amb in [(2, 8, 1), (1, 4, 2), (3, 5, 1)]:
        m = np.asarray(mbg.nt4_mat(a, b, 0, amb)).reshape(5, 5)
        exp = np.full((5, 5), -amb)
        exp[:4, :4] = -b
        np.fill_diagonal(exp[:4, :4], a)
        assert np.array_equal(m, exp)


class _OptHead(C.Structure):
    # mb_opt_t (minibwa.h:38-88) up to the alignment options
    _fields_ = ([("flag", C.c_uint64)] +
                [(n, C.c_int32) for n in ("min_len", "max_sub_occ", "max_occ", "bw", "bw_long", "max_gap",
                                          "max_sr_len", "max_chain_skip", "max_chain_iter", "min_chain_score")] +
                [("chain_gap_scale", C.c_float), ("mask_level", C.c_float), ("mask_len", C.c_int32),
                 ("pri_ratio", C.c_float), ("best_n", C.c_int32)] +
                [(n, C.c_int32) for n in ("a", "b", "b_ts", "b_ambi", "q", "q2", "e", "e2", "end_bonus",
                                          "min_dp_max", "zdrop", "zdrop_inv", "min_ksw_len")] +
                [("rest", C.c_uint8 * 512)])


def test_adaptive_options_vs_reference():
    """The step's options for a read length equal the reference's mb_opt_init + mb_opt_adap
    (options.c:47-116) run from the compiled reference, with ksw_w = (int)(bw*1.5+1) (align.c:564)
    and zdrop*a (align.c:374-376)."""
    O = pytest.importorskip("oracle.oracle")
    if not O.have_ref():
        pytest.skip("oracle/_ref not built")
    from minibwa_amd import mbg
    L = O.ref()
    L.mb_opt_init.argtypes = [C.POINTER(_OptHead)]
    L.mb_opt_adap.argtypes = [C.POINTER(_OptHead), C.c_int32, C.POINTER(_OptHead)]
    o0 = _OptHead()
    L.mb_opt_init(C.byref(o0))
    for ln in (50, 100, 101, 150, 250, 1000, 2000, 5000, 20000):
        oa = _OptHead()
        L.mb_opt_adap(C.byref(o0), ln, C.byref(oa))
        m = mbg.adaptive_opt(ln)
        assert m.ksw_w == int(oa.bw * 1.5 + 1.0), ln
        assert m.zdrop == oa.zdrop * oa.a, ln
        for f in ("min_len", "max_sub_occ", "max_occ", "a", "b", "q", "e", "q2", "e2", "end_bonus"):
            assert getattr(m, f) == getattr(oa, f), (ln, f)
        # mb_lchain_dp's arguments (map-algo.c:607-616, 634)
        pen = C.c_float(oa.chain_gap_scale * .01 * oa.min_len).value
        for rechain, bw in ((False, oa.bw), (True, oa.bw_long)):
            co = mbg.chain_opt(ln, rechain)
            assert (co.max_dist_x, co.max_dist_y, co.bw, co.max_skip, co.max_iter, co.min_sc) == \
                (oa.max_gap, oa.max_gap, bw, oa.max_chain_skip, oa.max_chain_iter, oa.min_chain_score), (ln, rechain)
            assert co.pen_gap == pen, ln


def test_index_builder_matches_reference_indexer(tmp_path):
    """minibwa_amd.index on the 16,569 bp seeded reference writes the same .mbw bytes as
    `minibwa index` (md5 recorded from the reference binary) and the same .l2b as the committed one."""
    from minibwa_amd import index, synth
    gold = json.load(open(os.path.join(GOLDEN, "golden.json")))
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    pre = str(tmp_path / "cfg0")
    index.build(ref, pre, n_ctg=1, ctg_prefix="cfg0_", device="cpu")
    assert hashlib.md5(open(pre + ".mbw", "rb").read()).hexdigest() == gold["cfg0.mbw.md5"]
    assert open(pre + ".l2b", "rb").read() == open(os.path.join(GOLDEN, "cfg0.l2b"), "rb").read()


def test_synth_deterministic():
    from minibwa_amd import synth
    a = synth.reference(5000, seed=11)
    b = synth.reference(5000, seed=11)
    c = synth.reference(5000, seed=12)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.max() <= 3
    r1, r2 = synth.read_pairs(a, 50, seed=3)
    s1, s2 = synth.read_pairs(a, 50, seed=3)
    assert np.array_equal(r1, s1) and np.array_equal(r2, s2) and r1.shape == (50, 150)


@pytest.mark.parametrize("L,rep,n_ctg,chunk", [(32, 0.0, 1, 10), (63, 0.0, 1, 20), (96, 0.0, 1, 50), (97, 0.0, 1, 50),
                                               (3000, 0.5, 2, 700), (200_000, 0.3, 3, 40_000)])
def test_index_builder_matches_reference_binary(tmp_path, L, rep, n_ctg, chunk):
    """The chunked suffix sort (several key-range chunks, repeat-rich texts, last occ block with at
    most 64 symbols) writes the same .mbw/.l2b bytes as `minibwa index` of the compiled reference."""
    O = pytest.importorskip("oracle.oracle")
    if not os.path.exists(O.REF_CLI):
        pytest.skip("oracle/_ref/minibwa not built")
    from minibwa_amd import index, synth
    ref = synth.reference(L, seed=5, rep_frac=rep, rep_min=20 if L < 10000 else 300, rep_max=200 if L < 10000 else 6000,
                          rep_div=0.02)
    pre = str(tmp_path / "g")
    index.build(ref, pre, n_ctg=n_ctg, device="cpu", max_chunk=chunk)
    synth.write_fasta(str(tmp_path / "r.fa"), ref, n_ctg=n_ctg, prefix="chr")
    subprocess.run([O.REF_CLI, "index", str(tmp_path / "r.fa"), str(tmp_path / "r")], check=True, capture_output=True)
    for ext in (".mbw", ".l2b"):
        assert open(pre + ext, "rb").read() == open(str(tmp_path / "r") + ext, "rb").read(), ext
