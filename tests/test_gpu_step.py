"""GPU parity of the seed -> locate -> extend step (mbg_map_step_dev) on the config-0 reference.

Checks every stage against the compiled reference or a restatement of the reference's rules:
seeding sets (mb_seed_intv_batch), located positions (mb_bwt_sa_batch), the anchor-row sampling
(seed.c:304-315), the extension jobs (windows of mb_align1, fast path and band cap of
mb_align_pair, align.c:321-372) and the DP results (ksw_extd2_sse on the emitted jobs).  The step's
fast-path and band decisions are also pinned against the reference's own mb_align_pair (static in
align.c, exported by oracle/ref_align_shim.c): every extension the step settles without DP, the
reference settles with one M, and every job it emits, the reference aligns to the same result.
"""
import os

import numpy as np
import pytest

from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

FIELDS = ("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score", "reach_end")


def sample_rows(x0, size, max_occ=250):
    if size <= max_occ:
        return list(range(x0, x0 + size))
    out, j, k = [], 0, 0
    while j < size and k < max_occ:
        step = max((size - j) // (max_occ - k), 1)
        out.append(x0 + j)
        j += step
        k += 1
    return out


def expected_jobs(reads, recs, pos_of_rec, ref, ctg_len, opt):
    """Restatement of the step's job builder from host data (mb_align1 / mb_align_pair rules)."""
    from minibwa_amd import mbg
    mat = np.array(list(opt.mat), np.int32)
    tot = int(sum(ctg_len))
    off = np.concatenate([[0], np.cumsum(ctg_len)])
    best = {}
    for i, r in enumerate(recs):
        if r["pass"] != 1 or r["size"] == 0:
            continue
        qs, qe = int(r["info"]) >> 32, int(r["info"]) & 0xffffffff
        key = (qe - qs, -qs)
        rid = int(r["rid"])
        if rid not in best or key > best[rid][0]:
            best[rid] = (key, i)
    jobs = {}
    fast = 0
    fast_calls = []
    for rid, (_, i) in best.items():
        r = recs[i]
        qs0, qe0 = int(r["info"]) >> 32, int(r["info"]) & 0xffffffff
        alen = qe0 - qs0
        st = pos_of_rec[i]
        en = st + alen
        if en > 2 * tot or (st < tot < en):
            continue
        rev = st >= tot
        sp = st if not rev else 2 * tot - en
        tid = int(np.searchsorted(off, sp, side="right") - 1)
        if not (off[tid] <= sp < off[tid + 1]) or sp + alen > off[tid + 1]:
            continue
        q = reads[rid]
        qlen = len(q)
        oq = q if not rev else np.where(q[::-1] < 4, 3 - q[::-1], 4).astype(np.uint8)
        coff, clen = int(off[tid]), int(ctg_len[tid])
        ts, te = sp - coff, sp - coff + alen
        qs, qe = (qlen - qe0, qlen - qs0) if rev else (qs0, qe0)

        def one(qq, tt, flag, side):
            nonlocal fast
            w, n_mm = opt.ksw_w, -1
            if len(tt) >= len(qq):
                sc = mat[tt[:len(qq)].astype(np.int32) * 5 + qq.astype(np.int32)]
                n_mm = -1 + int(np.sum((tt[:len(qq)] > 3) | (qq > 3) | (sc < 0)))
                cs = np.cumsum(sc)
                mx = max(0, int(cs.max()))
                if n_mm <= 2 and int(cs[-1]) + opt.end_bonus >= mx:
                    fast += 1
                    fast_calls.append((qq.copy(), tt.copy(), flag))
                    return
            if n_mm >= 0 and min(len(qq), len(tt)) < 100:
                x = n_mm * (opt.a + opt.b)
                m1 = (x - opt.q + opt.e - 1) // opt.e if x >= opt.q + opt.e else 0
                m2 = (x - opt.q2 + opt.e2 - 1) // opt.e2 if x >= opt.q2 + opt.e2 else 0
                w = min(w, max(m1, m2) + 4)
            jobs[(rid, side)] = (qq.copy(), tt.copy(), w, flag)

        if qs > 0 and ts > 0:
            l = qs
            l += (l * opt.a + opt.end_bonus - opt.q) // opt.e if l * opt.a + opt.end_bonus > opt.q else 0
            ts0 = max(ts - l, 0)
            one(oq[:qs][::-1], ref[coff + ts0: coff + ts][::-1], 0xc2, 0)
        if qe < qlen and te < clen:
            l = qlen - qe
            l += (l * opt.a + opt.end_bonus - opt.q) // opt.e if l * opt.a + opt.end_bonus > opt.q else 0
            te0 = min(te + l, clen)
            one(oq[qe:], ref[coff + te: coff + te0], 0x40, 1)
    return jobs, fast, len(best), fast_calls


def test_map_step_cfg0(gpu_ctx):
    from minibwa_amd import mbg, synth
    from oracle import oracle as O
    pre = os.path.join(GOLDEN, "cfg0")
    gpu_ctx.idx_load(pre, 10)
    ref = synth.reference(16569, seed=11, rep_frac=0.0)
    r1, r2 = synth.read_pairs(ref, 1500, seed=5, sub=0.02, indel=0.004)
    reads = [x for pair in zip(r1, r2) for x in pair]
    n = len(reads)
    rl = 150
    d_seq = gpu_ctx.to_device(np.concatenate(reads))
    d_off = gpu_ctx.to_device(np.arange(0, n + 1, dtype=np.int64) * rl)
    d_len = gpu_ctx.to_device(np.full(n, rl, np.int32))
    opt = mbg.adaptive_opt(rl)
    st, out = gpu_ctx.map_step_dev(n, d_off.ptr, d_len.ptr, d_seq.ptr, rl, opt, collect_stats=True)
    recs = gpu_ctx.to_host(out.recs, mbg.SEED_REC_DTYPE, st.n_rec)
    # 1. seeding == mb_seed_intv_batch (as sets per read)
    ri = O.RefIndex(pre + ".mbw")
    want = ri.seed_batch(reads)
    for i in range(n):
        g = recs[recs["rid"] == i]
        ga = sorted(zip(g["x0"].tolist(), g["x1"].tolist(), g["size"].tolist(), g["info"].tolist()))
        wa = sorted(map(tuple, want[i].tolist()))
        assert ga == wa, i
    # 2. anchor rows follow seed.c:304-315 and their positions equal mb_bwt_sa_batch
    pos = gpu_ctx.to_host(out.pos, np.uint64, st.n_rows)
    prec = gpu_ctx.to_host(out.pos_rec, np.uint32, st.n_rows)
    assert st.n_rows == sum(len(sample_rows(int(r["x0"]), int(r["size"]))) for r in recs)
    rows = np.concatenate([np.array(sample_rows(int(recs[k]["x0"]), int(recs[k]["size"])), np.uint64) for k in range(len(recs))])
    assert np.array_equal(np.sort(ri.sa(rows)), np.sort(pos))
    first = {}
    for s in range(len(pos)):
        first.setdefault(int(prec[s]), []).append(int(pos[s]))
    pos_of_rec = {k: ri.sa(np.array([recs[k]["x0"]], np.uint64))[0] for k in range(len(recs))}
    for k, v in first.items():
        assert int(pos_of_rec[k]) in v
    # 3. jobs == the restated builder
    jobs = gpu_ctx.to_host(out.jobs, mbg.DP_JOB_DTYPE, st.n_jobs)
    meta = gpu_ctx.to_host(out.meta, mbg.JOB_META_DTYPE, st.n_jobs)
    qb = gpu_ctx.to_host(out.qbuf, np.uint8, st.q_bytes)
    tb = gpu_ctx.to_host(out.tbuf, np.uint8, st.t_bytes)
    exp, fast, anchored, fast_calls = expected_jobs(reads, recs, {k: int(v) for k, v in pos_of_rec.items()}, ref, [16569], opt)
    assert st.n_fast == fast and st.n_anchored == anchored
    got = {}
    for j in range(st.n_jobs):
        jb = jobs[j]
        got[(int(meta["rid"][j]), int(meta["side"][j]))] = (qb[jb["qoff"]:jb["qoff"] + jb["qlen"]], tb[jb["toff"]:jb["toff"] + jb["tlen"]],
                                                            int(jb["w"]), int(jb["flag"]))
    assert set(got) == set(exp)
    for k in exp:
        assert np.array_equal(got[k][0], exp[k][0]) and np.array_equal(got[k][1], exp[k][1]) and got[k][2:] == exp[k][2:], k
    assert st.n_jobs > 50
    # 4. DP results == ksw_extd2_sse on each job
    res = gpu_ctx.to_host(out.res, mbg.DP_RES_DTYPE, st.n_jobs)
    cig = gpu_ctx.to_host(out.cigar, np.uint32, int(jobs["cigar_off"].max() + jobs["qlen"].max() + jobs["tlen"].max() + 1))
    mat = np.array(list(opt.mat), np.int8)
    for j in range(st.n_jobs):
        jb = jobs[j]
        q, t = got[(int(meta["rid"][j]), int(meta["side"][j]))][:2]
        r = O.ref_extd2(q, t, mat, opt.q, opt.e, opt.q2, opt.e2, int(jb["w"]), opt.zdrop, opt.end_bonus, int(jb["flag"]))
        assert {f: int(res[f][j]) for f in FIELDS} == {f: r[f] for f in FIELDS}, j
        assert cig[jb["cigar_off"]:jb["cigar_off"] + res["n_cigar"][j]].tolist() == r["cigar"], j
    # 5. the reference's own mb_align_pair, given the uncapped band and the unscaled z-drop, makes the
    # same decisions: a single M for every settled extension, the same DP result for every job
    if O.have_ref_pe():
        p = O.align_params(opt.a, opt.b, 0, opt.q, opt.e, opt.q2, opt.e2, 0)
        assert opt.zdrop % opt.a == 0
        for q, t, flag in fast_calls:
            r = O.ref_align_pair(p, 0, q, t, mat, opt.ksw_w, opt.end_bonus, opt.zdrop // opt.a, flag)
            assert r["cigar"] == [len(q) << 4] and r["reach_end"] == 1
        for j in range(st.n_jobs):
            jb = jobs[j]
            q, t = got[(int(meta["rid"][j]), int(meta["side"][j]))][:2]
            r = O.ref_align_pair(p, 0, q, t, mat, opt.ksw_w, opt.end_bonus, opt.zdrop // opt.a, int(jb["flag"]))
            assert {f: int(res[f][j]) for f in FIELDS} == {f: r[f] for f in FIELDS}, j
            assert cig[jb["cigar_off"]:jb["cigar_off"] + res["n_cigar"][j]].tolist() == r["cigar"], j
        assert len(fast_calls) > 50
