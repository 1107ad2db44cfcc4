"""FM-index construction for synthetic references, writing the reference's own file formats.

Index construction is not on the hot path; this exists so a benchmark host can build a
chromosome-scale index in seconds on the GPU.  The output is byte-identical to ``minibwa index``
for an N-free reference (the BWT of a text is unique), so the hot-path kernels read the same bytes
either way:

* text = forward strand + reverse complement, '$' smaller than every base (index.c:40-105);
* suffix order by prefix doubling on 64-bit keys (torch.argsort; runs on cuda when available);
* rows 0..n with row 0 = the '$' suffix, BWT without '$' and ``primary`` its row (index.c:22-38);
* 64-byte occ blocks (bwt.c:60-109), SA sampled every 2^sa_bit rows with sa[0] = 2^64-1;
* ``.mbw`` (bwt.c:631-646) and ``.l2b`` (l2bit.c:243-273) writers.
"""
from __future__ import annotations

import struct

import numpy as np


def _torch():
    import torch
    return torch


def suffix_array(text: np.ndarray, device=None):
    """Suffix array of an nt4 text (values 0..3), end-of-text smaller than any base.  Returns a
    torch int64 tensor on `device`."""
    torch = _torch()
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    n = len(text)
    t = torch.as_tensor(np.ascontiguousarray(text, np.uint8), device=device).to(torch.int64)
    K = 21  # 3 bits per symbol: 0 = past the end, 1..4 = A..T
    key = torch.zeros(n, dtype=torch.int64, device=device)
    for j in range(K):
        c = torch.zeros(n, dtype=torch.int64, device=device)
        if j < n:
            c[: n - j] = t[j:] + 1
        key = (key << 3) | c
    del t
    sa = torch.argsort(key)
    sk = key[sa]
    grp = torch.zeros(n, dtype=torch.int64, device=device)
    if n > 1:
        grp[1:] = torch.cumsum((sk[1:] != sk[:-1]).to(torch.int64), 0)
    rank = torch.empty_like(grp)
    rank[sa] = grp
    del sk, key
    h = K
    while n > 0 and int(grp[-1]) < n - 1:
        r2 = torch.zeros(n, dtype=torch.int64, device=device)
        if h < n:
            r2[: n - h] = rank[h:] + 1
        key = rank * (n + 1) + r2
        del r2
        sa = torch.argsort(key)
        sk = key[sa]
        del key
        grp = torch.zeros(n, dtype=torch.int64, device=device)
        grp[1:] = torch.cumsum((sk[1:] != sk[:-1]).to(torch.int64), 0)
        del sk
        rank[sa] = grp
        h *= 2
    return sa


def bwt_from_sa(text: np.ndarray, sa, sa_bit: int = 4):
    """(primary, L2[5], data u64 array, ssa u64 array) in the reference's layout."""
    torch = _torch()
    device = sa.device
    n = len(text)
    t = torch.as_tensor(np.ascontiguousarray(text, np.uint8), device=device)
    full = torch.empty(n + 1, dtype=torch.int64, device=device)
    full[0] = n
    full[1:] = sa
    prim_rows = torch.nonzero(full == 0)
    primary = int(prim_rows[0, 0])
    prev = full - 1
    prev[primary] = 0
    bw = t[prev]
    B = torch.cat([bw[:primary], bw[primary + 1:]]).to(torch.int64)  # '$' removed, length n
    mask = (1 << sa_bit) - 1
    ssa = full[0::(mask + 1)].clone()
    ssa_np = ssa.cpu().numpy().astype(np.int64).view(np.uint64).copy()
    ssa_np[0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    nblk = (n + 127) // 128
    data = np.zeros(nblk * 8 + 4, np.uint64)
    L2 = [0]
    for c in range(4):
        cs = torch.cumsum((B == c).to(torch.int64), 0)
        tot = int(cs[-1]) if n else 0
        L2.append(L2[-1] + tot)
        starts = torch.arange(nblk, device=device, dtype=torch.int64) * 128
        before = torch.zeros(nblk, dtype=torch.int64, device=device)
        if nblk > 1:
            before[1:] = cs[starts[1:] - 1]
        mid = starts + 64
        has_mid = mid < n
        half = torch.zeros(nblk, dtype=torch.int64, device=device)
        idx = torch.nonzero(has_mid).flatten()
        if len(idx):
            half[idx] = cs[mid[idx] - 1] - before[idx]
        blk = (before | (half << 56)).cpu().numpy().view(np.uint64)
        data[c:nblk * 8:8] = blk
        data[nblk * 8 + c] = np.uint64(tot)
    pad = torch.zeros(nblk * 128, dtype=torch.int64, device=device)
    pad[:n] = B
    sh = torch.arange(32, dtype=torch.int64, device=device) * 2
    words = (pad.view(nblk, 4, 32) << sh).sum(dim=2)  # bit patterns; sign bit wraps like u64
    w = words.cpu().numpy().astype(np.int64).view(np.uint64)
    for k in range(4):
        data[4 + k:nblk * 8:8] = w[:, k]
    return primary, L2, data, ssa_np


def write_mbw(path, primary, L2, data, ssa, sa_bit=4):
    with open(path, "wb") as f:
        f.write(b"MBW\x02")
        f.write(struct.pack("<I", sa_bit))
        f.write(struct.pack("<Q", primary))
        f.write(struct.pack("<4Q", *[int(v) for v in L2[1:5]]))
        f.write(np.ascontiguousarray(data, np.uint64).tobytes())
        f.write(struct.pack("<Q", len(ssa)))
        f.write(np.ascontiguousarray(ssa, np.uint64).tobytes())


def write_l2b(path, ref: np.ndarray, ctg_len, names):
    """.l2b of an N-free, upper-case reference (no ambiguous or soft-masked runs)."""
    tot = int(np.sum(ctg_len))
    assert tot == len(ref)
    n_pac = (tot + 31) // 32
    padded = np.zeros(n_pac * 32, np.uint64)
    padded[:tot] = ref
    sh = np.arange(32, dtype=np.uint64) * np.uint64(2)
    pac = (padded.reshape(n_pac, 32) << sh).sum(axis=1, dtype=np.uint64)
    cat_name = b"".join(n.encode() + b"\0" for n in names)
    cat_comm = b"\0" * len(names)
    with open(path, "wb") as f:
        f.write(b"L2B\x01")
        f.write(struct.pack("<I", 0))
        f.write(struct.pack("<7Q", len(names), tot, 0, 0, len(cat_name), len(cat_comm), n_pac))
        f.write(np.asarray(ctg_len, np.uint64).tobytes())
        f.write(pac.tobytes())
        f.write(cat_name)
        f.write(cat_comm)
    return pac


def build(ref: np.ndarray, prefix: str, n_ctg: int = 1, ctg_prefix: str = "chr", sa_bit: int = 4, device=None):
    """Index an N-free nt4 reference split into n_ctg equal contigs named <ctg_prefix>1.. (the
    split synth.write_fasta uses).  Writes prefix.mbw and prefix.l2b; returns a dict of the parts."""
    ref = np.ascontiguousarray(ref, np.uint8)
    L = len(ref)
    per = (L + n_ctg - 1) // n_ctg
    ctg_len = np.array([min(per, L - c * per) for c in range(n_ctg)], np.uint64)
    names = [f"{ctg_prefix}{c + 1}" for c in range(n_ctg)]
    pac = write_l2b(prefix + ".l2b", ref, ctg_len, names)
    text = np.concatenate([ref, (3 - ref[::-1]).astype(np.uint8)])
    sa = suffix_array(text, device)
    primary, L2, data, ssa = bwt_from_sa(text, sa, sa_bit)
    del sa
    write_mbw(prefix + ".mbw", primary, L2, data, ssa, sa_bit)
    return {"primary": primary, "L2": L2, "data": data, "sa": ssa, "seq_len": len(text), "sa_bit": sa_bit,
            "ctg_len": ctg_len, "pac": pac}


def _main():
    import argparse
    import time
    ap = argparse.ArgumentParser(description="index a seeded synthetic reference (minibwa_amd.synth.reference)")
    ap.add_argument("--len", type=int, required=True)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--rep-frac", type=float, default=0.05)
    ap.add_argument("--rep-min", type=int, default=300)
    ap.add_argument("--rep-max", type=int, default=6000)
    ap.add_argument("--rep-div", type=float, default=0.02)
    ap.add_argument("--contigs", type=int, default=1)
    ap.add_argument("--ctg-prefix", default="chr")
    ap.add_argument("--device", default=None)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    from minibwa_amd import synth
    t0 = time.time()
    ref = synth.reference(a.len, seed=a.seed, rep_frac=a.rep_frac, rep_min=a.rep_min, rep_max=a.rep_max, rep_div=a.rep_div)
    r = build(ref, a.out, n_ctg=a.contigs, ctg_prefix=a.ctg_prefix, device=a.device)
    print(f"indexed {a.len} bp ({r['seq_len']} BWT rows) into {a.out}.mbw/.l2b in {time.time() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    _main()
