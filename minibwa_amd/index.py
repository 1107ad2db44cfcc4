"""FM-index construction for synthetic references, writing the reference's own file formats.

Index construction is not on the hot path; this exists so a benchmark host can build up to a
genome-scale (3.1 Gbp) index on the GPU in a few minutes.  The output is byte-identical to
``minibwa index`` for an N-free reference (the BWT of a text is unique), so the hot-path kernels
read the same bytes either way:

* text T = forward strand + reverse complement, '$' smaller than every base (index.c:40-105);
* suffix order: every suffix gets a 63-bit key of its first 21 characters (3 bits each, '$' = 0);
  the suffixes are split into key-range chunks of a bounded size, each chunk is sorted by key on
  the device, and runs of equal keys are refined by the next 21 characters until no two suffixes
  tie (suffixes of a diverged repeat tie for a few rounds; random ones almost never);
* rows 0..n with row 0 = the '$' suffix, BWT without '$' and ``primary`` its row (index.c:22-38);
* 64-byte occ blocks (bwt.c:60-109), SA sampled every 2^sa_bit rows with sa[0] = 2^64-1;
* ``.mbw`` (bwt.c:631-646) and ``.l2b`` (l2bit.c:243-273) writers.

Device memory is about 9 bytes per text position (the key array and the BWT codes) plus the
chunk: 3.1 Gbp (n = 6.2e9) needs about 60 GB.
"""
from __future__ import annotations

import struct

import numpy as np

KEY_CHARS = 21                  # 3 bits per character in an int64 key
BLOCK = 1 << 27                 # positions per elementwise pass (bounds temporaries)


def _torch():
    import torch
    return torch


def _dev(device):
    torch = _torch()
    if device is None:
        device = "cuda" if torch.cuda.is_available() else "cpu"
    return device


def suffix_keys(t3, n):
    """key[p] = characters T3[p .. p+20] packed MSB first, 3 bits each, zeros past the end.
    t3 holds n+1 codes (bases 1..4, then '$' = 0)."""
    torch = _torch()
    dev = t3.device
    tp = torch.cat([t3, torch.zeros(KEY_CHARS, dtype=torch.uint8, device=dev)])
    key = torch.empty(n + 1, dtype=torch.int64, device=dev)
    for b in range(0, n + 1, BLOCK):
        e = min(n + 1, b + BLOCK)
        k = torch.zeros(e - b, dtype=torch.int64, device=dev)
        for j in range(KEY_CHARS):
            k = (k << 3) | tp[b + j: e + j].to(torch.int64)
        key[b:e] = k
    return key


def _resolve_ties(key, pos, sk, n):
    """pos is sorted by the 21-character keys sk; order every run of equal keys by the following
    characters, 21 at a time, until no run is left."""
    torch = _torch()
    m = pos.numel()
    if m < 2:
        return pos
    eq = sk[1:] == sk[:-1]
    if not bool(eq.any()):
        return pos
    dev = pos.device
    tied = torch.zeros(m, dtype=torch.bool, device=dev)
    tied[1:] |= eq
    tied[:-1] |= eq
    start = torch.ones(m, dtype=torch.bool, device=dev)
    start[1:] = ~eq
    idx = torch.nonzero(tied).flatten()
    gid = torch.cumsum(start.to(torch.int64), 0)[idx]
    del eq, tied, start
    off = KEY_CHARS
    while idx.numel():
        p = pos[idx]
        nk = key[p + off]           # tied suffixes hold no '$' in their keys, so p + off <= n
        o1 = torch.argsort(nk, stable=True)
        perm = o1[torch.argsort(gid[o1], stable=True)]
        p, nk, gid = p[perm], nk[perm], gid[perm]
        pos[idx] = p                # groups occupy contiguous index ranges in increasing gid order
        same = (gid[1:] == gid[:-1]) & (nk[1:] == nk[:-1])
        if not bool(same.any()):
            break
        t2 = torch.zeros(idx.numel(), dtype=torch.bool, device=dev)
        t2[1:] |= same
        t2[:-1] |= same
        ns = torch.ones(idx.numel(), dtype=torch.bool, device=dev)
        ns[1:] = ~same
        keep = torch.nonzero(t2).flatten()
        gid = torch.cumsum(ns.to(torch.int64), 0)[keep]
        idx = idx[keep]
        off += KEY_CHARS
    return pos


def sorted_suffix_chunks(key, n, max_chunk):
    """Yield the suffix positions in suffix order, as consecutive device tensors of at most
    about max_chunk positions (one per range of the keys' top 16 bits)."""
    torch = _torch()
    dev = key.device
    sh = 3 * KEY_CHARS - 16
    hist = torch.zeros(1 << 16, dtype=torch.int64, device=dev)
    for b in range(0, n + 1, BLOCK):
        hist += torch.bincount(key[b:min(n + 1, b + BLOCK)] >> sh, minlength=1 << 16)
    hc = hist.cpu().numpy()
    bounds, lo, acc = [], 0, 0
    for i in range(1 << 16):
        if acc and acc + hc[i] > max_chunk:
            bounds.append((lo, i))
            lo, acc = i, 0
        acc += int(hc[i])
    bounds.append((lo, 1 << 16))
    for lo, hi in bounds:
        lo_k = lo << sh
        hi_k = hi << sh if hi < (1 << 16) else None          # keys use 63 bits: the last range is open
        parts = []
        for b in range(0, n + 1, BLOCK):
            kk = key[b:min(n + 1, b + BLOCK)]
            sel = kk >= lo_k if hi_k is None else (kk >= lo_k) & (kk < hi_k)
            parts.append(torch.nonzero(sel).flatten() + b)
        pos = torch.cat(parts)
        del parts
        if pos.numel() == 0:
            continue
        sk, order = torch.sort(key[pos])
        pos = pos[order]
        del order
        yield _resolve_ties(key, pos, sk, n)


def bwt_index(text: np.ndarray, sa_bit: int = 4, device=None, max_chunk: int = 400_000_000):
    """(primary, L2[5], data u64 array, ssa u64 array) of nt4 text (values 0..3) in the reference's
    layout: rows of T$ sorted, BWT without '$', 64-byte occ blocks, SA every 2^sa_bit rows."""
    torch = _torch()
    dev = _dev(device)
    n = len(text)
    t = torch.as_tensor(np.ascontiguousarray(text, np.uint8), device=dev)
    t3 = torch.empty(n + 1, dtype=torch.uint8, device=dev)
    t3[:n] = t + 1
    t3[n] = 0
    key = suffix_keys(t3, n)
    del t3
    mask = (1 << sa_bit) - 1
    bw = torch.empty(n + 1, dtype=torch.uint8, device=dev)          # BWT with '$' as 4
    ssa = torch.empty((n + 1 + mask) >> sa_bit, dtype=torch.int64, device=dev)
    row, primary = 0, -1
    for pos in sorted_suffix_chunks(key, n, max_chunk):
        m = pos.numel()
        prev = pos - 1
        z = torch.nonzero(pos == 0).flatten()
        if z.numel():
            primary = row + int(z[0])
            prev[z] = 0
        c = t[prev]
        if z.numel():
            c[z] = 4
        bw[row:row + m] = c
        first = (row + mask) >> sa_bit << sa_bit             # first sampled row in this chunk
        if first < row + m:
            ssa[first >> sa_bit:((row + m - 1) >> sa_bit) + 1] = pos[first - row::mask + 1]
        row += m
        del pos, prev, c
    assert row == n + 1 and primary >= 0
    del key
    B = torch.cat([bw[:primary], bw[primary + 1:]])                # '$' removed, length n
    del bw
    ssa_np = ssa.cpu().numpy().view(np.uint64).copy()
    ssa_np[0] = np.uint64(0xFFFFFFFFFFFFFFFF)
    nblk = (n + 127) // 128
    pad = torch.full((nblk * 128,), 4, dtype=torch.uint8, device=dev)
    pad[:n] = B
    del B
    data = np.zeros(nblk * 8 + 4, np.uint64)
    L2 = [0]
    sh = torch.arange(32, dtype=torch.int64, device=dev) * 2
    tot = [0, 0, 0, 0]
    before = [0, 0, 0, 0]
    bs = BLOCK // 128 * 128                                       # positions per pass, whole blocks
    for b in range(0, nblk * 128, bs):
        e = min(nblk * 128, b + bs)
        blk = pad[b:e].view(-1, 128)
        nb = blk.shape[0]
        b0 = b // 128
        # the first-64 count is only recorded when the block holds more than 64 symbols (bwt.c:87-89)
        has_mid = (torch.arange(b0, b0 + nb, dtype=torch.int64, device=dev) * 128 + 64) < n
        for c in range(4):
            eq = (blk == c)
            full = eq.sum(dim=1, dtype=torch.int64)
            half = eq[:, :64].sum(dim=1, dtype=torch.int64) * has_mid
            cs = torch.cumsum(full, 0)
            bef = cs - full + before[c]
            before[c] += int(cs[-1])
            data[b0 * 8 + c:(b0 + nb) * 8:8] = (bef | (half << 56)).cpu().numpy().view(np.uint64)
        codes = blk.to(torch.int64) & 3                            # padding (4) encodes as 0
        words = (codes.view(nb, 4, 32) << sh).sum(dim=2)           # bit patterns; sign bit wraps like u64
        w = words.cpu().numpy().astype(np.int64).view(np.uint64)
        for k in range(4):
            data[b0 * 8 + 4 + k:(b0 + nb) * 8:8] = w[:, k]
    for c in range(4):
        tot[c] = before[c]
        L2.append(L2[-1] + tot[c])
        data[nblk * 8 + c] = np.uint64(tot[c])
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
    pac = np.zeros(n_pac, np.uint64)
    for b in range(0, n_pac, 1 << 24):                             # 2^29 bases per pass
        e = min(n_pac, b + (1 << 24))
        seg = np.zeros((e - b) * 32, np.uint64)
        part = ref[b * 32: min(tot, e * 32)]
        seg[:len(part)] = part
        sh = np.arange(32, dtype=np.uint64) * np.uint64(2)
        pac[b:e] = (seg.reshape(e - b, 32) << sh).sum(axis=1, dtype=np.uint64)
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


def build(ref: np.ndarray, prefix: str, n_ctg: int = 1, ctg_prefix: str = "chr", sa_bit: int = 4, device=None,
          max_chunk: int = 400_000_000):
    """Index an N-free nt4 reference split into n_ctg equal contigs named <ctg_prefix>1.. (the
    split synth.write_fasta uses).  Writes prefix.mbw and prefix.l2b; returns a dict of the parts."""
    ref = np.ascontiguousarray(ref, np.uint8)
    L = len(ref)
    per = (L + n_ctg - 1) // n_ctg
    ctg_len = np.array([min(per, L - c * per) for c in range(n_ctg)], np.uint64)
    names = [f"{ctg_prefix}{c + 1}" for c in range(n_ctg)]
    pac = write_l2b(prefix + ".l2b", ref, ctg_len, names)
    text = np.empty(2 * L, np.uint8)
    text[:L] = ref
    text[L:] = 3 - ref[::-1]
    primary, L2, data, ssa = bwt_index(text, sa_bit, device, max_chunk)
    del text
    write_mbw(prefix + ".mbw", primary, L2, data, ssa, sa_bit)
    return {"primary": primary, "L2": L2, "data": data, "sa": ssa, "seq_len": 2 * L, "sa_bit": sa_bit,
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
    ap.add_argument("--max-chunk", type=int, default=400_000_000)
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    from minibwa_amd import synth
    t0 = time.time()
    ref = synth.reference(a.len, seed=a.seed, rep_frac=a.rep_frac, rep_min=a.rep_min, rep_max=a.rep_max, rep_div=a.rep_div)
    print(f"reference {a.len} bp in {time.time() - t0:.1f}s", flush=True)
    r = build(ref, a.out, n_ctg=a.contigs, ctg_prefix=a.ctg_prefix, device=a.device, max_chunk=a.max_chunk)
    print(f"indexed {a.len} bp ({r['seq_len']} BWT rows) into {a.out}.mbw/.l2b in {time.time() - t0:.1f}s", flush=True)


if __name__ == "__main__":
    _main()
