"""Python host mirror of the reference's hot-path interface, bound to libmbg.so (include/mbg.h).

The names follow the reference functions each call replaces:

==========================  =======================================================
``Context.seed_intv_batch``  ``mb_seed_intv_batch`` (seed.c:56-103)
``Context.bwt_sa_batch``     ``mb_bwt_sa_batch`` (bwt.c:511-541)
``Context.bwt_rank2a``       ``mb_bwt_rank2a`` (bwt.c:195-231)
``Context.bwt_extend``       ``mb_bwt_extend`` (bwt.c:237-250)
``Context.ksw_extd2``        ``ksw_extd2_sse`` (ksw2_extd2_sse.c:27-403), batched
``Context.ksw_extz2``        ``ksw_extz2_sse`` (ksw2_extz2_sse.c:16-296), batched
``Context.l2b_getseq``       ``l2b_getseq`` (l2bit.c:52-77), batched
``Context.idx_load``         ``mb_idx_load`` (map-algo.c:19-42) minus chaining state
==========================  =======================================================

There is no CPU fallback: constructing a Context without the built library or without a GPU
raises.  Errors from the library raise ``MbgError`` with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBMBG = os.path.join(HERE, "libmbg.so")

MBG_EOVERFLOW = -8


class MbgError(RuntimeError):
    pass


class Sai(C.Structure):
    _fields_ = [("x0", C.c_uint64), ("x1", C.c_uint64), ("size", C.c_uint64), ("info", C.c_uint64)]


class BwtDesc(C.Structure):
    _fields_ = [("primary", C.c_uint64), ("L2", C.c_uint64 * 5), ("seq_len", C.c_uint64),
                ("data", C.POINTER(C.c_uint64)), ("sa_bit", C.c_uint32), ("n_sa", C.c_uint64),
                ("sa", C.POINTER(C.c_uint64))]


class L2bDesc(C.Structure):
    _fields_ = [("tot_len", C.c_uint64), ("n_ctg", C.c_uint64), ("ctg_len", C.POINTER(C.c_uint64)),
                ("n_pac", C.c_uint64), ("pac", C.POINTER(C.c_uint64)), ("n_ambi", C.c_uint64),
                ("ambi", C.POINTER(C.c_uint64))]


class IdxInfo(C.Structure):
    _fields_ = [("seq_len", C.c_uint64), ("primary", C.c_uint64), ("tot_len", C.c_uint64), ("n_ctg", C.c_uint64),
                ("pre_len", C.c_int32), ("sa_bit", C.c_int32), ("device_bytes", C.c_uint64)]


class Counters(C.Structure):       # mbg_counters_t
    _fields_ = ([(f, C.c_int64) for f in ("fronts", "front_reads", "front_records", "front_slots")] +
                [(f, C.c_double) for f in ("ms_seed1", "ms_seed2", "ms_group", "ms_anchor", "ms_chain", "ms_front_copy", "ms_front")] +
                [(f, C.c_uint64) for f in ("seed1_occ_blocks", "seed1_kmer_lookups")] +
                [(f, C.c_int64) for f in ("dp_launches", "dp_jobs", "dp_cells")] + [("ms_dp", C.c_double)] +
                [(f, C.c_int64) for f in ("ll_launches", "ll_jobs", "ll_cells")] + [("ms_ll", C.c_double)] +
                [(f, C.c_int64) for f in ("graph_launches", "graph_instantiations", "direct_chains")])


class AlignOpt(C.Structure):       # mbg_align_opt_t
    _fields_ = [(f, C.c_int32) for f in ("a", "b", "b_ts", "q", "e", "q2", "e2", "meth")] + [("max_sw_mat", C.c_int64)]


class MapOpt(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("min_len", "max_sub_occ", "max_occ", "a", "b", "q", "e", "q2", "e2",
                                         "end_bonus", "ksw_w", "zdrop")] + [("mat", C.c_int8 * 25)]


class MapStats(C.Structure):
    _fields_ = ([(n, C.c_int64) for n in ("n_reads", "n_rec", "n_rows", "n_anchored", "n_fast", "n_jobs", "q_bytes",
                                          "t_bytes", "cells")] +
                [("occ_blocks", C.c_uint64 * 2), ("kmer_lookups", C.c_uint64 * 2), ("extend_steps", C.c_uint64 * 2),
                 ("lf_steps", C.c_uint64)] +
                [(n, C.c_float) for n in ("ms_seed1", "ms_seed2", "ms_rows", "ms_sa", "ms_jobs", "ms_dp", "ms_total")])

    def as_dict(self):
        d = {}
        for n, t in self._fields_:
            v = getattr(self, n)
            d[n] = list(v) if hasattr(v, "__len__") else v
        return d


class MapOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("recs", "pos", "pos_rec", "jobs", "meta", "res", "cigar", "qbuf", "tbuf")]


def adaptive_opt(read_len: int, **over) -> MapOpt:
    """minibwa's defaults (options.c:5-88, preset "adap") with the length-adaptive band and z-drop
    (mb_opt_adap, options.c:90-116) turned into ksw arguments (align.c:564, 374-376)."""
    import math
    o = dict(min_len=19, max_sub_occ=10, max_occ=250, a=2, b=8, q=12, e=2, q2=23, e2=1, end_bonus=10, bw=100,
             zdrop=80, b_ts=0, b_ambi=1)
    o.update(over)
    a_ = -math.log(0.5) / (2000 - 100)
    b_ = math.exp(-a_ * ((read_len if read_len > 100 else 100) - 100))
    bw = int(500 - (500 - o["bw"]) * b_ + .499) if o["bw"] < 500 else o["bw"]
    zd = int(400 - (400 - o["zdrop"]) * b_ + .499) if o["zdrop"] < 400 else o["zdrop"]
    m = MapOpt()
    for k in ("min_len", "max_sub_occ", "max_occ", "a", "b", "q", "e", "q2", "e2", "end_bonus"):
        setattr(m, k, int(o[k]))
    m.ksw_w = int(bw * 1.5 + 1.)
    m.zdrop = zd * o["a"]
    mat = nt4_mat(o["a"], o["b"], o["b_ts"], o["b_ambi"])
    for i in range(25):
        m.mat[i] = int(mat[i])
    return m


def chain_opt(read_len: int, rechain: bool = False) -> "ChainOpt":
    """mb_lchain_dp's arguments for a read of read_len bases as mb_map_sai passes them (map-algo.c:607-616,
    634): minibwa's defaults (preset "adap": bw 100, max_gap 100, min_chain_score 25, max_chain_skip 25,
    max_chain_iter 5000, chain_gap_scale 0.8, min_len 19, bw_long 20000) through mb_opt_adap
    (options.c:90-116); rechain=True gives the re-chaining band bw_long."""
    import math
    a_ = -math.log(0.5) / (2000 - 100)
    b_ = math.exp(-a_ * ((read_len if read_len > 100 else 100) - 100))
    max_gap = int(5000 - (5000 - 100) * b_ + .499)
    if max_gap > read_len:
        max_gap = read_len
    bw = int(500 - (500 - 100) * b_ + .499)
    bw_long = min(20000, read_len * 5)
    bw_long = max(bw_long, bw)
    min_sc = int(40 - (40 - 25) * b_ + .499)
    pen = C.c_float(0.8).value * .01 * 19               # chain_gap_scale (float) * .01 * min_len, in double
    return ChainOpt(max_dist_x=max_gap, max_dist_y=max_gap, bw=bw_long if rechain else bw, max_skip=25, max_iter=5000,
                    min_sc=min_sc, pen_gap=pen)


def nt4_mat(a=2, b=8, b_ts=0, b_ambi=1):
    """ksw_gen_nt4_mat (ksw2.h:199-219) without the methylation cases."""
    a, b, b_ambi = abs(a), -abs(b), -abs(b_ambi)
    mat = np.zeros(25, np.int8)
    for i in range(4):
        for j in range(4):
            mat[i * 5 + j] = a if i == j else b
        mat[i * 5 + 4] = b_ambi
    mat[20:25] = b_ambi
    if b_ts != 0 and b_ts != b:
        bt = -abs(b_ts)
        mat[2] = mat[8] = mat[10] = mat[16] = bt
    return mat


DP_JOB_DTYPE = np.dtype([("qoff", "<i8"), ("toff", "<i8"), ("qlen", "<i4"), ("tlen", "<i4"),
                         ("q", "i1"), ("e", "i1"), ("q2", "i1"), ("e2", "i1"),
                         ("w", "<i4"), ("zdrop", "<i4"), ("end_bonus", "<i4"), ("flag", "<i4"),
                         ("pad", "<i4"), ("cigar_off", "<i8")])
assert DP_JOB_DTYPE.itemsize == 56
DP_RES_FIELDS = ("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score", "reach_end",
                 "n_cigar", "cells")
DP_RES_DTYPE = np.dtype([(f, "<i4") for f in DP_RES_FIELDS])
SEED_REC_DTYPE = np.dtype([("x0", "<u8"), ("x1", "<u8"), ("size", "<u8"), ("info", "<u8"), ("rid", "<u4"),
                           ("pass", "<u4")])

_P = C.c_void_p
_LIB = None


def lib():
    """Load libmbg.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBMBG):
            raise MbgError(f"{LIBMBG} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIBMBG)
        L.mbg_strerror.restype = C.c_char_p
        L.mbg_strerror.argtypes = [C.c_int]
        L.mbg_last_hip_error.restype = C.c_char_p
        L.mbg_init.argtypes = [C.c_int, C.POINTER(_P)]
        L.mbg_destroy.argtypes = [_P]
        L.mbg_idx_load.argtypes = [_P, C.c_char_p, C.c_int32]
        L.mbg_bwt_upload.argtypes = [_P, C.POINTER(BwtDesc), C.c_int32]
        L.mbg_l2b_upload.argtypes = [_P, C.POINTER(L2bDesc)]
        L.mbg_idx_info.argtypes = [_P, C.POINTER(IdxInfo)]
        L.mbg_idx_attach.argtypes = [_P, _P]
        L.mbg_pre_copy.argtypes = [_P, _P, C.c_uint64]
        L.mbg_rank2a.argtypes = [_P, C.c_int64, _P, _P, _P, _P]
        L.mbg_extend.argtypes = [_P, C.c_int64, _P, _P, _P]
        L.mbg_seed_batch.argtypes = [_P, C.c_int32, _P, _P, C.c_int32, C.c_int32, C.POINTER(_P), _P]
        L.mbg_sa_batch.argtypes = [_P, C.c_int64, _P]
        for fn in (L.mbg_extd2_batch, L.mbg_extz2_batch):
            fn.argtypes = [_P, C.c_int32, _P, _P, C.c_int64, _P, C.c_int64, C.c_int8, _P, _P, _P, C.c_int64]
        L.mbg_align_pair_batch.argtypes = [_P, C.POINTER(AlignOpt), C.c_int32, _P, _P, C.c_int64, _P, C.c_int64, _P, _P, _P,
                                           C.c_int64]
        L.mbg_getseq.argtypes = [_P, C.c_int64, _P, _P, _P, _P, _P, C.c_int64, _P]
        L.mbg_use_graphs.argtypes = [_P, C.c_int]
        L.mbg_counters.argtypes = [_P, C.POINTER(Counters), C.c_int]
        L.mbg_count_work.argtypes = [_P, C.c_int]
        L.mbg_seed_batch_dev.argtypes = [_P, _P, C.c_int32, _P, _P, _P, C.c_int32, C.c_int32, _P, C.c_int64,
                                         C.POINTER(C.c_int64)]
        L.mbg_seed_batch_dev_async.argtypes = [_P, _P, C.c_int32, _P, _P, _P, C.c_int32, C.c_int32, _P, C.c_int64, _P]
        L.mbg_sa_batch_dev.argtypes = [_P, _P, C.c_int64, _P]
        L.mbg_sa_batch_dev_noalloc.argtypes = [_P, _P, C.c_int64, _P]
        L.mbg_copy_to_host.argtypes = [_P, _P, _P, C.c_size_t]
        L.mbg_copy_to_device.argtypes = [_P, _P, _P, C.c_size_t]
        L.mbg_dev_alloc.argtypes = [_P, C.c_size_t, C.POINTER(_P)]
        L.mbg_dev_free.argtypes = [_P, _P]
        L.mbg_stream_copy.argtypes = [_P, _P, _P, _P, C.c_size_t]
        L.mbg_gather_bench.argtypes = [_P, C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_uint64)]
        L.mbg_map_step_dev.argtypes = [_P, _P, C.c_int32, _P, _P, _P, C.c_int32, C.POINTER(MapOpt), C.c_int,
                                       C.POINTER(MapStats), C.POINTER(MapOut)]
        L.mbg_extd2_batch_dev.argtypes = [_P, _P, C.c_int32, _P, _P, _P, C.c_int8, _P, _P, _P, C.c_int32, C.c_int32,
                                          C.c_int32]
        L.mbg_ll_batch.argtypes = [_P, C.c_int32, _P, _P, C.c_int64, _P, C.c_int64, C.c_int8, _P, C.c_int32, C.c_int32,
                                   C.c_int32, _P]
        L.mbg_matesw_batch.argtypes = [_P, C.c_int32, _P, _P, C.c_int64, C.POINTER(MateswOpt), _P, _P, C.c_int64,
                                       C.POINTER(MateswStats)]
        L.mbg_anchor_batch.argtypes = [_P, C.c_int32, _P, _P, _P, C.c_int32, _P, _P, _P, _P]
        L.mbg_anchor_batch_dev.argtypes = [_P, _P, C.c_int32, C.c_int64, _P, _P, _P, C.c_int32, _P, _P, _P, _P, _P]
        L.mbg_lchain_batch.argtypes = [_P, C.c_int32, _P, _P, _P, C.POINTER(ChainOpt), _P, _P, _P, _P]
        L.mbg_lchain_batch_dev.argtypes = [_P, _P, C.c_int32, C.c_int64, _P, _P, _P, C.POINTER(ChainOpt), _P, C.c_size_t, _P,
                                           _P, _P, _P]
        L.mbg_fq_open.argtypes = [C.c_char_p, C.c_int32, C.POINTER(_P)]
        L.mbg_fq_close.argtypes = [_P]
        L.mbg_fq_close.restype = None
        L.mbg_fq_buf_new.argtypes = [C.POINTER(_P)]
        L.mbg_fq_buf_free.argtypes = [_P]
        L.mbg_fq_buf_free.restype = None
        L.mbg_fq_read.argtypes = [_P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int64, _P,
                                  C.POINTER(FqChunk)]
        L.mbg_fq_read_frag.argtypes = [C.c_int32, C.POINTER(_P), C.c_int64, C.c_int32, C.c_int32, _P, C.POINTER(FqChunk)]
        L.mbg_update_extra_batch.argtypes = [_P, C.c_int64, _P, _P, C.c_int64, _P, C.c_int64, _P, C.c_int8, C.c_int8,
                                             C.c_int32, _P, C.c_int64]
        L.mbg_ll_batch_dev.argtypes = [_P, _P, C.c_int32, _P, _P, _P, C.c_int8, _P, C.c_int32, C.c_int32, C.c_int32, _P,
                                       C.c_int32, C.c_int64]
        _LIB = L
    return _LIB


EXPORTED = ("mbg_strerror", "mbg_last_hip_error", "mbg_init", "mbg_destroy", "mbg_idx_load", "mbg_bwt_upload",
            "mbg_l2b_upload", "mbg_idx_info", "mbg_pre_copy", "mbg_rank2a", "mbg_extend", "mbg_seed_batch",
            "mbg_sa_batch", "mbg_extd2_batch", "mbg_extz2_batch", "mbg_align_pair_batch", "mbg_getseq", "mbg_seed_batch_dev",
            "mbg_seed_batch_dev_async", "mbg_sa_batch_dev", "mbg_extd2_batch_dev", "mbg_map_step_dev", "mbg_copy_to_host", "mbg_copy_to_device", "mbg_dev_alloc",
            "mbg_dev_free", "mbg_stream_copy", "mbg_gather_bench", "mbg_idx_attach", "mbg_ll_batch",
            "mbg_ll_batch_dev", "mbg_matesw_batch", "mbg_update_extra_batch",
            "mbg_anchor_batch", "mbg_anchor_batch_dev", "mbg_lchain_batch", "mbg_lchain_batch_dev", "mbg_fq_open",
            "mbg_fq_close", "mbg_fq_buf_new", "mbg_fq_buf_free", "mbg_fq_read", "mbg_fq_read_frag", "mbg_use_graphs",
            "mbg_counters", "mbg_count_work")
ALIGN_JOB_DTYPE = np.dtype([("qoff", "<i8"), ("toff", "<i8"), ("qlen", "<i4"), ("tlen", "<i4"), ("w", "<i4"),
                            ("end_bonus", "<i4"), ("zdrop", "<i4"), ("flag", "<i4"), ("cigar_off", "<i8")], align=True)
LL_JOB_DTYPE = np.dtype([("qoff", "<i8"), ("toff", "<i8"), ("qlen", "<i4"), ("tlen", "<i4"), ("size", "<i4"),
                         ("pad", "<i4")], align=True)
LL_RES_DTYPE = np.dtype([(f, "<i4") for f in ("score", "te", "qe", "score2", "te2")])
assert LL_JOB_DTYPE.itemsize == 32 and LL_RES_DTYPE.itemsize == 20
KSW_LL_STOP = 0x20000
KSW_LL_SUBO = 0x40000
MATESW_JOB_DTYPE = np.dtype([("qoff", "<i8"), ("qlen", "<i4"), ("tid", "<i4"), ("ts", "<i8"), ("te", "<i8"),
                             ("min_sc", "<i4"), ("pad", "<i4"), ("cigar_off", "<i8")])
MATESW_RES_DTYPE = np.dtype([(f, "<i4") for f in ("max_ug", "max_i", "n_good", "n_kmer", "ts2", "te2", "aligned",
                                                    "rescued")] +
                            [("ll", LL_RES_DTYPE), ("ez", DP_RES_DTYPE)] +
                            [(f, "<i4") for f in ("qs", "qe", "ts", "te", "dp_max2", "ran_dp", "n_cigar", "blen", "mlen",
                                                  "n_ambi", "dp_max")])
assert MATESW_JOB_DTYPE.itemsize == 48 and MATESW_RES_DTYPE.itemsize == 144


EXTRA_JOB_DTYPE = np.dtype([("qoff", "<i8"), ("toff", "<i8"), ("qs", "<i4"), ("qe", "<i4"), ("ts", "<i8"), ("te", "<i8"),
                            ("rev", "<i4"), ("n_cigar", "<i4"), ("cigar_off", "<i8"), ("blen", "<i4"), ("mlen", "<i4"),
                            ("n_ambi", "<i4"), ("dp_max", "<i4")])
assert EXTRA_JOB_DTYPE.itemsize == 72


# mbg_anchor_t = mb_anchor_t (mbpriv.h:26-32)
ANCHOR_DTYPE = np.dtype([("sid", "<i4"), ("len", "<i4"), ("qpos", "<i4"), ("flag", "<u4"), ("tpos", "<i8")])
assert ANCHOR_DTYPE.itemsize == 24


def anchor_slots(intervals, max_occ):
    """Per read, the anchor slots mb_anchor preallocates (seed.c:292-293): the sum of min(size, max_occ)."""
    m = max(int(max_occ), 0)
    return np.array([int(np.minimum(np.asarray(u, np.uint64).reshape(-1, 4)[:, 2], m).sum()) for u in intervals], np.int64)


class FqChunk(C.Structure):
    """mbg_fq_chunk_t: views into a chunk's flat arrays."""
    _P8 = C.POINTER(C.c_uint8)
    _fields_ = [("n", C.c_int32), ("status", C.c_int32)] + \
               [(f, C.POINTER(C.c_int64)) for f in ("name_off", "com_off", "seq_off")] + \
               [(f, C.POINTER(C.c_char)) for f in ("name", "comment", "seq", "qual")] + \
               [("has_qual", C.POINTER(C.c_uint8)), ("nt4", C.POINTER(C.c_uint8))]


class FastxReader:
    """A FASTA/FASTQ file read in mb_bseq_read's chunks (bseq.c:85-161) by the library's reader; host only,
    no device needed."""

    def __init__(self, path, n_threads=4):
        self.L = lib()
        h = C.c_void_p()
        check(self.L.mbg_fq_open(path.encode() if path else None, n_threads, C.byref(h)), "mbg_fq_open")
        self.h = h
        b = C.c_void_p()
        check(self.L.mbg_fq_buf_new(C.byref(b)), "mbg_fq_buf_new")
        self.buf = b

    @staticmethod
    def _records(ch):
        n = ch.n
        if n == 0:
            return [], np.zeros(0, np.uint8), ch.status

        def arr(p, k, dt=np.int64):
            return np.ctypeslib.as_array(p, shape=(k,)).astype(dt, copy=True) if k else np.zeros(0, dt)
        no, co, so = arr(ch.name_off, n + 1), arr(ch.com_off, n + 1), arr(ch.seq_off, n + 1)
        names, coms = C.string_at(ch.name, int(no[-1])), C.string_at(ch.comment, int(co[-1]))
        seqs = C.string_at(ch.seq, int(so[-1]))
        quals = C.string_at(ch.qual, int(so[-1]))
        hq = arr(ch.has_qual, n, np.uint8)
        nt4 = np.frombuffer(C.string_at(ch.nt4, int(so[-1])), np.uint8).copy()
        recs = []
        for i in range(n):
            c = coms[co[i]:co[i + 1]]
            recs.append((names[no[i]:no[i + 1]], seqs[so[i]:so[i + 1]], quals[so[i]:so[i + 1]] if hq[i] else None,
                         c if c else None))
        return recs, nt4, ch.status

    def read(self, chunk_size, with_qual=True, with_comment=True, frag_mode=False, min_cnt=0, max_chunk_size=0):
        """-> ([(name, seq, qual|None, comment|None) bytes], nt4 codes of the chunk's bases, status)."""
        ch = FqChunk()
        check(self.L.mbg_fq_read(self.h, chunk_size, int(with_qual), int(with_comment), int(frag_mode), min_cnt,
                                 max_chunk_size, self.buf, C.byref(ch)), "mbg_fq_read")
        return self._records(ch)

    @staticmethod
    def read_frag(readers, chunk_size, with_qual=True, with_comment=True):
        L = readers[0].L
        hs = (C.c_void_p * len(readers))(*[r.h for r in readers])
        ch = FqChunk()
        check(L.mbg_fq_read_frag(len(readers), hs, chunk_size, int(with_qual), int(with_comment), readers[0].buf,
                                 C.byref(ch)), "mbg_fq_read_frag")
        return FastxReader._records(ch)

    def close(self):
        if getattr(self, "h", None):
            self.L.mbg_fq_close(self.h)
            self.L.mbg_fq_buf_free(self.buf)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ChainOpt(C.Structure):
    """mbg_chain_opt_t: mb_lchain_dp's parameters (lchain.c:145)."""
    _fields_ = [(f, C.c_int32) for f in ("max_dist_x", "max_dist_y", "bw", "max_skip", "max_iter", "min_sc")] + \
               [("pen_gap", C.c_float), ("pad", C.c_int32)]


def lchain_work_bytes(n_slots):
    """MBG_LCHAIN_WORK_BYTES (include/mbg.h)."""
    return int(n_slots) * 72 + 2048


class MateswStats(C.Structure):
    _fields_ = [(f, C.c_int64) for f in ("n_aligned", "n_rescued", "ll_cells", "dp_cells")] + \
               [(f, C.c_float) for f in ("ms_window", "ms_ungap", "ms_ll", "ms_dp", "ms_finish", "ms_total")]

    def as_dict(self):
        return {f: (round(getattr(self, f), 3) if f.startswith("ms") else int(getattr(self, f))) for f, _ in self._fields_}


class MateswOpt(C.Structure):
    """mbg_matesw_opt_t: the mb_opt_t fields mate rescue reads, and the batch's methylation mode."""
    _fields_ = [(f, C.c_int32) for f in ("a", "b", "b_ts", "b_ambi", "q", "e", "bw", "zdrop", "end_bonus", "min_dp_max",
                                         "min_len", "mt")]
JOB_META_DTYPE = np.dtype([("rid", "<u4"), ("side", "i1"), ("rev", "i1"), ("pad", "<i2"), ("tid", "<i4"), ("tpos", "<i8")],
                          align=True)
assert JOB_META_DTYPE.itemsize == 24


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


def check(rc: int, what: str = ""):
    if rc != 0:
        L = lib()
        msg = L.mbg_strerror(rc).decode()
        if rc == -2:
            msg += ": " + L.mbg_last_hip_error().decode()
        raise MbgError(f"{what}: {msg} ({rc})")


def pack_seqs(seqs):
    """Concatenate nt4 reads (a list, or a 2-D array of equal-length ones) -> (bytes, int64 offsets[n+1],
    int32 lengths[n])."""
    if isinstance(seqs, np.ndarray) and seqs.ndim == 2:
        n, rl = seqs.shape
        return (np.ascontiguousarray(seqs, np.uint8).reshape(-1), np.arange(n + 1, dtype=np.int64) * rl,
                np.full(n, rl, np.int32))
    lens = np.array([len(s) for s in seqs], np.int32)
    off = np.zeros(len(seqs) + 1, np.int64)
    np.cumsum(lens, out=off[1:])
    buf = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if len(seqs) else np.zeros(0, np.uint8)
    return np.ascontiguousarray(buf), off, lens


@dataclass
class DpBatchResult:
    res: np.ndarray       # DP_RES_DTYPE per job
    cigar: np.ndarray     # uint32 slots; job i's CIGAR at cigar_off[i] : cigar_off[i] + res[i].n_cigar
    cigar_off: np.ndarray

    def cigar_of(self, i):
        o = int(self.cigar_off[i])
        return [int(c) for c in self.cigar[o:o + int(self.res["n_cigar"][i])]]


class DeviceArray:
    """A device allocation made through libmbg (no other GPU runtime in the process)."""

    def __init__(self, ctx, ptr, dtype, shape):
        self.ctx, self.ptr, self.dtype, self.shape = ctx, ptr, np.dtype(dtype), tuple(shape)

    def numpy(self):
        return self.ctx.to_host(self.ptr, self.dtype, int(np.prod(self.shape))).reshape(self.shape)

    def free(self):
        if self.ptr and getattr(self.ctx, "h", None):
            check(self.ctx.L.mbg_dev_free(self.ctx.h, C.c_void_p(self.ptr)), "mbg_dev_free")
        self.ptr = None


class Context:
    """A device context holding one resident index (HBM) and the library's scratch pools."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        check(self.L.mbg_init(device, C.byref(h)), "mbg_init")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.mbg_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- index -------------------------------------------------------------------------------
    def idx_load(self, prefix: str, pre_len: int = 10):
        check(self.L.mbg_idx_load(self.h, prefix.encode(), pre_len), f"mbg_idx_load({prefix})")
        return self.info()

    def idx_attach(self, src: "Context"):
        """Serve from the index resident in `src` (which must outlive this context)."""
        check(self.L.mbg_idx_attach(self.h, src.h), "mbg_idx_attach")

    def bwt_upload(self, primary, L2, seq_len, data: np.ndarray, sa_bit, sa: np.ndarray, pre_len=10):
        data = np.ascontiguousarray(data, np.uint64)
        sa = np.ascontiguousarray(sa, np.uint64)
        d = BwtDesc(int(primary), (C.c_uint64 * 5)(*[int(v) for v in L2]), int(seq_len),
                    data.ctypes.data_as(C.POINTER(C.c_uint64)), int(sa_bit), len(sa), sa.ctypes.data_as(C.POINTER(C.c_uint64)))
        check(self.L.mbg_bwt_upload(self.h, C.byref(d), pre_len), "mbg_bwt_upload")

    def l2b_upload(self, ctg_len, pac: np.ndarray, ambi: np.ndarray | None = None):
        ctg_len = np.ascontiguousarray(ctg_len, np.uint64)
        pac = np.ascontiguousarray(pac, np.uint64)
        ambi = np.zeros((0, 2), np.uint64) if ambi is None else np.ascontiguousarray(ambi, np.uint64)
        d = L2bDesc(int(ctg_len.sum()), len(ctg_len), ctg_len.ctypes.data_as(C.POINTER(C.c_uint64)), len(pac),
                    pac.ctypes.data_as(C.POINTER(C.c_uint64)), len(ambi), ambi.ctypes.data_as(C.POINTER(C.c_uint64)))
        check(self.L.mbg_l2b_upload(self.h, C.byref(d)), "mbg_l2b_upload")

    def info(self) -> dict:
        i = IdxInfo()
        check(self.L.mbg_idx_info(self.h, C.byref(i)), "mbg_idx_info")
        return {f: getattr(i, f) for f, _ in IdxInfo._fields_}

    def pre_cache(self) -> np.ndarray:
        n = 1 << (2 * self.info()["pre_len"])
        out = np.zeros((n, 4), np.uint64)
        check(self.L.mbg_pre_copy(self.h, _ptr(out), n), "mbg_pre_copy")
        return out

    # ---- FM-index primitives -----------------------------------------------------------------
    def bwt_rank2a(self, k, l):
        k = np.ascontiguousarray(k, np.uint64)
        l = np.ascontiguousarray(l, np.uint64)
        ck = np.zeros((len(k), 4), np.uint64)
        cl = np.zeros((len(k), 4), np.uint64)
        check(self.L.mbg_rank2a(self.h, len(k), _ptr(k), _ptr(l), _ptr(ck), _ptr(cl)), "mbg_rank2a")
        return ck, cl

    def bwt_extend(self, ik: np.ndarray, is_back: np.ndarray):
        ik = np.ascontiguousarray(ik, np.uint64).reshape(-1, 4)
        ib = np.ascontiguousarray(is_back, np.uint8)
        ok = np.zeros((len(ik), 4, 4), np.uint64)
        check(self.L.mbg_extend(self.h, len(ik), _ptr(ik), _ptr(ib), _ptr(ok)), "mbg_extend")
        return ok

    def seed_intv_batch(self, seqs, min_len=19, max_sub_occ=10):
        """Per read, an (n,4) uint64 array of (x0, x1, size, info); order within a read unspecified."""
        n = len(seqs)
        lens = np.array([len(s) for s in seqs], np.int32)
        keep = [np.ascontiguousarray(s, np.uint8) for s in seqs]
        ptrs = (C.c_void_p * max(n, 1))(*[s.ctypes.data if len(s) else None for s in keep])
        out = C.c_void_p()
        off = np.zeros(n + 1, np.int64)
        check(self.L.mbg_seed_batch(self.h, n, _ptr(lens), ptrs, min_len, max_sub_occ, C.byref(out), _ptr(off)),
              "mbg_seed_batch")
        tot = int(off[-1])
        res = []
        try:
            flat = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint64)), shape=(max(tot, 1) * 4,))
            flat = flat[: tot * 4].reshape(tot, 4).copy()
        finally:
            C.CDLL(None).free(out)
        for i in range(n):
            res.append(flat[off[i]:off[i + 1]])
        return res

    def seed_intv_flat(self, reads: np.ndarray, min_len=19, max_sub_occ=10):
        """seed_intv_batch over equal-length reads (a 2-D uint8 array) -> ((m,4) uint64 intervals, int64
        offsets[n+1]): read i's intervals are rows off[i]..off[i+1]."""
        reads = np.ascontiguousarray(reads, np.uint8)
        n, rl = reads.shape
        lens = np.full(n, rl, np.int32)
        addrs = np.uint64(reads.ctypes.data) + np.arange(n, dtype=np.uint64) * np.uint64(rl)   # uint8_t *const seq[n]
        out = C.c_void_p()
        off = np.zeros(n + 1, np.int64)
        check(self.L.mbg_seed_batch(self.h, n, _ptr(lens), _ptr(addrs), min_len, max_sub_occ, C.byref(out), _ptr(off)),
              "mbg_seed_batch")
        tot = int(off[-1])
        try:
            flat = np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_uint64)), shape=(max(tot, 1) * 4,))
            flat = flat[: tot * 4].reshape(tot, 4).copy()
        finally:
            C.CDLL(None).free(out)
        return flat, off

    def bwt_sa_batch(self, x):
        x = np.ascontiguousarray(x, np.uint64).copy()
        if len(x):
            check(self.L.mbg_sa_batch(self.h, len(x), _ptr(x)), "mbg_sa_batch")
        return x

    def l2b_getseq(self, tid, st, en, ctg_len=None, with_len=False):
        """Windows [st,en) of contigs tid, clipped as l2b_getseq clips them; with ctg_len (the
        index's contig lengths) each output is sized to its clipped window, which a negative en
        (clipped to the contig end, l2bit.c:61) needs.  with_len: also l2b_getseq's return values
        (en - st after clipping, which may be negative; -1 for a bad contig id)."""
        tid = np.ascontiguousarray(tid, np.int64)
        st = np.ascontiguousarray(st, np.int64)
        en = np.ascontiguousarray(en, np.int64)
        en_c = en
        if ctg_len is not None:
            cl = np.asarray(ctg_len, np.uint64)
            ok = (tid >= 0) & (tid < len(cl))
            L = np.where(ok, cl[np.where(ok, tid, 0)], 0).astype(np.uint64)
            en_c = np.where(en.astype(np.uint64) > L, L.astype(np.int64), en)
        span = np.maximum(en_c - np.maximum(st, 0), 0)
        off = np.zeros(len(tid), np.int64)
        if len(tid) > 1:
            np.cumsum(span[:-1], out=off[1:])
        out = np.zeros(int(span.sum()) + 1, np.uint8)
        ln = np.zeros(len(tid), np.int64)
        check(self.L.mbg_getseq(self.h, len(tid), _ptr(tid), _ptr(st), _ptr(en), _ptr(off), _ptr(out), len(out),
                                _ptr(ln)), "mbg_getseq")
        seqs = [out[off[i]:off[i] + max(ln[i], 0)].copy() if ln[i] >= 0 else None for i in range(len(tid))]
        return (seqs, ln) if with_len else seqs

    # ---- seed -> locate -> extend step over device-resident reads ------------------------------
    def map_step_dev(self, n_reads, off_ptr, len_ptr, seq_ptr, max_len, opt: MapOpt, collect_stats=False, stream=None):
        """One pass of the hot path; pointers are device addresses (e.g. torch tensor data_ptr()).
        Returns (MapStats, MapOut of device addresses valid until the next call)."""
        st = MapStats()
        out = MapOut()
        check(self.L.mbg_map_step_dev(self.h, stream, n_reads, off_ptr, len_ptr, seq_ptr, max_len, C.byref(opt),
                                      int(collect_stats), C.byref(st), C.byref(out)), "mbg_map_step_dev")
        return st, out

    def to_device(self, a: np.ndarray) -> "DeviceArray":
        """Copy a host array into a new device allocation owned by the returned DeviceArray."""
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        check(self.L.mbg_dev_alloc(self.h, max(a.nbytes, 1), C.byref(p)), "mbg_dev_alloc")
        d = DeviceArray(self, p.value, a.dtype, a.shape)
        if a.nbytes:
            check(self.L.mbg_copy_to_device(self.h, C.c_void_p(p.value), _ptr(a), a.nbytes), "mbg_copy_to_device")
        return d

    def write_device(self, dev_addr, a: np.ndarray):
        """Copy a host array into existing device memory at dev_addr."""
        a = np.ascontiguousarray(a)
        if a.nbytes:
            check(self.L.mbg_copy_to_device(self.h, C.c_void_p(dev_addr), _ptr(a), a.nbytes), "mbg_copy_to_device")

    def step_results(self, st: MapStats, out: MapOut) -> dict:
        """Host copies of one step's outputs: seed records, located rows (and their records), the
        extension jobs with their metadata, DP results and CIGAR slots."""
        n_jobs = int(st.n_jobs)
        r = {"recs": self.to_host(out.recs, SEED_REC_DTYPE, st.n_rec),
             "pos": self.to_host(out.pos, np.uint64, st.n_rows),
             "pos_rec": self.to_host(out.pos_rec, np.uint32, st.n_rows),
             "jobs": self.to_host(out.jobs, DP_JOB_DTYPE, n_jobs),
             "meta": self.to_host(out.meta, JOB_META_DTYPE, n_jobs),
             "res": self.to_host(out.res, DP_RES_DTYPE, n_jobs)}
        r["cigar"] = self.to_host(out.cigar, np.uint32, int(st.q_bytes) + int(st.t_bytes))
        return r

    def gather_bench(self, n_chains, steps, block_threads=256):
        """(ms, checksum) of the dependent random 64-byte gather over the resident occ table."""
        ms = C.c_float()
        cs = C.c_uint64()
        check(self.L.mbg_gather_bench(self.h, n_chains, steps, block_threads, C.byref(ms), C.byref(cs)), "mbg_gather_bench")
        return ms.value, cs.value

    def to_host(self, dev_addr, dtype, count):
        """Copy `count` elements of `dtype` from a device address returned by the library."""
        a = np.empty(int(count), dtype)
        if a.nbytes:
            check(self.L.mbg_copy_to_host(self.h, _ptr(a), C.c_void_p(dev_addr), a.nbytes), "mbg_copy_to_host")
        return a

    def use_graphs(self, on: bool):
        """Launch chains as captured HIP graphs (on) or launch by launch (off); results are the same."""
        check(self.L.mbg_use_graphs(self.h, 1 if on else 0), "mbg_use_graphs")

    def counters(self, reset=False) -> dict:
        c = Counters()
        check(self.L.mbg_counters(self.h, C.byref(c), 1 if reset else 0), "mbg_counters")
        return {f: getattr(c, f) for f, _ in Counters._fields_}

    # ---- DP ----------------------------------------------------------------------------------
    def _dp(self, fn, name, queries, targets, params, mat, m=5):
        n = len(queries)
        qbuf, qoff, qlen = pack_seqs(queries)
        tbuf, toff, tlen = pack_seqs(targets)
        jobs = np.zeros(n, DP_JOB_DTYPE)
        jobs["qoff"] = qoff[:-1]
        jobs["toff"] = toff[:-1]
        jobs["qlen"] = qlen
        jobs["tlen"] = tlen
        for f in ("q", "e", "q2", "e2", "w", "zdrop", "end_bonus", "flag"):
            jobs[f] = params[f]
        span = qlen.astype(np.int64) + tlen
        coff = np.zeros(n, np.int64)
        if n > 1:
            np.cumsum(span[:-1], out=coff[1:])
        jobs["cigar_off"] = coff
        res = np.zeros(n, DP_RES_DTYPE)
        cap = int(span.sum()) + 1
        cig = np.zeros(cap, np.uint32)
        mat = np.ascontiguousarray(mat, np.int8)
        qb = qbuf if len(qbuf) else np.zeros(1, np.uint8)
        tb = tbuf if len(tbuf) else np.zeros(1, np.uint8)
        check(fn(self.h, n, _ptr(jobs), _ptr(qb), len(qbuf), _ptr(tb), len(tbuf), m, _ptr(mat), _ptr(res), _ptr(cig),
                 cap), name)
        return DpBatchResult(res, cig, coff)

    def ksw_extd2(self, queries, targets, params, mat, m=5) -> DpBatchResult:
        """params: dict of per-job arrays/scalars q, e, q2, e2, w, zdrop, end_bonus, flag."""
        return self._dp(self.L.mbg_extd2_batch, "mbg_extd2_batch", queries, targets, params, mat, m)

    def align_pair(self, p, max_sw_mat, queries, targets, params, mat) -> DpBatchResult:
        """mb_align_pair (align.c:329-390) over a batch: p = (a, b, b_ts, q, e, q2, e2, meth) as the
        oracle's align_params; params: per-call arrays/scalars w, end_bonus, zdrop, flag."""
        n = len(queries)
        qbuf, qoff, qlen = pack_seqs(queries)
        tbuf, toff, tlen = pack_seqs(targets)
        jobs = np.zeros(n, ALIGN_JOB_DTYPE)
        jobs["qoff"] = qoff[:-1]
        jobs["toff"] = toff[:-1]
        jobs["qlen"] = qlen
        jobs["tlen"] = tlen
        for f in ("w", "end_bonus", "zdrop", "flag"):
            jobs[f] = params[f]
        span = qlen.astype(np.int64) + tlen + 1
        coff = np.zeros(n, np.int64)
        if n > 1:
            np.cumsum(span[:-1], out=coff[1:])
        jobs["cigar_off"] = coff
        res = np.zeros(n, DP_RES_DTYPE)
        cap = int(span.sum()) + 1
        cig = np.zeros(cap, np.uint32)
        mat = np.ascontiguousarray(mat, np.int8)
        o = AlignOpt(*[int(x) for x in p], int(max_sw_mat))
        qb = qbuf if len(qbuf) else np.zeros(1, np.uint8)
        tb = tbuf if len(tbuf) else np.zeros(1, np.uint8)
        check(self.L.mbg_align_pair_batch(self.h, C.byref(o), n, _ptr(jobs), _ptr(qb), len(qbuf), _ptr(tb), len(tbuf), _ptr(mat),
                                          _ptr(res), _ptr(cig), cap), "mbg_align_pair_batch")
        return DpBatchResult(res, cig, coff)

    def ksw_ll(self, queries, targets, sizes, mat, gapo, gape, xtra, m=5) -> np.ndarray:
        """ksw_ll_qinit + ksw_ll_{u8,i16}_core over a batch (ksw2_ll_sse.c:49-333); sizes is per job
        (1 = u8 core, 2 = i16 core) or one value.  Returns an LL_RES_DTYPE array."""
        n = len(queries)
        qbuf, qoff, qlen = pack_seqs(queries)
        tbuf, toff, tlen = pack_seqs(targets)
        jobs = np.zeros(n, LL_JOB_DTYPE)
        jobs["qoff"] = qoff[:-1]
        jobs["toff"] = toff[:-1]
        jobs["qlen"] = qlen
        jobs["tlen"] = tlen
        jobs["size"] = sizes
        res = np.zeros(n, LL_RES_DTYPE)
        mat = np.ascontiguousarray(mat, np.int8)
        qb = qbuf if len(qbuf) else np.zeros(1, np.uint8)
        tb = tbuf if len(tbuf) else np.zeros(1, np.uint8)
        check(self.L.mbg_ll_batch(self.h, n, _ptr(jobs), _ptr(qb), len(qbuf), _ptr(tb), len(tbuf), m, _ptr(mat), gapo,
                                  gape, xtra, _ptr(res)), "mbg_ll_batch")
        return res

    def update_extra(self, hits, cigars, qseqs, tseqs, mat, q, e, log_gap):
        """mb_fix_cigar + mb_update_extra's scores (align.c:70-258) for hits = list of dict(qs, qe, ts, te,
        rev) with their CIGARs and the query/target bases from each hit's first base.  Returns
        (EXTRA_JOB_DTYPE array, fixed CIGARs)."""
        n = len(hits)
        qbuf, qoff, _ = pack_seqs(qseqs)
        tbuf, toff, _ = pack_seqs(tseqs)
        jobs = np.zeros(n, EXTRA_JOB_DTYPE)
        for f in ("qs", "qe", "ts", "te", "rev"):
            jobs[f] = [h.get(f, 0) for h in hits]
        jobs["qoff"], jobs["toff"] = qoff[:-1], toff[:-1]
        jobs["n_cigar"] = [len(c) for c in cigars]
        coff = np.zeros(n + 1, np.int64)
        np.cumsum(jobs["n_cigar"], out=coff[1:])
        jobs["cigar_off"] = coff[:-1]
        cig = np.zeros(int(coff[-1]) + 1, np.uint32)
        for i, c in enumerate(cigars):
            cig[coff[i]:coff[i + 1]] = c
        mat = np.ascontiguousarray(mat, np.int8)
        qb = qbuf if len(qbuf) else np.zeros(1, np.uint8)
        tb = tbuf if len(tbuf) else np.zeros(1, np.uint8)
        check(self.L.mbg_update_extra_batch(self.h, n, _ptr(jobs), _ptr(qb), len(qbuf), _ptr(tb), len(tbuf), _ptr(mat), q, e,
                                            log_gap, _ptr(cig), len(cig)), "mbg_update_extra_batch")
        return jobs, [cig[coff[i]:coff[i] + jobs["n_cigar"][i]].tolist() for i in range(n)]

    def anchor(self, intervals, qlens, max_occ=250):
        """mb_anchor (seed.c:268-339, no methylation conversion) of each read's SMEM intervals ((n,4)
        uint64: x0, x1, size, info) -> (per read an ANCHOR_DTYPE array, per read its intervals sorted and
        deduplicated as the reference leaves them)."""
        n = len(intervals)
        us = [np.asarray(u, np.uint64).reshape(-1, 4) for u in intervals]
        uoff = np.zeros(n + 1, np.int64)
        np.cumsum([len(u) for u in us], out=uoff[1:])
        u = np.ascontiguousarray(np.concatenate(us) if n else np.zeros((0, 4), np.uint64), np.uint64)
        aoff = np.zeros(n + 1, np.int64)
        np.cumsum(anchor_slots(us, max_occ), out=aoff[1:])
        anc = np.zeros(max(int(aoff[-1]), 1), ANCHOR_DTYPE)
        nanc = np.zeros(max(n, 1), np.int64)
        nu = np.zeros(max(n, 1), np.int64)
        ql = np.ascontiguousarray(qlens, np.int32)
        check(self.L.mbg_anchor_batch(self.h, n, _ptr(u), _ptr(uoff), _ptr(ql), max_occ, _ptr(anc), _ptr(aoff), _ptr(nanc),
                                      _ptr(nu)), "mbg_anchor_batch")
        return ([anc[aoff[i]:aoff[i] + nanc[i]].copy() for i in range(n)],
                [u[uoff[i]:uoff[i] + nu[i]].copy() for i in range(n)])

    def lchain(self, anchors, opt: ChainOpt):
        """mb_lchain_dp (lchain.c:145-240) of each read's anchors (ANCHOR_DTYPE arrays in mb_anchor's order)
        -> per read (uint64 chains score<<32 | count, chained anchors ANCHOR_DTYPE)."""
        n = len(anchors)
        cnt = np.array([len(a) for a in anchors], np.int64)
        off = np.zeros(n + 1, np.int64)
        np.cumsum(cnt, out=off[1:])
        anc = np.ascontiguousarray(np.concatenate(anchors) if n else np.zeros(0, ANCHOR_DTYPE), ANCHOR_DTYPE)
        if not len(anc):
            anc = np.zeros(1, ANCHOR_DTYPE)
        u = np.zeros(max(int(off[-1]), 1), np.uint64)
        out = np.zeros(max(int(off[-1]), 1), ANCHOR_DTYPE)
        n_u = np.zeros(max(n, 1), np.int32)
        n_out = np.zeros(max(n, 1), np.int64)
        cn = cnt if n else np.zeros(1, np.int64)
        check(self.L.mbg_lchain_batch(self.h, n, _ptr(anc), _ptr(off), _ptr(cn), C.byref(opt), _ptr(u), _ptr(n_u), _ptr(out),
                                      _ptr(n_out)), "mbg_lchain_batch")
        return [(u[off[i]:off[i] + n_u[i]].copy(), out[off[i]:off[i] + n_out[i]].copy()) for i in range(n)]

    def matesw(self, reads, tid, ts, te, min_sc, opt: MateswOpt, stats: "MateswStats | None" = None):
        """Mate rescue of candidate windows [ts, te) on contigs tid for reads (one per candidate, as they
        would align) -> (MATESW_RES_DTYPE array, list of extension CIGARs)."""
        n = len(reads)
        qbuf, qoff, qlen = pack_seqs(reads)
        jobs = np.zeros(n, MATESW_JOB_DTYPE)
        jobs["qoff"], jobs["qlen"], jobs["tid"], jobs["ts"], jobs["te"], jobs["min_sc"] = qoff[:-1], qlen, tid, ts, te, min_sc
        span = qlen.astype(np.int64) + np.maximum(np.asarray(te, np.int64) - np.maximum(np.asarray(ts, np.int64), 0), 0)
        coff = np.zeros(n, np.int64)
        if n > 1:
            np.cumsum(span[:-1], out=coff[1:])
        jobs["cigar_off"] = coff
        cap = int(span.sum()) + 1
        cig = np.zeros(cap, np.uint32)
        res = np.zeros(n, MATESW_RES_DTYPE)
        qb = qbuf if len(qbuf) else np.zeros(1, np.uint8)
        check(self.L.mbg_matesw_batch(self.h, n, _ptr(jobs), _ptr(qb), len(qbuf), C.byref(opt), _ptr(res), _ptr(cig), cap,
                                      C.byref(stats) if stats is not None else None), "mbg_matesw_batch")
        nb = np.where(res["rescued"] != 0, res["n_cigar"], np.where(res["ran_dp"] != 0, np.maximum(res["ez"]["n_cigar"], 0), 0))
        cigars = [cig[coff[i]:coff[i] + nb[i]].tolist() for i in range(n)]
        return res, cigars

    def ksw_extz2(self, queries, targets, params, mat, m=5) -> DpBatchResult:
        p = dict(params)
        p.setdefault("q2", 0)
        p.setdefault("e2", 0)
        return self._dp(self.L.mbg_extz2_batch, "mbg_extz2_batch", queries, targets, p, mat, m)
