"""ctypes bindings to the CPU checkers (TEST INFRASTRUCTURE ONLY).

Two libraries, both built by ``oracle/Makefile``:

* ``oracle/liboracle.so`` -- the C restatement in ``oracle/mb_oracle.c`` (the oracle);
* ``oracle/_ref/libmbref.so`` -- the reference's own sources compiled where they lie (present
  wherever ``oracle/_ref`` was built or copied; absent otherwise).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBORACLE = os.path.join(HERE, "liboracle.so")
LIBREF = os.path.join(HERE, "_ref", "libmbref.so")
LIBREF_PE = os.path.join(HERE, "_ref", "libmbref_pe.so")
REF_CLI = os.path.join(HERE, "_ref", "minibwa")


class OrcSai(C.Structure):
    _fields_ = [("x0", C.c_uint64), ("x1", C.c_uint64), ("size", C.c_uint64), ("info", C.c_uint64)]


class OrcBwt(C.Structure):
    _fields_ = [("primary", C.c_uint64), ("L2", C.c_uint64 * 5), ("seq_len", C.c_uint64),
                ("data_len", C.c_uint64), ("data", C.POINTER(C.c_uint64)), ("sa_bit", C.c_uint32),
                ("n_sa", C.c_uint64), ("sa", C.POINTER(C.c_uint64)), ("pre_len", C.c_int32),
                ("pre", C.POINTER(OrcSai))]


class OrcEz(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte",
                                         "mte_q", "score", "reach_end", "n_cigar")]


class LlRst(C.Structure):
    """ksw_llrst_t (ksw2.h:47-51)."""
    _fields_ = [(n, C.c_int32) for n in ("score", "te", "qe", "score2", "te2")]


LL_FIELDS = ("score", "te", "qe", "score2", "te2")


class OrcRescue(C.Structure):
    _fields_ = [("ran_ll", C.c_int32), ("ran_dp", C.c_int32), ("rescued", C.c_int32), ("ll", LlRst), ("ez", OrcEz),
                ("qs", C.c_int32), ("qe", C.c_int32), ("ts", C.c_int32), ("te", C.c_int32), ("dp_max2", C.c_int32)]


class Anchor(C.Structure):
    """mb_anchor_t (mbpriv.h:26-32)."""
    _fields_ = [("sid", C.c_int32), ("len", C.c_int32), ("qpos", C.c_int32), ("flag", C.c_uint32), ("tpos", C.c_int64)]


ANCHOR_DTYPE = np.dtype([("sid", "<i4"), ("len", "<i4"), ("qpos", "<i4"), ("flag", "<u4"), ("tpos", "<i8")])


class OrcExtra(C.Structure):
    _fields_ = [("qs", C.c_int32), ("qe", C.c_int32), ("rev", C.c_int32), ("n_cigar", C.c_int32), ("ts", C.c_int64),
                ("te", C.c_int64), ("blen", C.c_int32), ("mlen", C.c_int32), ("n_ambi", C.c_int32), ("dp_max", C.c_int32)]


EXTRA_FIELDS = ("qs", "qe", "ts", "te", "n_cigar", "blen", "mlen", "n_ambi", "dp_max")


class MbxHit(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("rescued", "qs", "qe", "score", "score0", "subsc", "dp_max", "dp_max2",
                                         "dp_score", "dp_max0", "mlen", "blen", "n_ambi", "n_cigar")] + \
               [("ts", C.c_int64), ("te", C.c_int64)]


HIT_FIELDS = ("rescued", "qs", "qe", "score", "score0", "subsc", "dp_max", "dp_max2", "dp_score", "dp_max0", "mlen",
              "blen", "n_ambi", "n_cigar", "ts", "te")
EZ11 = ("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score", "reach_end", "n_cigar")


def rescue_params(a=2, b=8, b_ts=0, b_ambi=1, q=12, e=2, bw=100, zdrop=100, end_bonus=5, min_dp_max=30, min_len=19):
    """The mb_opt_t fields mate rescue reads (pe.c:249-312), as the int32[11] the rescue functions take."""
    return np.array([a, b, b_ts, b_ambi, q, e, bw, zdrop, end_bonus, min_dp_max, min_len], np.int32)
KSW_LL_STOP = 0x20000
KSW_LL_SUBO = 0x40000

EZ_FIELDS = ("max", "zdropped", "max_q", "max_t", "mqe", "mqe_t", "mte", "mte_q", "score", "reach_end")

_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C")
_i8p = np.ctypeslib.ndpointer(dtype=np.int8, flags="C")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C")


def _load_oracle():
    if not os.path.exists(LIBORACLE):
        raise RuntimeError(f"{LIBORACLE} is not built (make -C oracle)")
    L = C.CDLL(LIBORACLE)
    L.orc_bwt_load.restype = C.POINTER(OrcBwt)
    L.orc_bwt_load.argtypes = [C.c_char_p]
    L.orc_bwt_destroy.argtypes = [C.POINTER(OrcBwt)]
    L.orc_bwt_cache.argtypes = [C.POINTER(OrcBwt), C.c_int32]
    L.orc_rank2a.argtypes = [C.POINTER(OrcBwt), C.c_uint64, C.c_uint64, _u64p, _u64p]
    L.orc_extend.argtypes = [C.POINTER(OrcBwt), C.POINTER(OrcSai), C.POINTER(OrcSai), C.c_int]
    L.orc_seed_intv.restype = C.c_int64
    L.orc_seed_intv.argtypes = [C.POINTER(OrcBwt), C.c_int32, _u8p, C.c_int32, C.c_int32, C.POINTER(OrcSai), C.c_int64]
    L.orc_smem.restype = C.c_int64
    L.orc_smem.argtypes = [C.POINTER(OrcBwt), C.c_uint32, _u8p, C.c_int64, C.c_int64, C.c_int64, C.POINTER(OrcSai)]
    L.orc_sa.restype = C.c_uint64
    L.orc_sa.argtypes = [C.POINTER(OrcBwt), C.c_uint64]
    L.orc_sa_batch.argtypes = [C.POINTER(OrcBwt), C.c_int64, _u64p]
    dp_common = [C.c_int, _u8p, C.c_int, _u8p, C.c_int8, _i8p]
    L.orc_extd2.restype = C.c_int
    L.orc_extd2.argtypes = dp_common + [C.c_int8] * 4 + [C.c_int] * 4 + [C.POINTER(OrcEz), _u32p, C.c_int]
    L.orc_extz2.restype = C.c_int
    L.orc_extz2.argtypes = dp_common + [C.c_int8] * 2 + [C.c_int] * 4 + [C.POINTER(OrcEz), _u32p, C.c_int]
    L.orc_bench.restype = C.c_uint64
    L.orc_bench.argtypes = [C.POINTER(OrcBwt), C.c_int, C.c_int64, C.c_int]
    L.orc_dp_cells.restype = C.c_int64
    L.orc_dp_cells.argtypes = [C.c_int, C.c_int, C.c_int]
    L.orc_align_pair.restype = C.c_int
    L.orc_align_pair.argtypes = [_i32p, C.c_int64, C.c_int, _u8p, C.c_int, _u8p, _i8p, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(OrcEz), _u32p, C.c_int]
    L.orc_ungap.restype = C.c_int32
    L.orc_ungap.argtypes = [C.c_int32, _u8p, C.c_int32, _u8p, C.c_int32, C.c_int32] + [C.POINTER(C.c_int32)] * 3
    L.orc_matesw_align.argtypes = [_i32p, C.c_int32, _u8p, C.c_int32, _u8p, C.c_int32, C.c_int32, C.POINTER(OrcRescue),
                                   _u32p, C.c_int]
    L.orc_matesw_cand.argtypes = [_i32p, C.c_int32, _u8p, C.c_int32, _u8p, C.c_int32, C.c_int32, _i32p, _i32p,
                                  C.POINTER(C.c_int32), C.POINTER(OrcRescue), _u32p, C.c_int]
    L.orc_anchor.restype = C.c_int64
    L.orc_lchain.restype = C.c_int32
    L.orc_lchain.argtypes = [C.c_int64, _u64p, _u64p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                             C.c_float, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
    L.orc_anchor.argtypes = [C.POINTER(OrcBwt), C.c_int64, _u64p, _u64p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                             C.c_void_p, C.c_int64]
    L.orc_update_extra.argtypes = [C.POINTER(OrcExtra), _u32p, _u8p, _u8p, _i8p, C.c_int8, C.c_int8, C.c_int]
    L.orc_ll.restype = C.c_int
    L.orc_ll.argtypes = [C.c_int, C.c_int, _u8p, C.c_int, _i8p, C.c_int, _u8p, C.c_int, C.c_int, C.c_int,
                         C.POINTER(LlRst)]
    return L


_ORC = None


def orc():
    global _ORC
    if _ORC is None:
        _ORC = _load_oracle()
    return _ORC


class OracleIndex:
    """An .mbw loaded by the restatement (bwt.c:648-683) plus its 10-mer cache (map-algo.c:36)."""

    def __init__(self, mbw_path: str, pre_len: int = 10):
        self.L = orc()
        self.b = self.L.orc_bwt_load(mbw_path.encode())
        if not self.b:
            raise IOError(f"cannot load {mbw_path}")
        if pre_len > 0:
            self.L.orc_bwt_cache(self.b, pre_len)
        bb = self.b.contents
        self.primary = bb.primary
        self.seq_len = bb.seq_len
        self.L2 = list(bb.L2)

    def close(self):
        if self.b:
            self.L.orc_bwt_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rank2a(self, k, l):
        ck = np.zeros(4, np.uint64)
        cl = np.zeros(4, np.uint64)
        self.L.orc_rank2a(self.b, int(k), int(l), ck, cl)
        return ck, cl

    def extend(self, ik, is_back):
        i = OrcSai(*[int(v) for v in ik])
        ok = (OrcSai * 4)()
        self.L.orc_extend(self.b, C.byref(i), ok, int(is_back))
        return [(o.x0, o.x1, o.size) for o in ok]

    def seed(self, seq: np.ndarray, min_len=19, max_sub_occ=10):
        """seed.c:24-54 -> array (n,4) uint64 of (x0, x1, size, info)."""
        seq = np.ascontiguousarray(seq, np.uint8)
        cap = 4 * len(seq) + 64
        buf = (OrcSai * cap)()
        n = self.L.orc_seed_intv(self.b, len(seq), seq, min_len, max_sub_occ, buf, cap)
        assert n <= cap
        out = np.frombuffer(buf, dtype=np.uint64).reshape(cap, 4)[:n].copy()
        return out

    def smem_step(self, seq: np.ndarray, x: int, min_len=19, min_occ=1, length=None):
        """One mb_bwt_smem call (bwt.c:280-305) from position x -> (next x, (x0, x1, size, info) or None)."""
        p = OrcSai()
        nx = self.L.orc_smem(self.b, len(seq) if length is None else length, seq, int(x), int(min_len), int(min_occ), C.byref(p))
        return int(nx), ((p.x0, p.x1, p.size, p.info) if p.size > 0 else None)

    def sa(self, rows):
        x = np.ascontiguousarray(rows, np.uint64).copy()
        self.L.orc_sa_batch(self.b, len(x), x)
        return x

    def bench_checksum(self, kind: str, n=1000000, intv=20) -> str:
        """`minibwa bench -b 2a|sa|msa` checksum (main.c:164-200) as lower-case hex."""
        t = {"2a": 0, "sa": 1, "msa": 2}[kind]
        return "%x" % self.L.orc_bench(self.b, t, n, intv)


def _dp_result(ez, cig, ret):
    d = {f: int(getattr(ez, f)) for f in EZ_FIELDS}
    d["cigar"] = [int(c) for c in cig[: ez.n_cigar]]
    d["cells"] = int(ret)
    return d


def extd2(query, target, mat, q, e, q2, e2, w, zdrop, end_bonus, flag, m=5):
    query = np.ascontiguousarray(query, np.uint8)
    target = np.ascontiguousarray(target, np.uint8)
    mat = np.ascontiguousarray(mat, np.int8)
    ez = OrcEz()
    cap = len(query) + len(target) + 8
    cig = np.zeros(cap, np.uint32)
    ret = orc().orc_extd2(len(query), query, len(target), target, m, mat, q, e, q2, e2, w, zdrop, end_bonus,
                          flag, C.byref(ez), cig, cap)
    return _dp_result(ez, cig, ret)


def extz2(query, target, mat, q, e, w, zdrop, end_bonus, flag, m=5):
    query = np.ascontiguousarray(query, np.uint8)
    target = np.ascontiguousarray(target, np.uint8)
    mat = np.ascontiguousarray(mat, np.int8)
    ez = OrcEz()
    cap = len(query) + len(target) + 8
    cig = np.zeros(cap, np.uint32)
    ret = orc().orc_extz2(len(query), query, len(target), target, m, mat, q, e, w, zdrop, end_bonus, flag,
                          C.byref(ez), cig, cap)
    return _dp_result(ez, cig, ret)


def align_params(a=2, b=8, b_ts=0, q=4, e=2, q2=24, e2=1, meth=0):
    """The mb_opt_t fields mb_align_pair reads (align.c:329-378), in the order orc_align_pair and
    mbx_align_pair take them."""
    return np.array([a, b, b_ts, q, e, q2, e2, meth], np.int32)


def orc_align_pair(p, max_sw_mat, query, target, mat, w, end_bonus, zdrop, flag):
    """mb_align_pair (align.c:329-390) restated: the fast paths, band cap, skip and DP dispatch."""
    query = np.ascontiguousarray(query, np.uint8)
    target = np.ascontiguousarray(target, np.uint8)
    mat = np.ascontiguousarray(mat, np.int8)
    p = np.ascontiguousarray(p, np.int32)
    ez = OrcEz()
    cap = len(query) + len(target) + 8
    cig = np.zeros(cap, np.uint32)
    ret = orc().orc_align_pair(p, max_sw_mat, len(query), query, len(target), target, mat, w, end_bonus, zdrop, flag, C.byref(ez),
                               cig, cap)
    return _dp_result(ez, cig, ret)


def gen_nt4_mat(a=2, b=8, b_ts=0, b_ambi=1, mt=0):
    """ksw2.h:199-219."""
    m = 5
    a = abs(a)
    b = -abs(b)
    b_ambi = -abs(b_ambi)
    mat = np.zeros(25, np.int8)
    for i in range(m - 1):
        for j in range(m - 1):
            mat[i * m + j] = a if i == j else b
        mat[i * m + m - 1] = b_ambi
    for j in range(m):
        mat[(m - 1) * m + j] = b_ambi
    if b_ts != 0 and b_ts != b:
        bt = -abs(b_ts)
        mat[0 * m + 2] = mat[1 * m + 3] = mat[2 * m + 0] = mat[3 * m + 1] = bt
    if mt == 1:
        mat[1 * m + 3] = a
    elif mt == 2:
        mat[2 * m + 0] = a
    return mat


# ---------------------------------------------------------------------------------------------
# The compiled reference (oracle/_ref/libmbref.so)
# ---------------------------------------------------------------------------------------------

class RefSaiV(C.Structure):
    _fields_ = [("n", C.c_size_t), ("m", C.c_size_t), ("a", C.POINTER(OrcSai))]


class RefExtz(C.Structure):
    _fields_ = [("maxz", C.c_uint32), ("max_q", C.c_int), ("max_t", C.c_int), ("mqe", C.c_int),
                ("mqe_t", C.c_int), ("mte", C.c_int), ("mte_q", C.c_int), ("score", C.c_int),
                ("m_cigar", C.c_int), ("n_cigar", C.c_int), ("reach_end", C.c_int),
                ("cigar", C.POINTER(C.c_uint32))]


def have_ref() -> bool:
    return os.path.exists(LIBREF)


_REF = None
_LIBC = None


def ref():
    global _REF, _LIBC
    if _REF is None:
        if not have_ref():
            raise RuntimeError(f"{LIBREF} is not built (make -C oracle ref)")
        L = C.CDLL(LIBREF)
        L.mb_bwt_load.restype = C.c_void_p
        L.mb_bwt_load.argtypes = [C.c_char_p]
        L.mb_bwt_destroy.argtypes = [C.c_void_p]
        L.mb_bwt_cache.argtypes = [C.c_void_p, C.c_int32]
        L.mb_bwt_rank2a.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, _u64p, _u64p]
        L.mb_bwt_extend.argtypes = [C.c_void_p, C.POINTER(OrcSai), C.POINTER(OrcSai), C.c_int]
        L.mb_bwt_sa.restype = C.c_uint64
        L.mb_bwt_sa.argtypes = [C.c_void_p, C.c_uint64]
        L.mb_bwt_sa_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, _u64p]
        L.mb_seed_intv_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_int32),
                                         C.POINTER(C.POINTER(C.c_uint8)), C.c_int32, C.c_int32, C.POINTER(RefSaiV)]
        L.mb_seed_intv.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, _u8p, C.c_int32, C.c_int32, C.POINTER(RefSaiV)]
        dp_common = [C.c_void_p, C.c_int, _u8p, C.c_int, _u8p, C.c_int8, _i8p]
        L.ksw_extd2_sse.argtypes = dp_common + [C.c_int8] * 4 + [C.c_int] * 4 + [C.POINTER(RefExtz)]
        L.ksw_extz2_sse.argtypes = dp_common + [C.c_int8] * 2 + [C.c_int] * 4 + [C.POINTER(RefExtz)]
        L.ksw_ll_qinit.restype = C.c_void_p
        L.ksw_ll_qinit.argtypes = [C.c_void_p, C.c_int, C.c_int, _u8p, C.c_int, _i8p]
        for fn in (L.ksw_ll_u8_core, L.ksw_ll_i16_core):
            fn.restype = LlRst
            fn.argtypes = [C.c_void_p, C.c_int, _u8p, C.c_int, C.c_int, C.c_int]
        _LIBC = C.CDLL(None)
        _LIBC.free.argtypes = [C.c_void_p]
        _REF = L
    return _REF


class RefIndex:
    """An .mbw loaded by the reference's own mb_bwt_load + mb_bwt_cache."""

    def __init__(self, mbw_path: str, pre_len: int = 10):
        self.L = ref()
        self.b = self.L.mb_bwt_load(mbw_path.encode())
        if not self.b:
            raise IOError(f"cannot load {mbw_path}")
        if pre_len > 0:
            self.L.mb_bwt_cache(self.b, pre_len)

    def close(self):
        if self.b:
            self.L.mb_bwt_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def rank2a(self, k, l):
        ck = np.zeros(4, np.uint64)
        cl = np.zeros(4, np.uint64)
        self.L.mb_bwt_rank2a(self.b, int(k), int(l), ck, cl)
        return ck, cl

    def extend(self, ik, is_back):
        i = OrcSai(*[int(v) for v in ik])
        ok = (OrcSai * 4)()
        self.L.mb_bwt_extend(self.b, C.byref(i), ok, int(is_back))
        return [(o.x0, o.x1, o.size) for o in ok]

    def sa(self, rows):
        x = np.ascontiguousarray(rows, np.uint64).copy()
        if len(x):
            self.L.mb_bwt_sa_batch(None, self.b, len(x), x)
        return x

    def seed_batch(self, seqs, min_len=19, max_sub_occ=10):
        """seed.c:56-103 over a list of uint8 arrays -> list of (n,4) uint64 arrays."""
        n = len(seqs)
        lens = (C.c_int32 * n)(*[len(s) for s in seqs])
        keep = [np.ascontiguousarray(s, np.uint8) for s in seqs]
        ptrs = (C.POINTER(C.c_uint8) * n)(*[s.ctypes.data_as(C.POINTER(C.c_uint8)) for s in keep])
        v = (RefSaiV * n)()
        self.L.mb_seed_intv_batch(None, self.b, n, lens, ptrs, min_len, max_sub_occ, v)
        out = []
        for i in range(n):
            k = v[i].n
            if k:
                a = np.ctypeslib.as_array(C.cast(v[i].a, C.POINTER(C.c_uint64)), shape=(k * 4,))
                out.append(a.reshape(k, 4).copy())
            else:
                out.append(np.zeros((0, 4), np.uint64))
            if v[i].a:
                _LIBC.free(C.cast(v[i].a, C.c_void_p))
        return out


def _ref_dp(fn, args):
    ez = RefExtz()
    fn(None, *args, C.byref(ez))
    d = {"max": ez.maxz & 0x7fffffff, "zdropped": ez.maxz >> 31}
    for f in EZ_FIELDS[2:]:
        d[f] = int(getattr(ez, f))
    d["cigar"] = [int(ez.cigar[i]) for i in range(ez.n_cigar)] if ez.n_cigar > 0 else []
    if ez.cigar:
        _LIBC.free(C.cast(ez.cigar, C.c_void_p))
    return d


def ref_extd2(query, target, mat, q, e, q2, e2, w, zdrop, end_bonus, flag, m=5):
    L = ref()
    query = np.ascontiguousarray(query, np.uint8)
    target = np.ascontiguousarray(target, np.uint8)
    mat = np.ascontiguousarray(mat, np.int8)
    return _ref_dp(L.ksw_extd2_sse, (len(query), query, len(target), target, m, mat, q, e, q2, e2, w, zdrop,
                                     end_bonus, flag))


def ref_extz2(query, target, mat, q, e, w, zdrop, end_bonus, flag, m=5):
    L = ref()
    query = np.ascontiguousarray(query, np.uint8)
    target = np.ascontiguousarray(target, np.uint8)
    mat = np.ascontiguousarray(mat, np.int8)
    return _ref_dp(L.ksw_extz2_sse, (len(query), query, len(target), target, m, mat, q, e, w, zdrop, end_bonus,
                                     flag))


def _ll_args(query, target, mat):
    return (np.ascontiguousarray(query, np.uint8), np.ascontiguousarray(target, np.uint8),
            np.ascontiguousarray(mat, np.int8))


def orc_ll(query, target, mat, size, gapo, gape, xtra, m=5):
    """The C restatement of ksw_ll_qinit + ksw_ll_{u8,i16}_core -> dict of ksw_llrst_t fields."""
    query, target, mat = _ll_args(query, target, mat)
    r = LlRst()
    if orc().orc_ll(size, len(query), query, m, mat, len(target), target, gapo, gape, xtra, C.byref(r)) != 0:
        raise ValueError("orc_ll: bad arguments")
    return {f: int(getattr(r, f)) for f in LL_FIELDS}


def ref_ll(query, target, mat, size, gapo, gape, xtra, m=5):
    """The compiled reference: ksw_ll_qinit (ksw2_ll_sse.c:49-96) then the u8 or i16 core."""
    L = ref()
    query, target, mat = _ll_args(query, target, mat)
    q = L.ksw_ll_qinit(None, size, len(query), query, m, mat)
    core = L.ksw_ll_u8_core if size <= 1 else L.ksw_ll_i16_core
    r = core(q, len(target), target, gapo, gape, xtra)
    _LIBC.free(C.c_void_p(q))
    return {f: int(getattr(r, f)) for f in LL_FIELDS}


def _rescue_dict(r, cig):
    ez = {f: int(getattr(r.ez, f)) for f in EZ11}
    d = {"ran_ll": r.ran_ll, "ran_dp": r.ran_dp, "rescued": r.rescued, "ll": {f: int(getattr(r.ll, f)) for f in LL_FIELDS},
         "ez": ez, "qs": r.qs, "qe": r.qe, "ts": r.ts, "te": r.te, "dp_max2": r.dp_max2,
         "cigar": [int(c) for c in cig[:max(0, ez["n_cigar"])]] if r.ran_dp else []}
    return d


def orc_ungap(query, target, mt=0, kmer=7):
    """mb_ungap restated -> (max, max_i, n_good, n_kmer)."""
    q = np.ascontiguousarray(query, np.uint8)
    t = np.ascontiguousarray(target, np.uint8)
    a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
    m = orc().orc_ungap(len(q), q, len(t), t, kmer, mt, C.byref(a), C.byref(b), C.byref(c))
    return int(m), a.value, b.value, c.value


def orc_matesw_cand(p, query, window, min_sc, mt=0):
    """One mate-rescue candidate restated (pe.c:342-349 + mb_matesw_align before mb_update_extra)."""
    q = np.ascontiguousarray(query, np.uint8)
    w = np.ascontiguousarray(window, np.uint8)
    ug = np.zeros(4, np.int32)
    w2 = np.zeros(2, np.int32)
    al = C.c_int32()
    r = OrcRescue()
    cig = np.zeros(len(q) + len(w) + 8, np.uint32)
    orc().orc_matesw_cand(np.ascontiguousarray(p, np.int32), len(q), q, len(w), w, min_sc, mt, ug, w2, C.byref(al),
                          C.byref(r), cig, len(cig))
    d = _rescue_dict(r, cig)
    d.update(ug=tuple(int(x) for x in ug), win2=(int(w2[0]), int(w2[1])), aligned=al.value)
    return d


_REFPE = None


def have_ref_pe():
    return os.path.exists(LIBREF_PE)


def ref_pe():
    global _REFPE
    if _REFPE is None:
        if not have_ref_pe():
            raise RuntimeError(f"{LIBREF_PE} is not built (make -C oracle ref)")
        L = C.CDLL(LIBREF_PE)
        L.mbx_ungap.restype = C.c_int32
        L.mbx_ungap.argtypes = [C.c_int32, _u8p, C.c_int32, _u8p, C.c_int32, C.c_int32, _i32p]
        L.mbx_matesw_align.argtypes = [_i32p, C.c_int32, _u8p, C.c_int32, _u8p, C.c_int32, C.c_int32, C.POINTER(MbxHit),
                                       _u32p, C.c_int32, _i32p, _u32p, C.c_int32]
        L.mb_idx_load.restype = C.c_void_p
        L.mb_idx_load.argtypes = [C.c_char_p, C.c_int]
        L.mb_idx_destroy.argtypes = [C.c_void_p]
        L.mbx_anchor.restype = C.c_int64
        L.mbx_lchain.restype = C.c_int32
        L.mbx_lchain.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                 C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int64)]
        L.mbx_anchor.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p, C.c_int64]
        L.mbx_update_extra.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_int32, C.c_int32, _u32p, _u8p, _u8p,
                                       _i8p, C.c_int8, C.c_int8, C.c_int32, C.POINTER(MbxHit), _u32p, C.c_int32]
        L.mbx_align_pair.argtypes = [_i32p, C.c_int64, C.c_int32, _u8p, C.c_int32, _u8p, _i8p, C.c_int32, C.c_int32, C.c_int32,
                                     C.c_int32, _i32p, _u32p, C.c_int32]
        L.mbx_finish.argtypes = [_i32p, C.c_int32, _u8p, _u8p, C.c_int32, C.c_int32, C.c_int64, C.c_int64, C.c_int32,
                                 C.c_int32, C.c_int32, C.c_int32, C.c_int32, _u32p, C.POINTER(MbxHit), _u32p, C.c_int32]
        _REFPE = L
    return _REFPE


def ref_ungap(query, target, mt=0, kmer=7):
    """The reference's mb_ungap (pe.c:209-247) -> (max, max_i, n_good, n_kmer)."""
    q = np.ascontiguousarray(query, np.uint8)
    t = np.ascontiguousarray(target, np.uint8)
    out = np.zeros(3, np.int32)
    m = ref_pe().mbx_ungap(len(q), q, len(t), t, kmer, mt, out)
    return int(m), int(out[0]), int(out[1]), int(out[2])


def ref_align_pair(p, max_sw_mat, query, target, mat, w, end_bonus, zdrop, flag):
    """The reference's own static mb_align_pair (align.c:329-390), through ref_align_shim.c."""
    L = ref_pe()
    query = np.ascontiguousarray(query, np.uint8)
    target = np.ascontiguousarray(target, np.uint8)
    mat = np.ascontiguousarray(mat, np.int8)
    p = np.ascontiguousarray(p, np.int32)
    out = np.zeros(11, np.int32)
    cap = len(query) + len(target) + 8
    cig = np.zeros(cap, np.uint32)
    L.mbx_align_pair(p, max_sw_mat, len(query), query, len(target), target, mat, w, end_bonus, zdrop, flag, out, cig, cap)
    d = {f: int(out[i]) for i, f in enumerate(EZ_FIELDS)}
    d["cigar"] = [int(c) for c in cig[: out[10]]]
    return d


def _hit_dict(h, cig):
    d = {f: int(getattr(h, f)) for f in HIT_FIELDS}
    d["cigar"] = [int(c) for c in cig[:d["n_cigar"]]]
    return d


def ref_matesw_align(p, query, window, min_sc, mt=0):
    """The reference's mb_matesw_align (pe.c:249-312) -> (hit after mb_update_extra, ez fields, ez CIGAR)."""
    q = np.ascontiguousarray(query, np.uint8).copy()
    w = np.ascontiguousarray(window, np.uint8).copy()
    h = MbxHit()
    cap = len(q) + len(w) + 8
    cig, ezc = np.zeros(cap, np.uint32), np.zeros(cap, np.uint32)
    ez = np.zeros(11, np.int32)
    ref_pe().mbx_matesw_align(np.ascontiguousarray(p, np.int32), len(q), q, len(w), w, min_sc, mt, C.byref(h), cig, cap,
                              ez, ezc, cap)
    ezd = {f: int(ez[i]) for i, f in enumerate(EZ11)}
    return _hit_dict(h, cig), ezd, [int(c) for c in ezc[:max(0, ezd["n_cigar"])]]


def ref_finish(p, query, window, mt, qs, qe, ts, te, ez_max, score, score2, dp_max2, cigar):
    """A rescued hit built from pre-update fields, then the reference's mb_append_cigar + mb_update_extra."""
    q = np.ascontiguousarray(query, np.uint8)
    w = np.ascontiguousarray(window, np.uint8)
    c0 = np.ascontiguousarray(cigar, np.uint32) if len(cigar) else np.zeros(1, np.uint32)
    h = MbxHit()
    cap = len(q) + len(w) + 8
    cig = np.zeros(cap, np.uint32)
    ref_pe().mbx_finish(np.ascontiguousarray(p, np.int32), mt, q, w, qs, qe, ts, te, ez_max, score, score2, dp_max2,
                        len(cigar), c0, C.byref(h), cig, cap)
    return _hit_dict(h, cig)


def orc_update_extra(hit, cigar, qseq, tseq, mat, q, e, log_gap):
    """mb_fix_cigar + mb_update_extra restated; hit = dict(qs, qe, ts, te, rev); qseq/tseq start at the
    hit's first bases.  Returns (dict of EXTRA_FIELDS, fixed CIGAR)."""
    r = OrcExtra(hit["qs"], hit["qe"], hit.get("rev", 0), len(cigar), hit["ts"], hit["te"], 0, 0, 0, 0)
    c = np.array(list(cigar) + [0], np.uint32)
    orc().orc_update_extra(C.byref(r), c, np.ascontiguousarray(qseq, np.uint8), np.ascontiguousarray(tseq, np.uint8),
                           np.ascontiguousarray(mat, np.int8), q, e, log_gap)
    return {f: int(getattr(r, f)) for f in EXTRA_FIELDS}, [int(x) for x in c[:r.n_cigar]]


def ref_update_extra(hit, cigar, qseq, tseq, mat, q, e, log_gap):
    """The reference's mb_update_extra (with mb_fix_cigar) -> (dict of EXTRA_FIELDS, fixed CIGAR)."""
    h = MbxHit()
    c0 = np.array(list(cigar) + [0], np.uint32)
    cap = len(cigar) + 8
    out = np.zeros(cap, np.uint32)
    ref_pe().mbx_update_extra(hit["qs"], hit["qe"], hit["ts"], hit["te"], hit.get("rev", 0), len(cigar), c0,
                              np.ascontiguousarray(qseq, np.uint8), np.ascontiguousarray(tseq, np.uint8),
                              np.ascontiguousarray(mat, np.int8), q, e, log_gap, C.byref(h), out, cap)
    d = {f: int(getattr(h, f)) for f in ("qs", "qe", "ts", "te", "n_cigar", "blen", "mlen", "dp_max")}
    d["n_ambi"] = int(h.n_ambi)
    return d, [int(x) for x in out[:h.n_cigar]]


def _sai_array(intervals):
    a = np.ascontiguousarray(np.asarray(intervals, np.uint64).reshape(-1, 4))
    return a


def orc_anchor(index, ctg_len, intervals, qlen, max_occ=250):
    """mb_anchor restated (no methylation) on one read's intervals ((n,4) uint64: x0, x1, size, info) with
    an OracleIndex -> ANCHOR_DTYPE array."""
    u = _sai_array(intervals).copy()
    ctg_len = np.ascontiguousarray(ctg_len, np.uint64)
    ctg_off = np.zeros(len(ctg_len), np.uint64)
    if len(ctg_len) > 1:
        np.cumsum(ctg_len[:-1], out=ctg_off[1:])
    cap = int(np.minimum(u[:, 2], max_occ).sum()) + 1 if len(u) else 1
    out = np.zeros(cap, ANCHOR_DTYPE)
    n = orc().orc_anchor(index.b, len(ctg_len), ctg_off, ctg_len, u.ctypes.data_as(C.c_void_p), len(u), qlen, max_occ,
                         out.ctypes.data_as(C.c_void_p), cap)
    if n < 0:
        raise RuntimeError("orc_anchor: capacity")
    return out[:n]


# mb_lchain_dp's parameters (lchain.c:145): max_dist_x, max_dist_y, bw, max_skip, max_iter, min_sc and the
# gap scale chn_pen_gap = chain_gap_scale * .01 * min_len (map-algo.c:607); short-read preset defaults
CHAIN_SR = dict(max_dist_x=100, max_dist_y=100, bw=100, max_skip=25, max_iter=5000, min_sc=25, pen_gap=0.8 * .01 * 19)


# option sets of the chaining tests: the short-read preset, the long-read preset (options.c:70-82), the
# re-chaining band (bw_long, map-algo.c:618-632), and tight skip/iteration limits that reach the breaks
CHAIN_PARAM_SETS = [
    CHAIN_SR,
    dict(max_dist_x=5000, max_dist_y=5000, bw=500, max_skip=25, max_iter=5000, min_sc=40, pen_gap=0.8 * .01 * 19),
    dict(max_dist_x=20000, max_dist_y=5000, bw=20000, max_skip=25, max_iter=5000, min_sc=40, pen_gap=0.8 * .01 * 19),
    dict(max_dist_x=400, max_dist_y=300, bw=60, max_skip=1, max_iter=12, min_sc=19, pen_gap=0.05),
    dict(max_dist_x=50, max_dist_y=50, bw=10, max_skip=0, max_iter=3, min_sc=1, pen_gap=0.5),
]


def _chain_args(p):
    return (int(p["max_dist_x"]), int(p["max_dist_y"]), int(p["bw"]), int(p["max_skip"]), int(p["max_iter"]), int(p["min_sc"]),
            float(p["pen_gap"]))


def orc_lchain(ctg_len, anchors, params=CHAIN_SR):
    """mb_lchain_dp restated on one read's anchors (ANCHOR_DTYPE, mb_anchor's order) -> (u: uint64 score<<32 |
    count per chain, chained anchors ANCHOR_DTYPE)."""
    a = np.ascontiguousarray(anchors, ANCHOR_DTYPE)
    ctg_len = np.ascontiguousarray(ctg_len, np.uint64)
    ctg_off = np.zeros(len(ctg_len), np.uint64)
    if len(ctg_len) > 1:
        np.cumsum(ctg_len[:-1], out=ctg_off[1:])
    u = np.zeros(max(len(a), 1), np.uint64)
    b = np.zeros(max(len(a), 1), ANCHOR_DTYPE)
    nb = C.c_int64(0)
    n_u = orc().orc_lchain(len(ctg_len), ctg_off, ctg_len, *_chain_args(params), len(a), a.ctypes.data_as(C.c_void_p),
                           u.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p), C.byref(nb))
    if n_u < 0:
        raise MemoryError("orc_lchain")
    return u[:n_u].copy(), b[:nb.value].copy()


class _Bseq1(C.Structure):       # mb_bseq1_t (bseq.h:14-18)
    _fields_ = [("l_seq", C.c_uint64), ("id", C.c_uint64), ("name", C.c_void_p), ("seq", C.c_void_p), ("qual", C.c_void_p),
                ("comment", C.c_void_p)]


class RefBseq:
    """The reference's own mb_bseq_open / mb_bseq_read / mb_bseq_read_frag (bseq.c:41-161), compiled in
    oracle/_ref/libmbref.so."""

    def __init__(self, path):
        L = ref()
        L.mb_bseq_open.restype = C.c_void_p
        L.mb_bseq_open.argtypes = [C.c_char_p]
        L.mb_bseq_close.argtypes = [C.c_void_p]
        L.mb_bseq_read.restype = C.POINTER(_Bseq1)
        L.mb_bseq_read.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, C.POINTER(C.c_int)]
        L.mb_bseq_read_frag.restype = C.POINTER(_Bseq1)
        L.mb_bseq_read_frag.argtypes = [C.c_int, C.POINTER(C.c_void_p), C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_int)]
        self.L = L
        self.p = L.mb_bseq_open(path.encode())
        if not self.p:
            raise IOError(path)

    @staticmethod
    def _take(a, n):
        free = C.CDLL(None).free
        free.argtypes = [C.c_void_p]
        out = []
        for i in range(n):
            r = a[i]
            get = (lambda p: C.string_at(p) if p else None)
            out.append((C.string_at(r.name), C.string_at(r.seq, r.l_seq), C.string_at(r.qual, r.l_seq) if r.qual else None,
                        get(r.comment)))
            for f in (r.name, r.seq, r.qual, r.comment):
                if f:
                    free(f)
        if a:
            free(C.cast(a, C.c_void_p))
        return out

    def read(self, chunk_size, with_qual=True, with_comment=True, frag_mode=False, min_cnt=0, max_chunk_size=0):
        n = C.c_int(0)
        a = self.L.mb_bseq_read(self.p, chunk_size, int(with_qual), int(with_comment), int(frag_mode), min_cnt, max_chunk_size,
                                C.byref(n))
        return self._take(a, n.value)

    @staticmethod
    def read_frag(readers, chunk_size, with_qual=True, with_comment=True):
        L = readers[0].L
        ps = (C.c_void_p * len(readers))(*[r.p for r in readers])
        n = C.c_int(0)
        a = L.mb_bseq_read_frag(len(readers), ps, chunk_size, int(with_qual), int(with_comment), C.byref(n))
        return RefBseq._take(a, n.value)

    def close(self):
        if self.p:
            self.L.mb_bseq_close(self.p)
            self.p = None


class RefMapIndex:
    """A prefix loaded by the reference's mb_idx_load (minibwa.h), for mb_anchor."""

    def __init__(self, prefix):
        self.L = ref_pe()
        self.p = self.L.mb_idx_load(prefix.encode(), 0)
        if not self.p:
            raise IOError(prefix)

    def anchor(self, intervals, qlen, max_occ=250):
        u = _sai_array(intervals)
        cap = int(np.minimum(u[:, 2], max_occ).sum()) + 1 if len(u) else 1
        out = np.zeros(cap, ANCHOR_DTYPE)
        n = self.L.mbx_anchor(C.c_void_p(self.p), u.ctypes.data_as(C.c_void_p), len(u), qlen, max_occ,
                              out.ctypes.data_as(C.c_void_p), cap)
        return out[:min(n, cap)]

    def lchain(self, anchors, params=CHAIN_SR):
        """The reference's own mb_lchain_dp on one read's anchors -> (u, chained anchors)."""
        a = np.ascontiguousarray(anchors, ANCHOR_DTYPE)
        u = np.zeros(max(len(a), 1), np.uint64)
        b = np.zeros(max(len(a), 1), ANCHOR_DTYPE)
        nb = C.c_int64(0)
        n_u = self.L.mbx_lchain(C.c_void_p(self.p), *_chain_args(params), len(a), a.ctypes.data_as(C.c_void_p),
                                u.ctypes.data_as(C.c_void_p), b.ctypes.data_as(C.c_void_p), C.byref(nb))
        return u[:n_u].copy(), b[:nb.value].copy()

    def close(self):
        if self.p:
            self.L.mb_idx_destroy(C.c_void_p(self.p))
            self.p = None
