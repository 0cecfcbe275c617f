"""ctypes binding of the CPU oracle (oracle/liboracle.so) for the tests."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "_build" / "liboracle.so"
CLI = ROOT / "oracle" / "_build" / "g3a_oracle"


class OrcOpt(C.Structure):
    _fields_ = [
        ("a", C.c_int), ("b", C.c_int), ("o_del", C.c_int), ("e_del", C.c_int), ("o_ins", C.c_int),
        ("e_ins", C.c_int), ("w", C.c_int), ("T", C.c_int), ("pen_clip5", C.c_int), ("pen_clip3", C.c_int),
        ("min_seed_len", C.c_int), ("split_width", C.c_int), ("max_occ", C.c_int), ("max_chain_gap", C.c_int),
        ("flag", C.c_int), ("max_mem_intv", C.c_uint64),
        ("split_factor", C.c_float), ("mask_level", C.c_float), ("drop_ratio", C.c_float),
        ("mask_level_redun", C.c_float), ("mapQ_coef_len", C.c_int), ("mapQ_coef_fac", C.c_int),
        ("mat", C.c_int8 * 25),
    ]


class OrcIntv(C.Structure):
    _fields_ = [("k", C.c_int64), ("l", C.c_int64), ("s", C.c_int64), ("m", C.c_int32), ("n", C.c_int32)]


class OrcSeed(C.Structure):
    _fields_ = [("rbeg", C.c_int64), ("qbeg", C.c_int32), ("len", C.c_int32), ("rid", C.c_int32)]


class OrcAln(C.Structure):
    _fields_ = [
        ("rid", C.c_int32), ("is_rev", C.c_int32), ("pos", C.c_int64), ("flag", C.c_int32), ("score", C.c_int32),
        ("nw_score", C.c_int32), ("NM", C.c_int32), ("qb", C.c_int32), ("qe", C.c_int32), ("rb", C.c_int64),
        ("re", C.c_int64), ("n_cigar", C.c_int32), ("n_core", C.c_int32), ("cigar_off", C.c_int64),
    ]


class OrcBatch(C.Structure):
    _fields_ = [
        ("n_reads", C.c_int32), ("aln_off", C.POINTER(C.c_int32)), ("n_alns", C.c_int64),
        ("alns", C.POINTER(OrcAln)), ("cigar", C.POINTER(C.c_uint32)), ("n_cigar_words", C.c_int64),
        ("fm_steps", C.c_int64), ("lf_steps", C.c_int64), ("sa_reads", C.c_int64), ("ext_cells", C.c_int64),
        ("ext_tasks", C.c_int64), ("tb_cells", C.c_int64),
    ]


class OrcPeStat(C.Structure):
    _fields_ = [("low", C.c_int), ("high", C.c_int), ("failed", C.c_int), ("avg", C.c_double), ("std", C.c_double)]


def read_ann(prefix) -> tuple[int, np.ndarray]:
    """l_pac and contig offsets from a BWA .ann file."""
    lines = Path(str(prefix) + ".ann").read_text().split("\n")
    l_pac, n_seqs = int(lines[0].split()[0]), int(lines[0].split()[1])
    offs = [int(lines[2 + 2 * i].split()[0]) for i in range(n_seqs)]
    return l_pac, np.array(offs, dtype=np.int64)


class Oracle:
    def __init__(self, lib_path: Path | None = None):
        path = Path(lib_path) if lib_path is not None else LIB
        if not path.exists():
            raise RuntimeError(f"{path} missing: build the oracle (make -C oracle)")
        lib = C.CDLL(str(path))
        lib.orc_opt_init.argtypes = [C.POINTER(OrcOpt)]
        lib.orc_index_load.argtypes = [C.c_char_p]
        lib.orc_index_load.restype = C.c_void_p
        lib.orc_index_free.argtypes = [C.c_void_p]
        lib.orc_occ.argtypes = [C.c_void_p, C.c_int64, C.c_int]
        lib.orc_occ.restype = C.c_int64
        lib.orc_sa.argtypes = [C.c_void_p, C.c_int64]
        lib.orc_sa.restype = C.c_int64
        lib.orc_index_primary.argtypes = [C.c_void_p]
        lib.orc_index_primary.restype = C.c_uint64
        lib.orc_seed_intervals.argtypes = [C.c_void_p, C.POINTER(OrcOpt), C.c_void_p, C.c_int, C.POINTER(OrcIntv),
                                           C.c_int, C.POINTER(C.c_int64)]
        lib.orc_read_seeds.argtypes = [C.c_void_p, C.POINTER(OrcOpt), C.c_void_p, C.c_int, C.POINTER(OrcSeed), C.c_int]
        lib.orc_region_dump.argtypes = [C.c_void_p, C.POINTER(OrcOpt), C.c_int] + [C.c_void_p] * 8
        lib.orc_chain_dump.argtypes = [C.c_void_p, C.POINTER(OrcOpt), C.POINTER(OrcSeed), C.c_int, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.POINTER(C.c_int)]
        lib.orc_align_batch.argtypes = [C.c_void_p, C.POINTER(OrcOpt), C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                        C.POINTER(OrcBatch)]
        lib.orc_batch_free.argtypes = [C.POINTER(OrcBatch)]
        ext_args = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                    C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        lib.orc_ksw_extend.argtypes = ext_args + [C.POINTER(C.c_int64)]
        lib.orc_ksw_extend_warp_emulate.argtypes = ext_args
        lib.orc_ksw_global2.argtypes = [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_int, C.POINTER(C.c_int), C.c_void_p, C.c_int]
        lib.orc_index_contig_name.argtypes = [C.c_void_p, C.c_int]
        lib.orc_index_contig_name.restype = C.c_char_p
        lib.orc_index_n_contigs.argtypes = [C.c_void_p]
        lib.orc_mapq.argtypes = [C.POINTER(OrcOpt), C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int64]
        lib.orc_infer_dir.argtypes = [C.c_int64, C.c_int64, C.c_int64, C.POINTER(C.c_int64)]
        lib.orc_cal_sub.argtypes = [C.POINTER(OrcOpt), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_mem_pestat.argtypes = [C.POINTER(OrcOpt), C.c_int64, C.c_int, C.c_int] + [C.c_void_p] * 6 + \
            [C.POINTER(OrcPeStat)]
        lib.orc_hash_64.argtypes = [C.c_uint64]
        lib.orc_hash_64.restype = C.c_uint64
        lib.orc_mem_pair.argtypes = [C.POINTER(OrcOpt), C.c_int64, C.c_void_p, C.POINTER(OrcPeStat), C.c_int,
                                     C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int)]
        self.lib = lib
        self.opt = OrcOpt()
        lib.orc_opt_init(C.byref(self.opt))

    # paired-end primitives ------------------------------------------------
    def infer_dir(self, l_pac, b1, b2):
        d = C.c_int64()
        r = self.lib.orc_infer_dir(l_pac, b1, b2, C.byref(d))
        return r, d.value

    def cal_sub(self, qb, qe, score):
        qb, qe, score = (np.ascontiguousarray(v, dtype=np.int32) for v in (qb, qe, score))
        return self.lib.orc_cal_sub(C.byref(self.opt), len(qb), qb.ctypes.data, qe.ctypes.data, score.ctypes.data)

    def pestat(self, l_pac, aln_off, rid, rb, qb, qe, score, max_ins=10000) -> list[dict]:
        """mem_pestat over reads 2i, 2i+1 as pairs; per-read hits aln_off[r]..aln_off[r+1]."""
        aln_off = np.ascontiguousarray(aln_off, dtype=np.int64)
        rid, qb, qe, score = (np.ascontiguousarray(v, dtype=np.int32) for v in (rid, qb, qe, score))
        rb = np.ascontiguousarray(rb, dtype=np.int64)
        pes = (OrcPeStat * 4)()
        self.lib.orc_mem_pestat(C.byref(self.opt), l_pac, max_ins, len(aln_off) - 1, aln_off.ctypes.data,
                                rid.ctypes.data, rb.ctypes.data, qb.ctypes.data, qe.ctypes.data, score.ctypes.data,
                                pes)
        return [{"low": p.low, "high": p.high, "failed": p.failed, "avg": p.avg, "std": p.std} for p in pes]

    def hash_64(self, key: int) -> int:
        return self.lib.orc_hash_64(key)

    def mem_pair(self, l_pac, ctg_off, pes, pair_id, rid, rb, score, n0):
        """mem_pair of one pair: rid/rb/score hold read 1's n0 hits, then read 2's.
        Returns (score, sub, n_sub, z0, z1)."""
        arr = (OrcPeStat * 4)(*[OrcPeStat(p["low"], p["high"], p["failed"], p["avg"], p["std"]) for p in pes])
        ctg_off = np.ascontiguousarray(ctg_off, dtype=np.int64)
        rid, score = (np.ascontiguousarray(v, dtype=np.int32) for v in (rid, score))
        rb = np.ascontiguousarray(rb, dtype=np.int64)
        sub, n_sub, z = C.c_int(), C.c_int(), (C.c_int * 2)(-1, -1)
        o = self.lib.orc_mem_pair(C.byref(self.opt), l_pac, ctg_off.ctypes.data, arr, pair_id, n0, len(rid) - n0,
                                  rid.ctypes.data, rb.ctypes.data, score.ctypes.data, C.byref(sub), C.byref(n_sub), z)
        return o, sub.value, n_sub.value, z[0], z[1]

    def mat(self, a=1, b=4) -> np.ndarray:
        m = np.full(25, -1, dtype=np.int8)
        for i in range(4):
            for j in range(4):
                m[i * 5 + j] = a if i == j else -b
        return m

    def load(self, prefix):
        h = self.lib.orc_index_load(str(prefix).encode())
        if not h:
            raise RuntimeError(f"oracle cannot load {prefix}")
        return h

    def seed_intervals(self, idx, read: np.ndarray):
        cap = 8192
        buf = (OrcIntv * cap)()
        read = np.ascontiguousarray(read, dtype=np.uint8)
        n = self.lib.orc_seed_intervals(idx, C.byref(self.opt), read.ctypes.data, len(read), buf, cap, None)
        return [(buf[i].k, buf[i].l, buf[i].s, buf[i].m, buf[i].n) for i in range(min(n, cap))]

    def extend(self, q, t, h0, w=100, end_bonus=5, mat=None, gaps=(6, 1, 6, 1), emulate=False):
        q = np.ascontiguousarray(q, dtype=np.uint8)
        t = np.ascontiguousarray(t, dtype=np.uint8)
        mat = self.mat() if mat is None else np.ascontiguousarray(mat, dtype=np.int8)
        qle, tle = C.c_int(), C.c_int()
        od, ed, oi, ei = gaps
        if emulate:
            s = self.lib.orc_ksw_extend_warp_emulate(len(q), q.ctypes.data, len(t), t.ctypes.data, mat.ctypes.data,
                                                     od, ed, oi, ei, w, end_bonus, h0, C.byref(qle), C.byref(tle))
        else:
            s = self.lib.orc_ksw_extend(len(q), q.ctypes.data, len(t), t.ctypes.data, mat.ctypes.data, od, ed, oi, ei,
                                        w, end_bonus, h0, C.byref(qle), C.byref(tle), None)
        return s, qle.value, tle.value

    def global2(self, q, t, w, mat=None, gaps=(6, 1, 6, 1)):
        q = np.ascontiguousarray(q, dtype=np.uint8)
        t = np.ascontiguousarray(t, dtype=np.uint8)
        mat = self.mat() if mat is None else np.ascontiguousarray(mat, dtype=np.int8)
        cap = len(q) + len(t) + 4
        cig = np.zeros(cap, dtype=np.uint32)
        nc = C.c_int()
        od, ed, oi, ei = gaps
        s = self.lib.orc_ksw_global2(len(q), q.ctypes.data, len(t), t.ctypes.data, mat.ctypes.data, od, ed, oi, ei, w,
                                     C.byref(nc), cig.ctypes.data, cap)
        return s, cig[: nc.value].copy()

    def align(self, idx, seq: np.ndarray, offsets: np.ndarray, threads: int = 4, opt: OrcOpt | None = None):
        seq = np.ascontiguousarray(seq, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        b = OrcBatch()
        rc = self.lib.orc_align_batch(idx, C.byref(opt if opt is not None else self.opt), len(offsets) - 1, seq.ctypes.data, offsets.ctypes.data,
                                      threads, C.byref(b))
        if rc != 0:
            raise RuntimeError("oracle alignment failed")
        try:
            n, na = b.n_reads, b.n_alns
            out = {"aln_off": np.ctypeslib.as_array(b.aln_off, shape=(n + 1,)).astype(np.int64)}
            recs = np.ctypeslib.as_array(b.alns, shape=(max(na, 1),))[:na]
            for f, _ in OrcAln._fields_:
                out[f] = recs[f].copy() if na else np.zeros(0, dtype=np.int64)
            nw = b.n_cigar_words
            out["cigar"] = np.ctypeslib.as_array(b.cigar, shape=(max(nw, 1),))[:nw].copy()
            out["work"] = {k: getattr(b, k) for k in ["fm_steps", "lf_steps", "sa_reads", "ext_cells", "ext_tasks",
                                                       "tb_cells"]}
            return out
        finally:
            self.lib.orc_batch_free(C.byref(b))
