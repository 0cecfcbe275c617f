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
        self.lib = lib
        self.opt = OrcOpt()
        lib.orc_opt_init(C.byref(self.opt))

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
