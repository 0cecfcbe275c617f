"""g3sa_amd: MI355X-native BWA-MEM short-read alignment hot path.

The product is the native library ``_build/libg3a.so`` (HIP kernels behind the
C ABI in ``include/g3a.h``).  This module binds it with ctypes; it never falls
back to a CPU implementation: if the library is missing it raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
BUILD = PKG / "_build"
LIB_PATH = BUILD / "libg3a.so"
CLI_PATH = BUILD / "g3a"
GEN_PATH = BUILD / "g3a-gen"

G3A_OK = 0
FORM_PLAIN = 0
FORM_BATCHED = 1
FORMS = {"plain": FORM_PLAIN, "batched": FORM_BATCHED}


class Opt(C.Structure):
    """g3a_opt (include/g3a.h)."""

    _fields_ = [
        ("a", C.c_int), ("b", C.c_int), ("o_del", C.c_int), ("e_del", C.c_int),
        ("o_ins", C.c_int), ("e_ins", C.c_int), ("w", C.c_int), ("T", C.c_int),
        ("pen_clip5", C.c_int), ("pen_clip3", C.c_int),
        ("min_seed_len", C.c_int), ("split_width", C.c_int), ("max_occ", C.c_int),
        ("max_chain_gap", C.c_int), ("flag", C.c_int),
        ("max_mem_intv", C.c_int64),
        ("split_factor", C.c_float), ("mask_level", C.c_float), ("drop_ratio", C.c_float),
        ("mask_level_redun", C.c_float),
        ("mat", C.c_int8 * 25),
    ]


class BatchOut(C.Structure):
    """g3a_batch_out (include/g3a.h)."""

    _fields_ = [
        ("n_reads", C.c_int32),
        ("aln_off", C.POINTER(C.c_int64)),
        ("n_alns", C.c_int64),
        ("rid", C.POINTER(C.c_int32)), ("pos", C.POINTER(C.c_int64)), ("flag", C.POINTER(C.c_int32)),
        ("is_rev", C.POINTER(C.c_int32)), ("score", C.POINTER(C.c_int32)), ("nw_score", C.POINTER(C.c_int32)),
        ("nm", C.POINTER(C.c_int32)), ("qb", C.POINTER(C.c_int32)), ("qe", C.POINTER(C.c_int32)),
        ("rb", C.POINTER(C.c_int64)), ("re", C.POINTER(C.c_int64)),
        ("n_cigar", C.POINTER(C.c_int32)), ("cigar_off", C.POINTER(C.c_int64)),
        ("n_core", C.POINTER(C.c_int32)), ("core_off", C.POINTER(C.c_int64)),
        ("cigar", C.POINTER(C.c_uint32)), ("n_cigar_words", C.c_int64),
        ("ms_seed", C.c_float), ("ms_sa", C.c_float), ("ms_chain", C.c_float),
        ("ms_extend", C.c_float), ("ms_traceback", C.c_float), ("ms_total", C.c_float),
        ("n_seeds", C.c_int64), ("n_tasks", C.c_int64), ("ext_cells", C.c_int64),
        ("fm_steps", C.c_int64), ("n_slow_reads", C.c_int64),
        ("mapq", C.POINTER(C.c_int32)), ("md_off", C.POINTER(C.c_int64)), ("md", C.POINTER(C.c_char)),
        ("n_md_bytes", C.c_int64),
    ]


class PeStat(C.Structure):
    """g3a_pestat: insert-size statistics of one pair orientation."""
    _fields_ = [("low", C.c_int32), ("high", C.c_int32), ("failed", C.c_int32), ("avg", C.c_double),
                ("std", C.c_double)]


class LearnedInfo(C.Structure):
    """g3a_learned_info (include/g3a.h)."""

    _fields_ = [
        ("rows", C.c_int64), ("distinct_keys", C.c_int64),
        ("leaves", C.c_int32), ("hidden_units", C.c_int32), ("window_cap", C.c_int32), ("pad", C.c_int32),
        ("model_bytes", C.c_int64), ("sa_bytes", C.c_int64), ("key_bytes", C.c_int64),
        ("max_err", C.c_int64), ("keys_over_window", C.c_int64), ("leaves_over_window", C.c_int64),
        ("mean_err", C.c_double), ("build_ms", C.c_float), ("pad2", C.c_float), ("knots", C.c_int64),
    ]


EXPORTS = [
    "g3a_opt_init", "g3a_device_alloc", "g3a_memcpy_index", "g3a_memcpy_input", "g3a_bwa_align",
    "g3a_memcpy_output", "g3a_align_batch", "g3a_batch_out_free", "g3a_destroy", "g3a_last_error",
    "g3a_build_index", "g3a_last_stats", "g3a_set_form", "g3a_set_graph", "g3a_graph_info", "g3a_set_batch_caps", "g3a_device_alloc_sibling", "g3a_set_seed_grid", "g3a_set_chain_serial", "g3a_set_tuning", "g3a_kernel_ms", "g3a_autotune",
    "g3a_set_seed_blocks", "g3a_autotune_seed", "g3a_set_seed_order", "g3a_set_learned", "g3a_learned_stats",
    "g3a_last_intervals", "g3a_set_ext_packed", "g3a_ext_packed", "g3a_build_index_dev", "g3a_memcpy_text",
    "g3a_format_records", "g3a_copy_text", "g3a_bwa_align_noalloc", "g3a_set_wide_layout",
    "g3a_mem_pestat", "g3a_mem_pair",
]

_lib = None


def load_library() -> C.CDLL:
    """Load libg3a.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
    lib = C.CDLL(str(LIB_PATH))
    lib.g3a_opt_init.argtypes = [C.POINTER(Opt)]
    lib.g3a_device_alloc.argtypes = [C.c_int, C.c_char_p, C.POINTER(Opt), C.c_int, C.c_int64, C.POINTER(C.c_void_p)]
    lib.g3a_memcpy_index.argtypes = [C.c_void_p]
    lib.g3a_memcpy_input.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    lib.g3a_bwa_align.argtypes = [C.c_void_p]
    lib.g3a_bwa_align_noalloc.argtypes = [C.c_void_p]
    lib.g3a_memcpy_output.argtypes = [C.c_void_p, C.POINTER(BatchOut)]
    lib.g3a_align_batch.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.POINTER(BatchOut)]
    lib.g3a_batch_out_free.argtypes = [C.POINTER(BatchOut)]
    lib.g3a_destroy.argtypes = [C.c_void_p]
    lib.g3a_last_error.argtypes = [C.c_void_p]
    lib.g3a_last_error.restype = C.c_char_p
    lib.g3a_build_index.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    lib.g3a_build_index_dev.argtypes = [C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.c_char_p, C.c_int]
    lib.g3a_memcpy_text.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.g3a_format_records.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int64)]
    lib.g3a_copy_text.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
    lib.g3a_last_stats.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_int64)]
    lib.g3a_set_form.argtypes = [C.c_void_p, C.c_int]
    lib.g3a_set_graph.argtypes = [C.c_void_p, C.c_int]
    lib.g3a_graph_info.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    lib.g3a_set_batch_caps.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
    lib.g3a_device_alloc_sibling.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_void_p)]
    lib.g3a_set_seed_grid.argtypes = [C.c_void_p, C.c_int, C.c_int]
    lib.g3a_set_chain_serial.argtypes = [C.c_void_p, C.c_int]
    lib.g3a_kernel_ms.argtypes = [C.c_void_p, C.POINTER(C.c_float)]
    lib.g3a_autotune.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)]
    lib.g3a_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    lib.g3a_set_seed_blocks.argtypes = [C.c_void_p, C.c_int]
    lib.g3a_autotune_seed.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(C.c_float)]
    lib.g3a_set_seed_order.argtypes = [C.c_void_p, C.c_int]
    lib.g3a_set_ext_packed.argtypes = [C.c_void_p, C.c_int]
    lib.g3a_set_wide_layout.argtypes = [C.c_void_p, C.c_int]
    lib.g3a_mem_pestat.argtypes = [C.c_void_p, C.c_int, C.POINTER(PeStat)]
    lib.g3a_mem_pair.argtypes = [C.c_void_p, C.POINTER(PeStat), C.c_int64, C.c_void_p]
    lib.g3a_ext_packed.argtypes = [C.c_void_p]
    lib.g3a_set_learned.argtypes = [C.c_void_p, C.c_int, C.POINTER(LearnedInfo)]
    lib.g3a_learned_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
    lib.g3a_last_intervals.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64)]
    _lib = lib
    return lib


def default_opt() -> Opt:
    o = Opt()
    load_library().g3a_opt_init(C.byref(o))
    return o


def build_index(fasta: str, prefix: str, sa_intv: int = 4) -> None:
    lib = load_library()
    err = C.create_string_buffer(512)
    rc = lib.g3a_build_index(str(fasta).encode(), str(prefix).encode(), sa_intv, err, 512)
    if rc != G3A_OK:
        raise RuntimeError(f"g3a_build_index failed ({rc}): {err.value.decode()}")


def build_index_dev(fasta: str, prefix: str, sa_intv: int = 4, device: int = 0) -> None:
    """The same index files, with the suffix sort and BWT built on a HIP device."""
    lib = load_library()
    err = C.create_string_buffer(512)
    rc = lib.g3a_build_index_dev(device, str(fasta).encode(), str(prefix).encode(), sa_intv, err, 512)
    if rc != G3A_OK:
        raise RuntimeError(f"g3a_build_index_dev failed ({rc}): {err.value.decode()}")


class Result:
    """Alignments of one batch as numpy arrays (copied out of g3a_batch_out)."""

    FIELDS = ["rid", "pos", "flag", "is_rev", "score", "nw_score", "nm", "qb", "qe", "rb", "re",
              "n_cigar", "cigar_off", "n_core", "core_off"]

    def __init__(self, out: BatchOut):
        n, na = out.n_reads, out.n_alns
        self.n_reads = n
        self.n_alns = na
        self.aln_off = np.ctypeslib.as_array(out.aln_off, shape=(n + 1,)).copy()
        for f in self.FIELDS:
            ptr = getattr(out, f)
            setattr(self, f, np.ctypeslib.as_array(ptr, shape=(max(na, 1),))[:na].copy())
        nw = out.n_cigar_words
        self.cigar = np.ctypeslib.as_array(out.cigar, shape=(max(nw, 1),))[:nw].copy()
        self.timing = {k: getattr(out, k) for k in
                       ["ms_seed", "ms_sa", "ms_chain", "ms_extend", "ms_traceback", "ms_total"]}
        self.work = {k: getattr(out, k) for k in ["n_seeds", "n_tasks", "ext_cells", "fm_steps", "n_slow_reads"]}
        self.mapq = np.ctypeslib.as_array(out.mapq, shape=(max(na, 1),))[:na].copy()
        md_off = np.ctypeslib.as_array(out.md_off, shape=(na + 1,)).copy()
        blob = C.string_at(out.md, out.n_md_bytes)
        self.md = [blob[md_off[i]:md_off[i + 1] - 1].decode() for i in range(na)]


class Aligner:
    """A device context: index in HBM, batches aligned through the C ABI."""

    def __init__(self, prefix: str, device: int = 0, max_reads: int = 1 << 20, max_bases: int | None = None,
                 opt: Opt | None = None, form: str = "batched", tuning: tuple[int, int, int] | None = None,
                 wide_layout: bool = False):
        self.lib = load_library()
        self.ctx = C.c_void_p()
        if max_bases is None:
            max_bases = max_reads * 320
        rc = self.lib.g3a_device_alloc(device, str(prefix).encode(), C.byref(opt) if opt else None,
                                       max_reads, max_bases, C.byref(self.ctx))
        if rc != G3A_OK:
            raise RuntimeError(f"g3a_device_alloc failed ({rc})")
        if wide_layout:
            self._check(self.lib.g3a_set_wide_layout(self.ctx, 1))
        self._check(self.lib.g3a_memcpy_index(self.ctx))
        self.set_form(form)
        if tuning is not None:
            self.set_tuning(*tuning)

    def sibling(self, max_reads: int, max_bases: int | None = None) -> "Aligner":
        """A second context sharing this one's device index (its own stream and
        workspace, this one's form and tuning); batches on the two can run at
        the same time from two threads."""
        if max_bases is None:
            max_bases = max_reads * 320
        ctx = C.c_void_p()
        self._check(self.lib.g3a_device_alloc_sibling(self.ctx, max_reads, max_bases, C.byref(ctx)))
        al = Aligner.__new__(Aligner)
        al.lib = self.lib
        al.ctx = ctx
        return al

    def set_graph(self, on: bool) -> None:
        """Run each batch as one HIP graph (captured again when its shape changes)."""
        self._check(self.lib.g3a_set_graph(self.ctx, int(bool(on))))

    def graph_info(self) -> dict:
        """Whether the last batch ran as a graph, captures so far, overflow reruns."""
        v = (C.c_int64 * 3)()
        self._check(self.lib.g3a_graph_info(self.ctx, v))
        return {"graph_used": bool(v[0]), "captures": v[1], "overflow_reruns": v[2]}

    def set_batch_caps(self, seeds: float, tasks: float, alns: float, cig_used: float, z_per_base: float,
                       cig_per_base: float, slow_reads: int) -> None:
        """Capacities the next batch starts from (later batches raise them)."""
        v = (C.c_double * 7)(seeds, tasks, alns, cig_used, z_per_base, cig_per_base, slow_reads)
        self._check(self.lib.g3a_set_batch_caps(self.ctx, v))

    def set_form(self, form: str) -> None:
        """'plain' (one wavefront per extension task) or 'batched'."""
        self._check(self.lib.g3a_set_form(self.ctx, FORMS[form]))

    def autotune(self) -> dict:
        """Pick the batched form's fastest tile shape on the last aligned batch."""
        sr, lb, ms = C.c_int(), C.c_int(), C.c_float()
        self._check(self.lib.g3a_autotune(self.ctx, C.byref(sr), C.byref(lb), C.byref(ms)))
        return {"strip_rows": sr.value, "lane_blocks_per_cu": lb.value, "packed": self.lib.g3a_ext_packed(self.ctx),
                "extension_ms": ms.value}

    def set_ext_packed(self, packed: bool) -> None:
        """Lane engine with two strip rows per register (True) or one cell per lane."""
        self._check(self.lib.g3a_set_ext_packed(self.ctx, int(bool(packed))))

    def set_tuning(self, strip_rows: int, lane_blocks_per_cu: int, wave_blocks_per_cu: int) -> None:
        self._check(self.lib.g3a_set_tuning(self.ctx, strip_rows, lane_blocks_per_cu, wave_blocks_per_cu))

    def set_seed_blocks(self, blocks_per_cu: int) -> None:
        """Persistent seeding blocks per compute unit (0: occupancy maximum)."""
        self._check(self.lib.g3a_set_seed_blocks(self.ctx, blocks_per_cu))

    def mem_pestat(self, max_ins: int = 10000) -> list[dict]:
        """Insert-size statistics of the last batch, read as pairs (2p, 2p+1),
        per orientation FF, FR, RF, RR (bwa/src/aux.cu:435-500)."""
        pes = (PeStat * 4)()
        self._check(self.lib.g3a_mem_pestat(self.ctx, int(max_ins), pes))
        return [{"low": p.low, "high": p.high, "failed": p.failed, "avg": p.avg, "std": p.std} for p in pes]

    def mem_pair(self, pes: list[dict], id0: int = 0) -> np.ndarray:
        """Best pair of hits per pair of the last batch (bwa/src/aux.cu:1000-1070):
        rows of (score, sub, n_sub, hit of read 1, hit of read 2, proper-pair bit)."""
        arr = (PeStat * 4)(*[PeStat(p["low"], p["high"], p["failed"], p["avg"], p["std"]) for p in pes])
        n_pairs = self.n_reads // 2
        out = np.zeros((max(n_pairs, 1), 6), dtype=np.int32)
        self._check(self.lib.g3a_mem_pair(self.ctx, arr, int(id0), out.ctypes.data))
        return out[:n_pairs]

    def set_chain_serial(self, on: bool) -> None:
        """Chain every read by the B-tree code on one thread (True) or by the tiers (False)."""
        self._check(self.lib.g3a_set_chain_serial(self.ctx, int(bool(on))))

    def set_seed_grid(self, threads_per_block: int, blocks_per_cu: int) -> None:
        """Seeding threads per block (128-1024) and blocks per compute unit."""
        self._check(self.lib.g3a_set_seed_grid(self.ctx, threads_per_block, blocks_per_cu))

    def set_seed_order(self, ordered: bool) -> None:
        """Seed reads in minimizer order (True) or input order (False)."""
        self._check(self.lib.g3a_set_seed_order(self.ctx, int(bool(ordered))))

    def autotune_seed(self) -> dict:
        """Re-align the last batch per seeding grid and order; keep the fastest."""
        t, b, o, ms = C.c_int(), C.c_int(), C.c_int(), C.c_float()
        self._check(self.lib.g3a_autotune_seed(self.ctx, C.byref(t), C.byref(b), C.byref(o), C.byref(ms)))
        return {"seed_threads_per_block": t.value, "seed_blocks_per_cu": b.value, "seed_ordered": o.value,
                "seed_stage_ms": ms.value}

    def set_learned(self, on: bool | int) -> dict:
        """Switch seeding to the learned-index form (built on first use) or back
        (2: full SA without the model's lookups); returns the model's figures."""
        info = LearnedInfo()
        self._check(self.lib.g3a_set_learned(self.ctx, int(on), C.byref(info)))
        return {k: getattr(info, k) for k, _ in LearnedInfo._fields_ if not k.startswith("pad")}

    def learned_stats(self) -> dict:
        """Learned-form work of the last batch."""
        v = (C.c_int64 * 6)()
        self._check(self.lib.g3a_learned_stats(self.ctx, v))
        return dict(zip(["lookups", "fm_answered", "text_runs", "text_run_bases", "model_narrowed",
                         "model_window_misses"], list(v)))

    def intervals(self) -> tuple[np.ndarray, np.ndarray]:
        """Seeding intervals of the last batch: per-read counts and (k, s, m, n) rows."""
        n = self.n_reads
        counts = np.zeros(max(n, 1), dtype=np.int32)
        total = C.c_int64()
        self._check(self.lib.g3a_last_intervals(self.ctx, counts.ctypes.data, None, 0, C.byref(total)))
        kmsn = np.zeros((max(total.value, 1), 4), dtype=np.int64)
        self._check(self.lib.g3a_last_intervals(self.ctx, counts.ctypes.data, kmsn.ctypes.data, total.value,
                                                C.byref(total)))
        return counts[:n], kmsn[:total.value]

    def _check(self, rc: int) -> None:
        if rc != G3A_OK:
            raise RuntimeError(f"g3a call failed ({rc}): {self.lib.g3a_last_error(self.ctx).decode()}")

    def upload(self, seq: np.ndarray, offsets: np.ndarray) -> None:
        seq = np.ascontiguousarray(seq, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.int32)
        self._keep = (seq, offsets)
        self.n_reads = len(offsets) - 1
        self._check(self.lib.g3a_memcpy_input(self.ctx, len(offsets) - 1, seq.ctypes.data, offsets.ctypes.data))

    def run(self) -> None:
        self._check(self.lib.g3a_bwa_align(self.ctx))

    def run_noalloc(self) -> None:
        """The allocation-free entry point on the resident batch."""
        self._check(self.lib.g3a_bwa_align_noalloc(self.ctx))

    def download(self) -> Result:
        out = BatchOut()
        self._check(self.lib.g3a_memcpy_output(self.ctx, C.byref(out)))
        try:
            return Result(out)
        finally:
            self.lib.g3a_batch_out_free(C.byref(out))

    def stats(self) -> dict:
        ms = (C.c_float * 6)()
        work = (C.c_int64 * 6)()
        self._check(self.lib.g3a_last_stats(self.ctx, ms, work))
        keys_ms = ["seed", "sa", "chain", "extend", "traceback", "total"]
        keys_w = ["n_seeds", "n_tasks", "ext_cells", "fm_steps", "n_slow_reads", "n_alns"]
        kms = (C.c_float * 4)()
        self._check(self.lib.g3a_kernel_ms(self.ctx, kms))
        keys_k = ["k_seed", "ext_lane", "ext_wave", "k_traceback"]
        return {"ms": dict(zip(keys_ms, list(ms))), "work": dict(zip(keys_w, list(work))),
                "kernel_ms": dict(zip(keys_k, list(kms)))}

    def upload_text(self, names: bytes, name_off: np.ndarray, quals: bytes, qual_off: np.ndarray) -> None:
        """Names and qualities of the resident batch, for the device formatter."""
        name_off = np.ascontiguousarray(name_off, dtype=np.int64)
        qual_off = np.ascontiguousarray(qual_off, dtype=np.int64)
        nb = np.frombuffer(names, dtype=np.uint8) if len(names) else np.zeros(1, np.uint8)
        qb = np.frombuffer(quals, dtype=np.uint8) if len(quals) else np.zeros(1, np.uint8)
        self._keep_text = (nb, qb, name_off, qual_off)
        self._check(self.lib.g3a_memcpy_text(self.ctx, len(name_off) - 1, nb.ctypes.data, name_off.ctypes.data,
                                             qb.ctypes.data, qual_off.ctypes.data))

    def format(self, sam: bool = True) -> int:
        """Format the last batch's records on the device; returns their bytes."""
        n = C.c_int64()
        self._check(self.lib.g3a_format_records(self.ctx, 1 if sam else 0, C.byref(n)))
        self._text_bytes = n.value
        return n.value

    def text(self) -> bytes:
        """Copy the records of the last format() to the host."""
        buf = C.create_string_buffer(max(self._text_bytes, 1))
        self._check(self.lib.g3a_copy_text(self.ctx, buf, self._text_bytes))
        return buf.raw[:self._text_bytes]

    def align(self, seq: np.ndarray, offsets: np.ndarray) -> Result:
        self.upload(seq, offsets)
        self.run()
        return self.download()

    def close(self) -> None:
        if self.ctx:
            self.lib.g3a_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def read_fastq(path: str, limit: int = 0):
    """Parse FASTQ with the reference's rules: name up to the first space,
    sequence encoded A0 C1 G2 T3 (case-insensitive), anything else 4."""
    names, seqs, quals = [], [], []
    table = np.full(256, 4, dtype=np.uint8)
    for i, ch in enumerate(b"ACGT"):
        table[ch] = i
        table[ch + 32] = i
    with open(path, "rb") as f:
        data = f.read()
    lines = data.split(b"\n")
    i = 0
    while i + 3 < len(lines) + 1 and i < len(lines) and lines[i].startswith(b"@"):
        hdr = lines[i][1:]
        names.append(hdr.split(b" ", 1)[0])
        seqs.append(table[np.frombuffer(lines[i + 1], dtype=np.uint8)])
        quals.append(lines[i + 3] if i + 3 < len(lines) else b"")
        i += 4
        if limit and len(names) >= limit:
            break
    offsets = np.zeros(len(seqs) + 1, dtype=np.int32)
    offsets[1:] = np.cumsum([len(s) for s in seqs]) if seqs else []
    seq = np.concatenate(seqs) if seqs else np.zeros(0, dtype=np.uint8)
    return names, seq, offsets, quals
