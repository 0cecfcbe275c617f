"""GPU parity: the HIP pipeline, called through the C ABI, against the oracle.

Every alignment field and both CIGARs must be identical (integer work:
bit-exact).  Datasets: config 1 exactly (1 Mbp, 10,000 x 150 bp), config-3
style variable-length reads with divergent tails, a repeat-rich 5-contig
reference with paired reads aligned as single-end, and edge-case batches.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import g3sa_amd
from conftest import oracle_threads

pytestmark = pytest.mark.gpu

FIELDS = [("rid", "rid"), ("pos", "pos"), ("flag", "flag"), ("is_rev", "is_rev"), ("score", "score"),
          ("nw_score", "nw_score"), ("nm", "NM"), ("qb", "qb"), ("qe", "qe"), ("rb", "rb"), ("re", "re"),
          ("n_cigar", "n_cigar"), ("n_core", "n_core")]


def compare(gpu: g3sa_amd.Result, ref: dict, label: str) -> None:
    assert gpu.n_reads == len(ref["aln_off"]) - 1, label
    np.testing.assert_array_equal(gpu.aln_off, ref["aln_off"], err_msg=f"{label}: alignments per read differ")
    for g, r in FIELDS:
        a, b = getattr(gpu, g), ref[r]
        if not np.array_equal(a, b):
            bad = np.nonzero(a != b)[0][:5]
            raise AssertionError(f"{label}: field {g} differs at alignments {bad.tolist()}: gpu {a[bad]} oracle {b[bad]}")
    for i in range(gpu.n_alns):
        gc = gpu.cigar[gpu.cigar_off[i]: gpu.cigar_off[i] + gpu.n_cigar[i]]
        rc = ref["cigar"][ref["cigar_off"][i]: ref["cigar_off"][i] + ref["n_cigar"][i]]
        assert np.array_equal(gc, rc), f"{label}: CIGAR of alignment {i}"
        gk = gpu.cigar[gpu.core_off[i]: gpu.core_off[i] + gpu.n_core[i]]
        rk = ref["cigar"][ref["cigar_off"][i] + ref["n_cigar"][i]:
                          ref["cigar_off"][i] + ref["n_cigar"][i] + ref["n_core"][i]]
        assert np.array_equal(gk, rk), f"{label}: global CIGAR of alignment {i}"


# (form, tuning, seed order, seeding grid (threads per block, blocks per
# CU), packed lane arithmetic): the plain form and the batched form at both
# strip heights with two strip rows per register, and at 16 rows with one
# cell per lane; reads seeded in minimizer order and in input order; seeding
# grids of 1024, 512, 256 and 128 threads per block, whose LDS shares hold
# different numbers of LEP entries per thread (the rest of a list spills to
# its global slice)
VARIANTS = [("plain", None, 1, (1024, 1), 1), ("batched", (16, 8, 8), 1, (128, 10), 1),
            ("batched", (32, 4, 8), 0, (256, 4), 1), ("batched", (16, 8, 8), 0, (512, 0), 0)]


def run_both(ds, oracle, seq, off, label, max_reads=None, wide=False, variants=None):
    n = len(off) - 1
    idx = oracle.load(ds.prefix)
    try:
        ref = oracle.align(idx, seq, off, threads=oracle_threads())
    finally:
        oracle.lib.orc_index_free(idx)
    with g3sa_amd.Aligner(str(ds.prefix), max_reads=max(max_reads or n, 1),
                          max_bases=max(int(off[-1]), 1), wide_layout=wide) as al:
        for form, tuning, ordered, blocks, packed in variants or VARIANTS:
            print(f"[parity] {label}: {n} reads, form {form} {tuning} seed order {ordered} blocks {blocks} "
                  f"packed {packed}", flush=True)
            al.set_form(form)
            if tuning:
                al.set_tuning(*tuning)
            al.set_seed_order(ordered)
            al.set_seed_grid(*blocks)
            al.set_ext_packed(packed)
            got = al.align(seq, off)
            compare(got, ref, f"{label} [{form} {tuning} order {ordered}]")
            # seeds resolved by the GPU equal the oracle's SA lookups
            assert got.work["n_seeds"] == ref["work"]["sa_reads"], label
    return got, ref


def test_config1_parity(ds_c1, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["main"]))
    got, ref = run_both(ds_c1, oracle, seq, off, "config1")
    assert got.n_alns >= 9900


def test_config3_style_parity(ds_c1, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["c3"]))
    run_both(ds_c1, oracle, seq, off, "config3-style")


def test_repeats_parity(ds_rep, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_rep.reads["main"]))
    got, _ = run_both(ds_rep, oracle, seq, off, "repeats")


def _batch(reads):
    off = np.zeros(len(reads) + 1, dtype=np.int32)
    off[1:] = np.cumsum([len(r) for r in reads])
    seq = np.concatenate([np.asarray(r, dtype=np.uint8) for r in reads]) if off[-1] else np.zeros(1, np.uint8)
    return seq, off


def test_edge_cases_parity(ds_c1, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["main"]), limit=50)
    reads = [seq[off[i]:off[i + 1]] for i in range(50)]
    rng = np.random.default_rng(7)
    cases = [
        np.zeros(0, np.uint8),                              # empty read
        np.full(150, 4, np.uint8),                          # all N
        reads[0][:18],                                      # shorter than a seed
        reads[1][:19],                                      # exactly one seed long
        np.concatenate([reads[2][:70], [4], reads[2][71:]]),  # N inside
        np.concatenate([reads[3], reads[4], reads[5]]),     # 450 bp: over the 320 bp limit
        np.concatenate([reads[6], reads[7][:170]]),         # 320 bp: at the limit
        rng.integers(0, 4, 150).astype(np.uint8),           # random (unmapped)
        reads[8][::-1].copy(),                              # reversed (not complemented)
        np.tile(np.array([0, 1], np.uint8), 75),            # low complexity
    ] + reads[9:20]
    seq2, off2 = _batch(cases)
    run_both(ds_c1, oracle, seq2, off2, "edge-cases")


def test_single_read_and_repeated_batches(ds_c1, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["main"]), limit=300)
    idx = oracle.load(ds_c1.prefix)
    try:
        with g3sa_amd.Aligner(str(ds_c1.prefix), max_reads=300, max_bases=int(off[-1])) as al:
            for lo, hi in [(0, 1), (0, 300), (17, 18), (100, 250)]:
                s, o = seq[off[lo]:off[hi]], (off[lo:hi + 1] - off[lo]).astype(np.int32)
                got = al.align(s, o)
                ref = oracle.align(idx, s, o, threads=2)
                compare(got, ref, f"batch {lo}:{hi}")
    finally:
        oracle.lib.orc_index_free(idx)


def _with_scoring(oracle, w=None, b=None):
    """Matching option blocks for the oracle and the library."""
    oo = type(oracle.opt)()
    C.pointer(oo)[0] = oracle.opt
    go = g3sa_amd.default_opt()
    for o in (oo, go):
        if w is not None:
            o.w = w
        if b is not None:
            o.b = b
            for i in range(4):
                for j in range(4):
                    o.mat[i * 5 + j] = o.a if i == j else -b
    return oo, go


@pytest.mark.parametrize("w,b", [(20, None), (None, 9)])
def test_scoring_variants_parity(ds_c1, oracle, w, b):
    """A narrow band (the column limit, corner rule and quirk rows of the
    reference's cell set apply to many cells of the lane engine's general
    columns) and a mismatch score outside the lane engine's 4-bit range
    (every task takes the one-wavefront kernel)."""
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["c3"]), limit=1500)
    oo, go = _with_scoring(oracle, w, b)
    idx = oracle.load(ds_c1.prefix)
    try:
        ref = oracle.align(idx, seq, off, threads=oracle_threads(), opt=oo)
    finally:
        oracle.lib.orc_index_free(idx)
    with g3sa_amd.Aligner(str(ds_c1.prefix), max_reads=len(off) - 1, max_bases=int(off[-1]), opt=go) as al:
        for form, tuning, ordered, blocks, packed in VARIANTS:
            print(f"[parity] scoring w={w} b={b}: form {form} {tuning} seed order {ordered} blocks {blocks} "
                  f"packed {packed}", flush=True)
            al.set_form(form)
            if tuning:
                al.set_tuning(*tuning)
            al.set_seed_order(ordered)
            al.set_seed_grid(*blocks)
            al.set_ext_packed(packed)
            compare(al.align(seq, off), ref, f"scoring w={w} b={b} [{form} {tuning} order {ordered}]")


def test_empty_batch(ds_c1):
    with g3sa_amd.Aligner(str(ds_c1.prefix), max_reads=4, max_bases=1000) as al:
        got = al.align(np.zeros(1, np.uint8), np.zeros(1, np.int32))
        assert got.n_reads == 0 and got.n_alns == 0


def test_autotune_keeps_output(ds_c1, oracle):
    """The autotuners rerun the extension stage under every tile shape and the
    whole batch under every seeding grid and order; the batch's output and
    every later batch stay identical to the oracle."""
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["c3"]), limit=2000)
    idx = oracle.load(ds_c1.prefix)
    try:
        ref = oracle.align(idx, seq, off, threads=oracle_threads())
    finally:
        oracle.lib.orc_index_free(idx)
    with g3sa_amd.Aligner(str(ds_c1.prefix), max_reads=len(off) - 1, max_bases=int(off[-1])) as al:
        al.upload(seq, off)
        al.run()
        print("[parity] autotune", flush=True)
        t = al.autotune()
        assert t["strip_rows"] in (16, 32) and t["lane_blocks_per_cu"] in (2, 4, 8, 16) and t["packed"] in (0, 1)
        compare(al.download(), ref, "after autotune")
        compare(al.align(seq, off), ref, "next batch after autotune")
        print("[parity] seeding autotune", flush=True)
        st = al.autotune_seed()
        assert (st["seed_threads_per_block"], st["seed_blocks_per_cu"]) in ((1024, 1), (512, 2), (256, 4), (128, 6))
        assert st["seed_ordered"] in (0, 1)
        compare(al.download(), ref, "after seeding autotune")
        compare(al.align(seq, off), ref, "next batch after seeding autotune")


def test_api_error_paths(ds_c1):
    """Documented failures come back as codes with a message; the context
    stays usable after each."""
    lib = g3sa_amd.load_library()
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["main"]), limit=50)
    ctx = C.c_void_p()
    assert lib.g3a_device_alloc(0, str(ds_c1.prefix).encode(), None, 50, int(off[-1]), C.byref(ctx)) == 0
    try:
        assert lib.g3a_bwa_align(ctx) == -4                     # index not on the device yet
        assert lib.g3a_memcpy_index(ctx) == 0
        assert lib.g3a_bwa_align(ctx) == -4                     # no input yet
        out = g3sa_amd.BatchOut()
        assert lib.g3a_memcpy_output(ctx, C.byref(out)) == -4   # nothing aligned
        bad = off.copy()
        bad[3], bad[4] = bad[4], bad[3]                          # offsets not monotonic
        assert lib.g3a_memcpy_input(ctx, 50, seq.ctypes.data, bad.ctypes.data) == -1
        assert b"non-decreasing" in lib.g3a_last_error(ctx)
        big = np.concatenate([off, off[-1:] + 150]).astype(np.int32)
        assert lib.g3a_memcpy_input(ctx, 51, seq.ctypes.data, big.ctypes.data) == -1   # > max_reads
        assert lib.g3a_set_form(ctx, 7) == -1
        assert lib.g3a_set_tuning(ctx, 24, 8, 8) == -1
        assert lib.g3a_set_seed_grid(ctx, 100, 1) == -1          # not a multiple of 64
        assert lib.g3a_set_seed_grid(ctx, 2048, 1) == -1         # over 1024 threads
        assert lib.g3a_set_seed_grid(ctx, 256, -1) == -1
        assert lib.g3a_set_chain_serial(ctx, 2) == -1
        assert lib.g3a_set_seed_grid(ctx, 256, 4) == 0
        # still usable
        assert lib.g3a_memcpy_input(ctx, 50, seq.ctypes.data, off.ctypes.data) == 0
        assert lib.g3a_bwa_align(ctx) == 0
        assert lib.g3a_memcpy_output(ctx, C.byref(out)) == 0
        assert out.n_reads == 50
        lib.g3a_batch_out_free(C.byref(out))
    finally:
        lib.g3a_destroy(ctx)


def _tandem_dataset(tmp_path):
    """A 200 kbp reference with three 4 kbp tandem repeats (1% point
    divergence) and 400 reads, three in four inside the repeats."""
    rng = np.random.default_rng(17)
    L = 200_000
    g = rng.integers(0, 4, L).astype(np.uint8)
    spans = [(20_000, [0, 1, 2]), (80_000, [3, 1, 0, 2]), (150_000, [2, 2, 0, 1, 3, 3])]
    for start, unit in spans:  # 4 kbp tandem repeats at 1% point divergence
        rep = np.tile(np.array(unit, np.uint8), 4000 // len(unit) + 1)[:4000]
        mut = rng.random(4000) < 0.01
        rep[mut] = rng.integers(0, 4, int(mut.sum()))
        g[start:start + 4000] = rep
    fa = tmp_path / "tandem.fa"
    fa.write_text(">chrT\n" + "\n".join("".join("ACGT"[b] for b in g[i:i + 60]) for i in range(0, L, 60)) + "\n")
    prefix = tmp_path / "tandem"
    g3sa_amd.build_index(str(fa), str(prefix), 8)
    reads = []
    for k in range(400):
        if k % 4 == 3:
            p = int(rng.integers(0, L - 150))
        else:
            s0 = spans[k % 3][0]
            p = int(rng.integers(s0, s0 + 4000 - 150))
        r = g[p:p + 150].copy()
        if k % 2:
            r = (3 - r[::-1]).astype(np.uint8)
        sub = rng.random(150) < 0.005
        r[sub] = rng.integers(0, 4, int(sub.sum()))
        reads.append(r)
    seq, off = _batch(reads)
    return prefix, seq, off


def test_slow_seeding_path_parity(built, oracle, tmp_path):
    """Reads inside long tandem repeats grow their LEP lists past the fast
    seeding path's capacity and are seeded again on the slow path (LDS lists
    of 400 entries); the output still equals the oracle's."""
    from types import SimpleNamespace
    prefix, seq, off = _tandem_dataset(tmp_path)
    got, _ = run_both(SimpleNamespace(prefix=prefix), oracle, seq, off, "tandem-repeats")
    assert got.work["n_slow_reads"] > 0, "no read took the slow seeding path"
    got, ref = run_both(SimpleNamespace(prefix=prefix), oracle, seq, off, "tandem-repeats, wide layout", wide=True,
                        variants=VARIANTS[:2])
    assert got.work["n_slow_reads"] > 0, "no read took the slow seeding path (wide layout)"
    # the slow path of the learned-index form (text runs, k-mer and model lookups)
    with g3sa_amd.Aligner(str(prefix), max_reads=len(off) - 1, max_bases=int(off[-1])) as al:
        al.set_learned(1)
        print("[parity] tandem-repeats, learned form", flush=True)
        got = al.align(seq, off)
        compare(got, ref, "tandem-repeats, learned form")
        assert got.work["n_slow_reads"] > 0, "no read took the slow seeding path (learned form)"


@pytest.mark.parametrize("which", ["c1", "c3", "rep"])
def test_wide_layout_parity(ds_c1, ds_rep, oracle, which):
    """The layout of texts of 2^32 rows or more (32-byte Occ lines, 64-bit
    intervals, no 12-mer table) on texts small enough for the oracle: the
    seeding, SA lookup and the rest of the path that a 3.1 Gbp index takes,
    against the oracle in the plain and batched forms."""
    ds = ds_rep if which == "rep" else ds_c1
    fq = ds.reads["main" if which != "c3" else "c3"]
    names, seq, off, _ = g3sa_amd.read_fastq(str(fq), limit=4000)
    got, _ = run_both(ds, oracle, seq, off, f"wide layout {which}", wide=True, variants=VARIANTS[:2])
    assert got.n_alns > 0


def test_wide_band_global_alignment_parity(built, oracle, tmp_path):
    """Reads of 250 bases carrying one 15-60 base insertion or deletion: their
    global alignments have bands of 2w+1 from about 40 up to 127 columns,
    computed by the one-wavefront DP (one and two columns per lane) and
    walked per thread; the CIGARs must equal the oracle's."""
    from types import SimpleNamespace
    rng = np.random.default_rng(23)
    L = 300_000
    g = rng.integers(0, 4, L).astype(np.uint8)
    fa = tmp_path / "wide.fa"
    fa.write_text(">chrW\n" + "\n".join("".join("ACGT"[b] for b in g[i:i + 60]) for i in range(0, L, 60)) + "\n")
    prefix = tmp_path / "wide"
    g3sa_amd.build_index(str(fa), str(prefix), 8)
    reads = []
    for k in range(600):
        gap = 15 + k % 46
        p = int(rng.integers(0, L - 400))
        cut = 90 + int(rng.integers(0, 70))
        if k % 2:  # deletion from the read
            r = np.concatenate([g[p:p + cut], g[p + cut + gap:p + 250 + gap]])
        else:      # insertion into the read
            r = np.concatenate([g[p:p + cut], rng.integers(0, 4, gap).astype(np.uint8), g[p + cut:p + 250 - gap]])
        r = r[:250].copy()
        if k % 4 >= 2:
            r = (3 - r[::-1]).astype(np.uint8)
        sub = rng.random(len(r)) < 0.005
        r[sub] = rng.integers(0, 4, int(sub.sum()))
        reads.append(r)
    seq, off = _batch(reads)
    run_both(SimpleNamespace(prefix=prefix), oracle, seq, off, "wide-bands")
    # the oracle's alignments include bands wider than 64 columns
    idx = oracle.load(prefix)
    try:
        ref = oracle.align(idx, seq, off, threads=oracle_threads())
    finally:
        oracle.lib.orc_index_free(idx)
    lq = ref["qe"] - ref["qb"]
    rl = ref["re"] - ref["rb"]
    assert int(np.abs(lq - rl).max()) >= 32, "no alignment with a band over 64 columns"


def _oracle_sam_records(ds, fq, tmp_path):
    """The oracle command line's SAM for fq: header and mapped records."""
    import subprocess
    from conftest import ROOT
    out = tmp_path / "oracle.sam"
    subprocess.run([str(ROOT / "oracle" / "_build" / "g3a_oracle"), "-t", str(oracle_threads()), "-m", "sam", "-o",
                    str(out), str(ds.prefix), str(fq)], check=True, capture_output=True)
    return out.read_bytes()


@pytest.mark.parametrize("which", ["rep", "c3"])
def test_abi_mapq_md_and_device_records(ds_rep, ds_c1, oracle, tmp_path, which):
    """MAPQ and MD at the C ABI (g3a_batch_out) against the oracle's SAM, and
    the device-formatted records (g3a_memcpy_text + g3a_format_records) byte
    for byte against the oracle's command line in both formats: paired reads
    on a repeat-rich text, and variable-length reads with divergent tails
    (whose short, low-scoring alignments give MAPQ below 60)."""
    ds_rep = ds_rep if which == "rep" else ds_c1
    fq = ds_rep.reads["main" if which == "rep" else "c3"]
    names, seq, off, quals = g3sa_amd.read_fastq(str(fq))
    sam = _oracle_sam_records(ds_rep, fq, tmp_path)
    recs = [ln.split(b"\t") for ln in sam.split(b"\n") if ln and not ln.startswith(b"@")]
    mapped = [r for r in recs if r[2] != b"*"]
    md_ref = [next(f[5:].decode() for f in r if f.startswith(b"MD:Z:")) for r in mapped]
    mapq_ref = np.array([int(r[4]) for r in mapped], dtype=np.int32)
    n = len(off) - 1
    name_off = np.zeros(n + 1, dtype=np.int64)
    name_off[1:] = np.cumsum([len(x) for x in names])
    qual_off = np.zeros(n + 1, dtype=np.int64)
    qual_off[1:] = np.cumsum([len(x) for x in quals])
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=n, max_bases=int(off[-1])) as al:
        print(f"[parity] ABI mapq/md: {n} reads", flush=True)
        got = al.align(seq, off)
        assert got.n_alns == len(mapped)
        np.testing.assert_array_equal(got.mapq, mapq_ref, err_msg="MAPQ differs from the oracle's")
        bad = [i for i in range(len(md_ref)) if got.md[i] != md_ref[i]]
        assert not bad, f"MD differs at alignments {bad[:5]}: {[(got.md[i], md_ref[i]) for i in bad[:3]]}"
        if which == "c3":
            assert len(set(mapq_ref.tolist())) > 3, "the oracle's MAPQ takes few values here"
        al.upload_text(b"".join(names), name_off, b"".join(quals), qual_off)
        for mode in ("sam", "compat"):
            print(f"[parity] device records: {mode}", flush=True)
            nb = al.format(sam=mode == "sam")
            text = al.text()
            assert len(text) == nb
            ref = sam if mode == "sam" else _oracle_compat(ds_rep, fq, tmp_path)
            body = b"".join(ln + b"\n" for ln in ref.split(b"\n") if ln and not ln.startswith(b"@"))
            assert text == body, f"{mode}: device records differ from the oracle's"


def _oracle_compat(ds, fq, tmp_path):
    import subprocess
    from conftest import ROOT
    out = tmp_path / "oracle.txt"
    subprocess.run([str(ROOT / "oracle" / "_build" / "g3a_oracle"), "-t", str(oracle_threads()), "-m", "compat",
                    "-o", str(out), str(ds.prefix), str(fq)], check=True, capture_output=True)
    return out.read_bytes()


@pytest.mark.parametrize("which", ["rep", "tandem"])
def test_chain_tiers_equal_serial(ds_rep, tmp_path, which):
    """Chaining by the tiers (a thread per read of at most 16 seeds, a
    wavefront per read of at most 128, the B-tree code on one lane above)
    equals the B-tree code for every read, field by field, in the learned and
    the FM-index forms, on repeat-rich references whose reads reach every
    tier."""
    if which == "rep":
        prefix = ds_rep.prefix
        names, seq, off, _ = g3sa_amd.read_fastq(str(ds_rep.reads["main"]), limit=20000)
    else:
        prefix, seq, off = _tandem_dataset(tmp_path)
    n = len(off) - 1
    with g3sa_amd.Aligner(str(prefix), max_reads=n, max_bases=int(off[-1])) as al:
        for learned in (0, 1):
            if learned:
                al.set_learned(1)
            al.set_chain_serial(False)
            tiers = al.align(seq, off)
            counts, kmsn = al.intervals()
            seeds = np.bincount(np.repeat(np.arange(n), counts), weights=np.minimum(kmsn[:, 1], 500),
                                minlength=n)
            print(f"[chain] {which} learned {learned}: {n} reads, seeds per read max {int(seeds.max())}, "
                  f"over 16: {int((seeds > 16).sum())}, over 64: {int((seeds > 64).sum())}", flush=True)
            al.set_chain_serial(True)
            serial = al.align(seq, off)
            al.set_chain_serial(False)
            np.testing.assert_array_equal(tiers.aln_off, serial.aln_off)
            for f in g3sa_amd.Result.FIELDS:
                np.testing.assert_array_equal(getattr(tiers, f), getattr(serial, f), err_msg=f"{which}: {f}")
            np.testing.assert_array_equal(tiers.cigar, serial.cigar)
            assert (seeds > 16).sum() > 0, "no read reached the wavefront tier"
            if which == "tandem":
                assert (seeds > 64).sum() > 0, "no read reached the wavefront tier's two-chunk form"
