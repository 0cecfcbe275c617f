"""GPU parity of the learned-index form: the device-built full SA, keys and
two-stage model, the text comparison of unique matches and the model's
intervals in the last seeding round must give the same seeding intervals
(k, s, m, n) as the FM-index-only path, and alignments identical to the
oracle, on uniform, repeat-rich and tandem-repeat references."""
from __future__ import annotations

import numpy as np
import pytest

import g3sa_amd
from conftest import oracle_threads
from test_gpu_parity import _batch, compare

pytestmark = pytest.mark.gpu


def _intervals_equal(a, b, label):
    ca, ia = a
    cb, ib = b
    np.testing.assert_array_equal(ca, cb, err_msg=f"{label}: intervals per read differ")
    if not np.array_equal(ia, ib):
        bad = np.nonzero((ia != ib).any(axis=1))[0][:5]
        raise AssertionError(f"{label}: intervals differ at {bad.tolist()}: fm {ia[bad].tolist()} learned {ib[bad].tolist()}")


def _run(ds, oracle, seq, off, label, forms=(("batched", 1, (1024, 1)), ("batched", 0, (128, 10)), ("plain", 1, (256, 0))), modes=(1,)):
    idx = oracle.load(ds.prefix)
    try:
        ref = oracle.align(idx, seq, off, threads=oracle_threads())
    finally:
        oracle.lib.orc_index_free(idx)
    with g3sa_amd.Aligner(str(ds.prefix), max_reads=max(len(off) - 1, 1), max_bases=max(int(off[-1]), 1)) as al:
        info = None
        for form, ordered, blocks in forms:
            al.set_form(form)
            al.set_seed_order(ordered)
            al.set_seed_grid(*blocks)
            al.set_learned(False)
            print(f"[learned] {label}: {len(off) - 1} reads, FM-only, form {form} order {ordered} blocks {blocks}",
                  flush=True)
            got_fm = al.align(seq, off)
            iv_fm = al.intervals()
            compare(got_fm, ref, f"{label} FM-only [{form}]")
            for mode in modes:
                info = al.set_learned(mode)
                print(f"[learned] {label}: learned form (mode {mode}); model {info}", flush=True)
                got = al.align(seq, off)
                iv = al.intervals()
                st = al.learned_stats()
                print(f"[learned] {label}: work {st}", flush=True)
                compare(got, ref, f"{label} learned mode {mode} [{form}]")
                _intervals_equal(iv_fm, iv, f"{label} mode {mode} [{form}]")
                assert got.work["n_seeds"] == ref["work"]["sa_reads"], label
            al.set_learned(1)
            st = al.learned_stats() if modes[-1] == 1 else st
        return info, st


def test_learned_config1(ds_c1, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["main"]))
    info, st = _run(ds_c1, oracle, seq, off, "config1", modes=(2, 1))
    assert info["leaves"] == 4096 and info["rows"] == 2 * 1_000_000 + 1
    assert info["distinct_keys"] > 0.99 * info["rows"]
    assert info["max_err"] < 256 and info["keys_over_window"] == 0
    assert info["model_bytes"] < 2 << 20
    assert st["lookups"] > 0 and st["text_runs"] > 0 and st["fm_answered"] == 0


def test_learned_config3_style(ds_c1, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["c3"]))
    _run(ds_c1, oracle, seq, off, "config3-style", forms=(("batched", 1, (1024, 1)),))


def test_learned_repeats(ds_rep, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_rep.reads["main"]))
    info, st = _run(ds_rep, oracle, seq, off, "repeats", forms=(("batched", 1, (1024, 1)),))
    assert info["distinct_keys"] < info["rows"]  # repeated 32-mers share keys


def test_learned_edge_cases(ds_c1, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["main"]), limit=40)
    reads = [seq[off[i]:off[i + 1]] for i in range(40)]
    rng = np.random.default_rng(11)
    cases = [
        np.zeros(0, np.uint8),
        np.full(150, 4, np.uint8),
        reads[0][:19],
        reads[1][:20],                                        # one last-round window exactly
        np.concatenate([reads[2][:40], [4], reads[2][41:]]),  # N inside a unique match
        np.concatenate([reads[6], reads[7][:170]]),           # 320 bp
        rng.integers(0, 4, 150).astype(np.uint8),
        np.tile(np.array([0, 1], np.uint8), 75),
        np.full(150, 0, np.uint8),                            # poly-A: long runs of equal keys
    ] + reads[9:40]
    seq2, off2 = _batch(cases)
    _run(ds_c1, oracle, seq2, off2, "edge-cases", forms=(("batched", 1, (1024, 1)),))


@pytest.mark.parametrize("min_seed_len", [9, 11, 15, 26, 31])
def test_learned_seed_lengths(ds_rep, oracle, min_seed_len):
    """The last round's window is min_seed_len + 1 bases: below 12 the model's
    window alone is searched (a miss goes to the FM-index), from 12 on inside
    the 12-mer bucket, and above 32 the round runs on the FM-index; seeding
    intervals and alignments stay identical."""
    import ctypes as C
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_rep.reads["main"]), limit=3000)
    oo = type(oracle.opt)()
    C.pointer(oo)[0] = oracle.opt
    oo.min_seed_len = min_seed_len
    go = g3sa_amd.default_opt()
    go.min_seed_len = min_seed_len
    idx = oracle.load(ds_rep.prefix)
    try:
        ref = oracle.align(idx, seq, off, threads=oracle_threads(), opt=oo)
    finally:
        oracle.lib.orc_index_free(idx)
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=len(off) - 1, max_bases=int(off[-1]), opt=go) as al:
        print(f"[learned] min_seed_len {min_seed_len}: FM-only", flush=True)
        got_fm = al.align(seq, off)
        iv_fm = al.intervals()
        compare(got_fm, ref, f"min_seed_len {min_seed_len} FM-only")
        al.set_learned(True)
        print(f"[learned] min_seed_len {min_seed_len}: learned", flush=True)
        got = al.align(seq, off)
        iv = al.intervals()
        st = al.learned_stats()
        print(f"[learned] min_seed_len {min_seed_len}: work {st}", flush=True)
        compare(got, ref, f"min_seed_len {min_seed_len} learned")
        _intervals_equal(iv_fm, iv, f"min_seed_len {min_seed_len}")
        assert (st["lookups"] > 0) == (min_seed_len + 1 <= 32)


@pytest.mark.parametrize("which", ["c1", "rep"])
def test_seeding_intervals_against_oracle(ds_c1, ds_rep, oracle, which):
    """Stage-level parity of seeding: the GPU's intervals (k, s, m, n) of
    every read, in each read's order, against the oracle's seeding rounds
    (seed.cu restated), for the FM-index-only and the learned form."""
    ds = ds_c1 if which == "c1" else ds_rep
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds.reads["main"]), limit=3000)
    n = len(off) - 1
    idx = oracle.load(ds.prefix)
    try:
        ref = [oracle.seed_intervals(idx, seq[off[i]:off[i + 1]]) for i in range(n)]
    finally:
        oracle.lib.orc_index_free(idx)
    counts_ref = np.array([len(r) for r in ref], dtype=np.int32)
    flat_ref = np.array([(k, s, m, e) for r in ref for (k, _l, s, m, e) in r], dtype=np.int64).reshape(-1, 4)
    with g3sa_amd.Aligner(str(ds.prefix), max_reads=n, max_bases=int(off[-1])) as al:
        for mode in (0, 1):
            al.set_learned(mode)
            print(f"[seeding] {which}: {n} reads, learned {mode}", flush=True)
            al.align(seq[:off[n]], off)
            counts, kmsn = al.intervals()
            _intervals_equal((counts_ref, flat_ref), (counts, kmsn), f"{which} oracle vs GPU (learned {mode})")
