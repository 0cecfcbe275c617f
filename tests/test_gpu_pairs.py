"""Paired-end primitives on the GPU (g3a_mem_pestat, g3a_mem_pair) against the
oracle's restatement of bwa/src/aux.cu:411-500, 703-714, 1000-1070, on the
pipeline's own alignment lists of a paired dataset.  Statistics must match
bit for bit (the doubles included); pair choices field by field."""
import ctypes as C

import numpy as np
import pytest

import g3sa_amd
import oracle_lib

pytestmark = pytest.mark.gpu


def _wrap32(x: int) -> int:
    return ((x + (1 << 31)) % (1 << 32)) - (1 << 31)


def _revcomp_pairs(seq, off, every):
    """Reverse-complement both ends of every `every`-th pair (FR pairs become RF)."""
    seq = seq.copy()
    n_pairs = (len(off) - 1) // 2
    for p in range(0, n_pairs, every):
        for r in (2 * p, 2 * p + 1):
            s = seq[off[r]:off[r + 1]]
            seq[off[r]:off[r + 1]] = np.where(s < 4, 3 - s, 4)[::-1]
    return seq


def _check(al, got, orc, l_pac, ctg, max_ins, id0, label):
    pes = al.mem_pestat(max_ins)
    ref = orc.pestat(l_pac, got.aln_off, got.rid, got.rb, got.qb, got.qe, got.score, max_ins)
    print(f"[pairs] {label}: pes {pes}", flush=True)
    assert pes == ref, label
    out = al.mem_pair(pes, id0)
    n_pairs = got.n_reads // 2
    assert out.shape == (n_pairs, 6)
    off = got.aln_off
    stats = {"paired": 0, "n_sub": 0, "proper": 0, "multi": 0}
    for p in range(n_pairs):
        b0, b1, b2 = int(off[2 * p]), int(off[2 * p + 1]), int(off[2 * p + 2])
        want = [0, 0, 0, -1, -1, 0]
        if b1 > b0 and b2 > b1:
            want[:5] = orc.mem_pair(l_pac, ctg, pes, _wrap32(id0 + p), got.rid[b0:b2], got.rb[b0:b2],
                                    got.score[b0:b2], b1 - b0)
            if got.rid[b0] == got.rid[b1] and got.rid[b0] >= 0:
                d, dist = orc.infer_dir(l_pac, int(got.rb[b0]), int(got.rb[b1]))
                if not pes[d]["failed"] and pes[d]["low"] <= dist <= pes[d]["high"]:
                    want[5] = 2
            stats["multi"] += (b2 - b0) > 2
        assert list(out[p]) == want, f"{label} pair {p}: {list(out[p])} != {want}"
        stats["paired"] += want[0] > 0
        stats["n_sub"] += want[2] > 0
        stats["proper"] += want[5] > 0
    print(f"[pairs] {label}: {n_pairs} pairs, {stats}", flush=True)
    return pes, stats


def test_pairs_parity(ds_rep, oracle):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_rep.reads["main"]))
    l_pac, ctg = oracle_lib.read_ann(ds_rep.prefix)
    n = len(off) - 1
    assert n % 2 == 0
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=n, max_bases=int(off[-1])) as al:
        got = al.align(seq, off)
        pes, st = _check(al, got, oracle, l_pac, ctg, 10000, 0, "rep FR")
        assert pes[1]["failed"] == 0 and [p["failed"] for i, p in enumerate(pes) if i != 1] == [1, 1, 1]
        assert 350 < pes[1]["avg"] < 450 and 30 < pes[1]["std"] < 70
        assert st["paired"] > 0.8 * (n // 2) and st["proper"] > 0.8 * (n // 2)
        # sizes cut at 350: the statistics of the truncated distribution
        _check(al, got, oracle, l_pac, ctg, 350, 0, "rep FR max_ins 350")
        # a third of the pairs turned into RF pairs: two live orientations;
        # pair ids past 2^31 wrap as the reference's int does
        got = al.align(_revcomp_pairs(seq, off, 3), off)
        pes, st = _check(al, got, oracle, l_pac, ctg, 10000, (1 << 31) - 1000, "rep FR+RF")
        assert pes[1]["failed"] == 0 and pes[2]["failed"] == 0
        # one pair in 40 RF: at least 10, but under 5% of FR, so RF fails
        got = al.align(_revcomp_pairs(seq, off, 40), off)
        pes, _ = _check(al, got, oracle, l_pac, ctg, 10000, 5, "rep FR + few RF")
        assert pes[2]["failed"] == 1 and pes[2]["std"] > 0
    # secondary hits kept (flag 0x8): ends with several hits, several
    # consistent pairs per pair of ends, and the unique-hit rule at work
    go = g3sa_amd.default_opt()
    go.flag |= 0x8
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=n, max_bases=int(off[-1]), opt=go) as al:
        got = al.align(seq, off)
        _, st = _check(al, got, oracle, l_pac, ctg, 10000, 77, "rep FR, secondary hits kept")
        assert st["multi"] > 0 and st["n_sub"] > 0


def test_pairs_error_paths(ds_rep):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_rep.reads["main"]))
    lib = g3sa_amd.load_library()
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=64, max_bases=int(off[64])) as al:
        pes = (g3sa_amd.PeStat * 4)()
        out = np.zeros((32, 6), dtype=np.int32)
        assert lib.g3a_mem_pestat(al.ctx, 10000, pes) == -4  # no batch yet
        al.align(seq[: off[63]], off[:64])  # 63 reads
        assert lib.g3a_mem_pestat(al.ctx, 10000, pes) == -1
        assert lib.g3a_mem_pair(al.ctx, pes, 0, out.ctypes.data) == -1
        al.align(seq[: off[64]], off[:65])
        assert lib.g3a_mem_pestat(al.ctx, -1, pes) == -1
        assert lib.g3a_mem_pestat(al.ctx, (1 << 20) + 1, pes) == -1
        assert lib.g3a_mem_pestat(al.ctx, 10000, pes) == 0
        assert lib.g3a_mem_pair(al.ctx, pes, 0, out.ctypes.data) == 0
        assert lib.g3a_mem_pair(al.ctx, pes, 0, C.c_void_p(0)) == -1
        # an empty batch holds no pairs
        al.align(seq[:0], off[:1])
        assert lib.g3a_mem_pestat(al.ctx, 10000, pes) == 0
        assert all(p.failed == 1 for p in pes)
        assert lib.g3a_mem_pair(al.ctx, pes, 0, out.ctypes.data) == 0
