"""Pins for the oracle functions that the brute-force tests leave open.

Each check below compares the oracle with something it was not written from:

* Global alignment (bwa/src/ksw.cu:706-812).  A full-matrix restatement of
  ksw_global2's recurrences over explicit H, E, F and M tables (gaps open from
  M only, band [i-w, i+w+1), first row defined up to column w), with the
  traceback re-derived as a state machine over those tables, and small cases
  whose CIGARs are worked out by hand (the M >= E >= F tie order).
* Chaining with equal chain positions (bwa/src/kbtree_CUDA.cu:22-144,
  bwa/src/aux.cu:1755-1825).  Seed sets whose chain order and seed ownership
  are worked out by hand from klib's B-tree rules: a duplicate key is
  inserted after the first equal key of its node, and the interval query
  returns the first equal key it meets on the way down.
* Round-2 reseeding (bwa/src/seed.cu:639-913).  Planted texts where the
  round's output is fixed by hand: the pivot > 77 skip (seed.cu:718), the
  `collected` flag (one seed per backward step, seed.cu:836-859) and the
  append after the loop (seed.cu:870-873).
* Extension quirk rows (bwa/src/ksw.cu:924-1003).  A case where the only
  cells above zero after row 31 descend from the row-32 cells that read the
  untouched first-row buffer, with the result worked out by hand.
* Record and SAM emitters (bwa/src/pipeline.cpp:287-351 and standard SAM).
  Hand-written expected lines for unmapped, forward, reverse, clipped and
  mismatching reads.

The checks take the oracle as an argument so that tests/test_oracle_mutants.py
can run them against deliberately broken builds of the oracle.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

import g3sa_amd
from oracle_lib import OrcSeed

NEG = -0x40000000  # MINUS_INF, bwa/src/ksw.cu:1040
A, C_, G, T = 0, 1, 2, 3


def enc(s: str) -> np.ndarray:
    return np.array(["ACGT".index(ch) for ch in s], dtype=np.uint8)


def cigar_str(cig) -> str:
    return "".join(f"{int(c) >> 4}{'MIDSH'[int(c) & 0xF]}" for c in cig)


# --------------------------------------------------------------------------
# global alignment
# --------------------------------------------------------------------------
def global_tables(q, t, mat, od, ed, oi, ei, w):
    """ksw_global2 as tables.  Rows are target positions i, columns query
    positions j.  Cell (i, j) is in the band iff max(i-w, 0) <= j < min(i+w+1, Q).
      H(-1,-1) = 0; H(-1, j) = -(oi + ei*(j+1)) for j+1 <= w, else -inf;
      H(i, -1) = -(od + ed*(i+1)) for i <= w, else -inf;
      M(i,j) = H(i-1,j-1) + S(t_i, q_j);
      E(i,j) = max(M(i-1,j) - od - ed, E(i-1,j) - ed) when (i-1, j) is in the band, else -inf;
      F(i,j) = max(M(i,j-1) - oi - ei, F(i,j-1) - ei) when (i, j-1) is in the band, else -inf;
      H(i,j) = max(M, E, F).
    Gaps open from M, not from H."""
    Q, Tn = len(q), len(t)
    oed, oei = od + ed, oi + ei

    def band(i, j):
        return 0 <= i < Tn and max(i - w, 0) <= j < min(i + w + 1, Q)

    H, E, F, M = {}, {}, {}, {}

    def h(i, j):
        if i == -1 and j == -1:
            return 0
        if i == -1:
            return -(oi + ei * (j + 1)) if j + 1 <= w else NEG
        if j == -1:
            return -(od + ed * (i + 1)) if i <= w else NEG
        return H[i, j]

    for i in range(Tn):
        for j in range(max(i - w, 0), min(i + w + 1, Q)):
            m = h(i - 1, j - 1) + int(mat[t[i] * 5 + q[j]])
            e = max(M[i - 1, j] - oed, E[i - 1, j] - ed) if band(i - 1, j) else NEG
            f = max(M[i, j - 1] - oei, F[i, j - 1] - ei) if band(i, j - 1) else NEG
            M[i, j], E[i, j], F[i, j] = m, e, f
            H[i, j] = max(m, e, f)
    return H, E, F, M


def global_reference(q, t, mat, od, ed, oi, ei, w):
    """Score and CIGAR of ksw_global2 from the tables: the traceback starts
    at the last row's last band column and explains each value by the move
    that produced it, preferring M over E on a tie and M or E over F; an E
    (F) run continues while extending beats opening from M strictly."""
    Q, Tn = len(q), len(t)
    H, E, F, M = global_tables(q, t, mat, od, ed, oi, ei, w)
    oed, oei = od + ed, oi + ei
    score = H[Tn - 1, Q - 1] if Tn - 1 + w + 1 >= Q else NEG
    ops = []
    i, k = Tn - 1, min(Tn + w, Q) - 1
    state = "H"
    while i >= 0 and k >= 0:
        if state == "H":
            m, e, f = M[i, k], E[i, k], F[i, k]
            state = "M" if m >= e else "E"
            if max(m, e) < f:
                state = "F"
        if state == "M":
            ops.append("M")
            i, k, state = i - 1, k - 1, "H"
        elif state == "E":
            ops.append("D")
            # E(i,k) extended E(i-1,k) iff that beats opening from M(i-1,k) strictly
            ext = i - 1 >= 0 and E[i - 1, k] - ed > M[i - 1, k] - oed
            i -= 1
            state = "E" if ext else "M"
        else:
            ops.append("I")
            ext = k - 1 >= 0 and F[i, k - 1] - ei > M[i, k - 1] - oei
            k -= 1
            state = "F" if ext else "M"
    ops += ["D"] * (i + 1) + ["I"] * (k + 1)
    ops.reverse()
    runs = []
    for op in ops:
        if runs and runs[-1][0] == op:
            runs[-1][1] += 1
        else:
            runs.append([op, 1])
    return score, "".join(f"{n}{op}" for op, n in runs)


# Hand-worked cases (default scoring: match 1, mismatch -4, gaps 6+1k).
#   q=A,  t=AA: row 1 has M(1,0) = H(0,-1)+1 = -7+1 = -6 and E(1,0) = M(0,0)-7 = -6;
#               the tie goes to M, so the last target base is matched: 1D1M.
#   q=AA, t=A : M(0,1) = H(-1,0)+1 = -6 ties F(0,1) = M(0,0)-7 = -6; M wins: 1I1M.
#   q=AC, t=C : M(0,1) = -7+1 = -6 beats F(0,1) = (1-4)-7 = -10: 1I1M, score -6.
#   q=ACGT, t=ACGT, w=0: the diagonal only, 4M, score 4.
#   q=A,  t=C : one mismatch, 1M, score -4.
HAND_GLOBAL = [
    ("A", "AA", 100, -6, "1D1M"),
    ("AA", "A", 100, -6, "1I1M"),
    ("AC", "C", 100, -6, "1I1M"),
    ("ACGT", "ACGT", 0, 4, "4M"),
    ("A", "C", 100, -4, "1M"),
]


def check_global(orc) -> None:
    mat = orc.mat()
    for qs, ts, w, score, cig in HAND_GLOBAL:
        s, c = orc.global2(enc(qs), enc(ts), w=w)
        assert (s, cigar_str(c)) == (score, cig), (qs, ts, w, s, cigar_str(c))
    rng = np.random.default_rng(77)
    for it in range(400):
        Q = int(rng.integers(1, 40))
        q = rng.integers(0, 4, Q).astype(np.uint8)
        t = q.copy()
        for _ in range(int(rng.integers(0, 6))):  # edits: substitutions, insertions, deletions
            p = int(rng.integers(0, len(t) + 1))
            r = rng.random()
            if r < 0.4 and p < len(t):
                t[p] = rng.integers(0, 4)
            elif r < 0.7:
                t = np.insert(t, p, rng.integers(0, 4, int(rng.integers(1, 4))))
            elif len(t) > 1:
                t = np.delete(t, slice(p, p + int(rng.integers(1, 3))))
        if len(t) == 0:
            continue
        t = t.astype(np.uint8)
        # the band always reaches the end cell (the reference's callers
        # guarantee w >= |tlen - qlen| + 3)
        w = abs(len(t) - Q) + int(rng.choice([0, 1, 2, 3, 5, 20]))
        gaps = (6, 1, 6, 1) if rng.random() < 0.6 else tuple(int(x) for x in rng.integers(1, 7, 4))
        s, c = orc.global2(q, t, w=w, gaps=gaps)
        want = global_reference(q, t, mat, *gaps, w)
        assert (s, cigar_str(c)) == want, (it, q.tolist(), t.tolist(), w, gaps, s, cigar_str(c), want)


def test_global_alignment_pinned(oracle):
    check_global(oracle)


def test_global_reference_hand_cases():
    """The table restatement itself reproduces the hand-worked cases."""
    import oracle_lib
    mat = np.full(25, -1, dtype=np.int8)
    for i in range(4):
        for j in range(4):
            mat[i * 5 + j] = 1 if i == j else -4
    for qs, ts, w, score, cig in HAND_GLOBAL:
        assert global_reference(enc(qs), enc(ts), mat, 6, 1, 6, 1, w) == (score, cig)
    assert oracle_lib  # imported for the same scoring convention


# --------------------------------------------------------------------------
# chaining with equal chain positions
# --------------------------------------------------------------------------
def chain_dump(orc, idx, seeds):
    n = len(seeds)
    arr = (OrcSeed * n)()
    for i, (qbeg, ln, rbeg) in enumerate(seeds):
        arr[i].rbeg, arr[i].qbeg, arr[i].len, arr[i].rid = rbeg, qbeg, ln, 0
    owner = np.zeros(n, np.int32)
    w = np.zeros(n, np.int32)
    kept = np.zeros(n, np.int32)
    nk = C.c_int()
    nc = orc.lib.orc_chain_dump(idx, C.byref(orc.opt), arr, n, owner.ctypes.data, w.ctypes.data, kept.ctypes.data,
                                C.byref(nk))
    return nc, owner.tolist()


# Seeds (qbeg, len, rbeg) in sort order; every seed is on contig 0, forward strand.
#
# Case 1, one leaf.  s1, s2, s3 start chains A, B, C at the same position 1000
# (their query offsets differ by more than w = 100).  Inserting B after A's
# equal key gives [A, B]; the query for s3 meets A first; C goes right after
# A: [A, C, B].  s4 at 1010: the last key <= 1010 in the leaf is B (s2's
# chain), whose query offset is 210 away, so s4 starts chain D although it is
# colinear with C.  In-order: A C B D.
CHAIN_CASE_1 = ([(0, 20, 1000), (200, 20, 1000), (400, 20, 1000), (410, 20, 1010)], 4, [0, 2, 1, 3])
# Case 2, across a split.  Eleven chains c0..c10 at 100, 200, ..., 1100 fill
# the root leaf (2t-1 = 11 keys, t = 6).  s11 at 600 (c5's position) does
# not merge with c5 and starts c11: the root splits into [c0..c4] | c5 |
# [c6..c10], and c11 descends right of c5 (equal keys go to the child after
# the equal key) to the front of the right leaf.  s12 at 610: the root gives
# c5 as a candidate, the right leaf's last key <= 610 is c11, which s12
# extends colinearly.  In-order: c0..c5, c11, c6..c10.
_ELEVEN = [(250 * i, 20, 100 * (i + 1)) for i in range(11)]
CHAIN_CASE_2 = (_ELEVEN + [(2750, 20, 600), (2760, 20, 610)], 12, [0, 1, 2, 3, 4, 5, 7, 8, 9, 10, 11, 6, 6])
# Case 3, an equal key in an internal node.  As case 2 up to c11; s12 at 600
# again: the query stops at the root's c5 (equal key, returned at once) and
# does not descend to c11, with which s12 would be colinear (50 apart on the
# read, 0 on the reference).  So s12 starts c12, inserted after c11 (the
# first equal key of the right leaf).  In-order: c0..c5, c11, c12, c6..c10.
CHAIN_CASE_3 = (_ELEVEN + [(2750, 20, 600), (2800, 20, 600)], 13, [0, 1, 2, 3, 4, 5, 8, 9, 10, 11, 12, 6, 7])


def check_chain(orc, idx) -> None:
    for seeds, nc, owner in (CHAIN_CASE_1, CHAIN_CASE_2, CHAIN_CASE_3):
        got = chain_dump(orc, idx, seeds)
        assert got == (nc, owner), (seeds[-2:], got)


@pytest.fixture(scope="module")
def chain_idx(built, oracle, tmp_path_factory):
    tmp = tmp_path_factory.mktemp("pin_chain")
    rng = np.random.default_rng(3)
    fa = tmp / "c.fa"
    fa.write_text(">c0\n" + "".join("ACGT"[v] for v in rng.integers(0, 4, 6000)) + "\n")
    g3sa_amd.build_index(str(fa), str(tmp / "c"), 4)
    return tmp / "c"


def test_chaining_equal_positions_pinned(chain_idx, oracle):
    idx = oracle.load(chain_idx)
    try:
        check_chain(oracle, idx)
    finally:
        oracle.lib.orc_index_free(idx)


# --------------------------------------------------------------------------
# round-2 reseeding
# --------------------------------------------------------------------------
def brute_count(text: bytes, pat: np.ndarray) -> int:
    p = bytes((np.asarray(pat) + 1).tolist())
    n, k, i = 0, -1, 0
    while True:
        k = text.find(p, k + 1)
        if k < 0:
            return n
        n += 1


class PlantedReseed:
    """Reference: 6000 random bases with a 60-mer X planted once,
    Y = X[20:45] (25 bases) planted three more times and Y2 = X[20:41]
    (21 bases) twice more, each copy's flanks differing from X's.  Reads
    (2-bit codes):

      read_a = 10 random + X + 10 random.  Round 1: the one SMEM read[10..69]
        (X, 1 occurrence).  Round 2: pivot (10+69+1)>>1 = 40 <= 77.  Forward
        LEPs from read[40] = X[30] grow until X[30..45] leaves Y; backward,
        every LEP survives down to m = 30 (read[30..] = Y), where extending by
        X[19] leaves one occurrence: all LEPs fail in that step and, with the
        `collected` flag, only the first (longest) one, read[30..54] = Y
        (4 occurrences), is collected.  No LEP is left, so nothing is
        appended after the loop.  Round 2 = [Y at 30..54].
      read_b = X[20:60] + 10 random: the SMEM read[0..39] (X[20..59]); pivot
        (0+39+1)>>1 = 20.  Backward extension stays inside Y down to read[0],
        the loop ends at m = 0 with LEPs left, and the longest, read[0..24]
        = Y, is appended after the loop.  Round 2 = [Y at 0..24].
      read_c = 60 random + X + 10 random: the SMEM read[60..119], pivot 90 >
        77, so round 2 is skipped: round 2 = [].
    The random flanks are chosen so that no flank base extends a match."""

    def __init__(self, tmp: Path):
        rng = np.random.default_rng(41)
        L = 6000
        ref = rng.integers(0, 4, L).astype(np.uint8)
        x = rng.integers(0, 4, 60).astype(np.uint8)
        y = x[20:45]
        ref[1000:1060] = x
        for p in (2500, 3700, 5000):
            ref[p:p + 25] = y
            ref[p - 1] = (x[19] + 1) % 4   # Y copies cannot extend left into X[19]
            ref[p + 25] = (x[45] + 1) % 4  # or right into X[45]
        for p in (2000, 4400):
            ref[p:p + 21] = x[20:41]
            ref[p - 1] = (x[19] + 1) % 4
            ref[p + 21] = (x[41] + 1) % 4
        self.ref, self.x, self.y = ref, x, y
        fa = tmp / "rs.fa"
        fa.write_text(">rs\n" + "".join("ACGT"[v] for v in ref) + "\n")
        self.prefix = tmp / "rs"
        g3sa_amd.build_index(str(fa), str(self.prefix), 4)
        text = np.concatenate([ref, 3 - ref[::-1]])
        self.text = bytes((text + 1).tolist())

        def flank(n, before=None, after=None):
            f = rng.integers(0, 4, n).astype(np.uint8)
            if after is not None:   # base right before X must differ from ref[999]
                f[-1] = (after + 1) % 4
            if before is not None:  # base right after X must differ from ref[1060]
                f[0] = (before + 1) % 4
            return f

        self.read_a = np.concatenate([flank(10, after=ref[999]), x, flank(10, before=ref[1060])])
        self.read_b = np.concatenate([x[20:], flank(10, before=ref[1060])])
        self.read_c = np.concatenate([flank(60, after=ref[999]), x, flank(10, before=ref[1060])])

    def preconditions(self) -> None:
        cnt = lambda s: brute_count(self.text, s)  # noqa: E731
        assert cnt(self.x) == 1 and cnt(self.y) == 4 and cnt(self.x[20:41]) == 6
        assert cnt(self.x[19:45]) == 1 and cnt(self.x[20:46]) == 1
        assert cnt(self.x[19:41]) == 1 and cnt(self.x[20:42]) == 4 and cnt(self.x[30:41]) == 6
        for r, m, n in ((self.read_a, 10, 69), (self.read_b, 0, 39), (self.read_c, 60, 119)):
            assert cnt(r[m:n + 1]) == 1
            if m > 0:
                assert cnt(r[m - 1:n + 1]) == 0
            if n + 1 < len(r):
                assert cnt(r[m:n + 2]) == 0

    def round2(self, orc, idx, read, smem):
        """Round-2 part of the oracle's intervals: round 1 is the hand-known
        SMEM, round 3 the trailing 20-base forward scan, so round 2 sits
        between them."""
        got = [(s, m, n) for (_k, _l, s, m, n) in orc.seed_intervals(idx, read)]
        assert got[0] == smem, got
        r3 = round3_intervals(self.text, read)
        if r3:
            assert got[len(got) - len(r3):] == r3, (got, r3)
        return got[1:len(got) - len(r3)]


def round3_intervals(text: bytes, q: np.ndarray, min_len=20, max_intv=20):
    """reseedLastRound (bwa/src/seed.cu:1518-1550) by forward scan with
    brute-force counts."""
    out, m, L = [], 0, len(q)
    while m < L:
        nxt = m + 1
        if q[m] < 4:
            for n in range(m + 1, L):
                nxt = n + 1
                c = brute_count(text, q[m:n + 1])
                if c < max_intv and n - m + 1 >= min_len:
                    if c > 0:
                        out.append((c, m, n))
                    break
        m = nxt
    return out


def check_reseed(orc, pr: PlantedReseed) -> None:
    idx = orc.load(pr.prefix)
    try:
        assert pr.round2(orc, idx, pr.read_a, (1, 10, 69)) == [(4, 30, 54)]
        assert pr.round2(orc, idx, pr.read_b, (1, 0, 39)) == [(4, 0, 24)]
        assert pr.round2(orc, idx, pr.read_c, (1, 60, 119)) == []
    finally:
        orc.lib.orc_index_free(idx)


@pytest.fixture(scope="module")
def planted_reseed(built, tmp_path_factory):
    pr = PlantedReseed(tmp_path_factory.mktemp("pin_reseed"))
    return pr


def test_reseed_planted_preconditions(planted_reseed):
    planted_reseed.preconditions()


def test_reseed_hand_cases(planted_reseed, oracle):
    check_reseed(oracle, planted_reseed)


# --------------------------------------------------------------------------
# extension quirk rows
# --------------------------------------------------------------------------
# Query: 100 x A.  Target: 32 x C, then 68 x A.  h0 = 60, w = 0, default
# scoring (match 1, mismatch -4, gaps 6+1k), end_bonus 5 (w stays 0).
#
# Rows 0..31 mismatch everywhere.  With w = 0 a cell (i, j) needs j >= i, so
# no deletion run leaves the diagonal band; the best value there,
# H(0,0) = 60-4 = 56, falls by 4 per row along the diagonal and the first-row
# injections at columns 31..33 (row 0, outside the corner) die within three
# rows: every cell of rows 14..31 is 0.
#
# Row 32 (target A, a match) fills columns 32..65.  Columns 32 and 33 read
# zeros.  Columns j >= 34 lie beyond i+w+1 = 33, where the tile sweep's
# lane 0 reads the first-row buffer that no row above overwrote: the diagonal
# input is max(h0 - 7 - (j-1), 0) = 54 - j, so H(32,34) = 21, H(32,35) = 20,
# ... (F from the left is 7 lower, E input 0).
#
# From there the DP is ordinary: diagonal d = j - i = 2 grows by 1 per row,
# H(32+k, 34+k) = 21+k, up to (63, 65) = 52 at the tile's column limit 65.
# Diagonal 1 takes E from diagonal 2 (open 7): H(33,34) = 14 and then
# H(32+k, 33+k) = 13+k, i.e. H(63,64) = 44.  Row 64 is again a quirk row:
# (64, 66) reads the first-row value 0, so diagonal 2 is cut; (64, 65) takes
# max(H(63,64)+1, E from (63,65) = 53-7-... ) = 45.  Diagonal 1 then runs
# freely: H(64+k, 65+k) = 45+k to its end (98, 99) = 79 > h0.  Row 96 is a
# quirk row, but diagonal 1's column 97 there is i+w+1 itself, a normal cell.
#
# Result: score 79 at target row 98, query column 99: qle = 100, tle = 99.
# Without the quirk (zeros beyond row 31) nothing exceeds h0 and the result
# would be (60, 0, 0).
QUIRK_CASE = dict(q="A" * 100, t="C" * 32 + "A" * 68, h0=60, w=0, want=(79, 100, 99))


def check_quirk(orc) -> None:
    c = QUIRK_CASE
    got = orc.extend(enc(c["q"]), enc(c["t"]), c["h0"], w=c["w"])
    assert got == c["want"], got


def test_extension_quirk_row_hand_case(oracle):
    check_quirk(oracle)


def test_extension_quirk_row_emulation_agrees(oracle):
    c = QUIRK_CASE
    assert oracle.extend(enc(c["q"]), enc(c["t"]), c["h0"], w=c["w"], emulate=True) == c["want"]


# --------------------------------------------------------------------------
# emitters
# --------------------------------------------------------------------------
def revcomp(s: str) -> str:
    return s[::-1].translate(str.maketrans("ACGT", "TGCA"))


class EmitterCase:
    """Two contigs (chrA 1000 bp, chrB 800 bp, random) and five reads whose
    records are written out by hand below."""

    def __init__(self, tmp: Path):
        rng = np.random.default_rng(99)
        a = "".join("ACGT"[v] for v in rng.integers(0, 4, 1000))
        b = "".join("ACGT"[v] for v in rng.integers(0, 4, 800))
        self.fa = tmp / "em.fa"
        self.fa.write_text(f">chrA first contig\n{a}\n>chrB\n{b}\n")
        self.prefix = tmp / "em"
        g3sa_amd.build_index(str(self.fa), str(self.prefix), 4)

        def other(base):  # a base different from `base`
            return "ACGT"[("ACGT".index(base) + 1) % 4]

        r0 = "".join("ACGT"[v] for v in np.random.default_rng(5).integers(0, 4, 50))  # not in the reference
        r1 = b[100:180]                                                # forward, exact
        r2 = revcomp(a[200:290])                                       # reverse, exact
        tail = "".join(other(b[360 + k]) for k in range(20))
        r3 = b[300:360] + tail                                         # forward, 3' tail clipped
        junk = "".join(other(c) for c in revcomp(a[660:675]))
        r4 = junk + revcomp(a[600:660])                                # reverse, read's 5' junk clipped
        r5 = a[500:520] + other(a[520]) + other(a[521]) + a[522:540]   # two mismatches
        self.reads = [("r0", r0), ("r1", r1), ("r2", r2), ("r3", r3), ("r4", r4), ("r5", r5)]
        self.quals = {n: "".join(chr(33 + 2 + (i * 7) % 38) for i in range(len(s))) for n, s in self.reads}
        self.fq = tmp / "em.fq"
        self.fq.write_text("".join(f"@{n} comment {k}\n{s}\n+\n{self.quals[n]}\n" for k, (n, s) in enumerate(self.reads)))
        q = self.quals
        # The reference's record lines: NAME RNAME POS0 CIGAR SEQ QUAL, the
        # read as given; POS0 is the 0-based leftmost contig position; clips
        # follow the genome orientation (clip5 = len-qe for reverse hits).
        self.compat = "".join([
            f"r0 * 0 * {r0} {q['r0']}\n",
            f"r1 chrB 100 80M {r1} {q['r1']}\n",
            f"r2 chrA 200 90M {r2} {q['r2']}\n",
            f"r3 chrB 300 60M20S {r3} {q['r3']}\n",
            f"r4 chrA 600 60M15S {r4} {q['r4']}\n",
            f"r5 chrA 500 40M {r5} {q['r5']}\n",
        ])
        # Standard SAM: 1-based POS, SEQ/QUAL reverse-complemented/reversed
        # for flag 16.  MAPQ (aux.cu:676-700 with sub = 19): for l >= 50,
        # 6.02*(score-19)*(3/ln l)^2 is far above 60, so 60; for r5,
        # l = 40 < 50, identity = 1 - (40-30)/5/40 = 0.95, and
        # int(6.02 * 11 * 0.95^4 + .499) = int(53.94 + .499) = 54.
        md5 = f"20{a[520]}0{a[521]}18"
        self.sam = "".join([
            "@HD\tVN:1.6\tSO:unsorted\n@SQ\tSN:chrA\tLN:1000\n@SQ\tSN:chrB\tLN:800\n@PG\tID:g3a\tPN:g3a\n",
            f"r0\t4\t*\t0\t0\t*\t*\t0\t0\t{r0}\t{q['r0']}\n",
            f"r1\t0\tchrB\t101\t60\t80M\t*\t0\t0\t{r1}\t{q['r1']}\tNM:i:0\tMD:Z:80\tAS:i:80\n",
            f"r2\t16\tchrA\t201\t60\t90M\t*\t0\t0\t{a[200:290]}\t{q['r2'][::-1]}\tNM:i:0\tMD:Z:90\tAS:i:90\n",
            f"r3\t0\tchrB\t301\t60\t60M20S\t*\t0\t0\t{r3}\t{q['r3']}\tNM:i:0\tMD:Z:60\tAS:i:60\n",
            f"r4\t16\tchrA\t601\t60\t60M15S\t*\t0\t0\t{revcomp(r4)}\t{q['r4'][::-1]}\tNM:i:0\tMD:Z:60\tAS:i:60\n",
            f"r5\t0\tchrA\t501\t54\t40M\t*\t0\t0\t{r5}\t{q['r5']}\tNM:i:2\tMD:Z:{md5}\tAS:i:30\n",
        ])


def run_oracle_cli(cli: Path, ec: EmitterCase, mode: str, out: Path) -> str:
    subprocess.run([str(cli), "-t", "1", "-m", mode, "-o", str(out), str(ec.prefix), str(ec.fq)], check=True,
                   capture_output=True)
    return out.read_text()


def check_emitters(cli: Path, ec: EmitterCase, tmp: Path) -> None:
    assert run_oracle_cli(cli, ec, "compat", tmp / "o.txt") == ec.compat
    assert run_oracle_cli(cli, ec, "sam", tmp / "o.sam") == ec.sam


@pytest.fixture(scope="module")
def emitter_case(built, tmp_path_factory):
    return EmitterCase(tmp_path_factory.mktemp("pin_emit"))


def test_emitters_hand_lines(emitter_case, tmp_path):
    import oracle_lib
    check_emitters(oracle_lib.CLI, emitter_case, tmp_path)
