"""The oracle's seeding against brute force on a small text.

Round 1 (SMEMs of length >= 19 in end order) and round 3 (forward scan for
matches of length >= 20 with fewer than 20 occurrences) are recomputed with
naive substring counting and suffix-array ranges, so a wrong Occ, count,
sentinel or bidirectional update shows up.  Round 2 is recomputed with the
reference's control flow over brute-force interval sizes and ranges.
"""
from __future__ import annotations

import bisect

import numpy as np
import pytest

import g3sa_amd
from test_index import small_reference, true_sa

MIN_SEED = 19


class Brute:
    def __init__(self, text: np.ndarray, sa: np.ndarray):
        self.t = bytes((text + 1).tolist()) + b"\x00"
        self.sa = sa
        self.keys = [self.t[i:] for i in sa]

    def interval(self, pat: np.ndarray):
        p = bytes((np.asarray(pat) + 1).tolist())
        lo = bisect.bisect_left(self.keys, p)
        hi = bisect.bisect_left(self.keys, p + b"\x06")
        return lo, hi - lo

    def count(self, pat):
        return self.interval(pat)[1]


def round1(b: Brute, q: np.ndarray):
    n_len = len(q)
    info = {}
    for n in range(MIN_SEED - 1, n_len):
        if q[n] > 3:
            continue
        m = n
        while m > 0 and q[m - 1] < 4 and b.count(q[m - 1:n + 1]) > 0:
            m -= 1
        if n - m + 1 >= MIN_SEED:
            info[n] = m
    out = []
    for n in sorted(info):
        if n + 1 in info and info[n + 1] == info[n]:
            continue
        m = info[n]
        k, s = b.interval(q[m:n + 1])
        out.append((k, s, m, n))
    return out


def round3(b: Brute, q: np.ndarray, max_intv=20):
    out = []
    m = 0
    L = len(q)
    while m < L:
        next_m = m + 1
        if q[m] < 4:
            for n in range(m + 1, L):
                next_m = n + 1
                if q[n] > 3:
                    break
                c = b.count(q[m:n + 1])
                if c < max_intv and n - m + 1 >= MIN_SEED + 1:
                    if c > 0:
                        k, s = b.interval(q[m:n + 1])
                        out.append((k, s, m, n))
                    break
        m = next_m
    return out


def round2(b: Brute, q: np.ndarray, smems):
    """reseedV2 (bwa/src/seed.cu:639-913) over brute-force intervals."""
    out = []
    split_len = int(MIN_SEED * 1.5 + .499)
    for (_, s, m0, n0) in smems:
        if n0 - m0 + 1 < split_len or s > 10:
            continue
        min_intv = s + 1
        pivot = (m0 + n0 + 1) >> 1
        if pivot > 77 or q[pivot] > 3:
            continue
        leps = []
        cur = (pivot, pivot)
        for n in range(pivot, len(q) - 1):
            old = cur
            if q[n + 1] > 3:
                break
            cur = (pivot, n + 1)
            if b.count(q[cur[0]:cur[1] + 1]) < b.count(q[old[0]:old[1] + 1]):
                leps.append(old)
                if b.count(q[cur[0]:cur[1] + 1]) < min_intv:
                    break
        if b.count(q[cur[0]:cur[1] + 1]) >= min_intv:
            leps.append(cur)
        leps = leps[::-1]
        seeds = []
        m = pivot
        while m > 0:
            if q[m - 1] > 3:
                break
            curr = min_intv
            new, collected = [], False
            for (a, e) in leps:
                c = b.count(q[m - 1:e + 1])
                if c >= curr:
                    new.append((m - 1, e))
                    curr = c
                elif not collected and e - a + 1 >= MIN_SEED:
                    seeds.append((a, e))
                    collected = True
            leps = new
            if not leps:
                break
            m -= 1
        if leps and leps[0][1] - leps[0][0] + 1 >= MIN_SEED:
            seeds.append(leps[0])
        for (a, e) in seeds:
            k, s2 = b.interval(q[a:e + 1])
            out.append((k, s2, a, e))
    return out


@pytest.fixture(scope="module")
def small(built, tmp_path_factory):
    tmp = tmp_path_factory.mktemp("seed")
    fa, fwd = small_reference(tmp)
    prefix = tmp / "small"
    g3sa_amd.build_index(str(fa), str(prefix), 4)
    text, sa = true_sa(fwd)
    return prefix, fwd, Brute(text, sa)


def reads_for(fwd: np.ndarray, rng, count=40):
    text = np.concatenate([fwd, 3 - fwd[::-1]])
    reads = []
    for _ in range(count):
        L = int(rng.integers(30, 121))
        p = int(rng.integers(0, len(text) - L))
        r = text[p:p + L].copy()
        for _ in range(int(rng.integers(0, 4))):
            r[int(rng.integers(0, L))] = rng.integers(0, 4)
        if rng.random() < 0.2:
            r[int(rng.integers(0, L))] = 4
        reads.append(r.astype(np.uint8))
    # a tandem read and a read crossing the contig 0/1 boundary
    reads.append(np.tile(np.array([0, 1, 2], np.uint8), 30))
    reads.append(text[1480:1580].astype(np.uint8))
    return reads


def test_seeding_matches_bruteforce(small, oracle):
    prefix, fwd, b = small
    idx = oracle.load(prefix)
    rng = np.random.default_rng(11)
    try:
        for q in reads_for(fwd, rng):
            got = [(k, s, m, n) for (k, _l, s, m, n) in oracle.seed_intervals(idx, q)]
            r1 = round1(b, q)
            r2 = round2(b, q, r1)
            r3 = round3(b, q)
            assert got == r1 + r2 + r3, q.tolist()
    finally:
        oracle.lib.orc_index_free(idx)


@pytest.fixture(scope="module")
def planted(built, tmp_path_factory):
    """A text where a 50 bp segment occurs 3 times and a 30 bp one 6 times, so
    that long unique SMEMs have repeated middles (round-2 seeds)."""
    tmp = tmp_path_factory.mktemp("planted")
    rng = np.random.default_rng(17)
    fwd = rng.integers(0, 4, 4000)
    x, y = rng.integers(0, 4, 50), rng.integers(0, 4, 30)
    for p in (300, 1500, 2900):
        fwd[p:p + 50] = x
    for p in (700, 1100, 2000, 2300, 3300, 3700):
        fwd[p:p + 30] = y
    fa = tmp / "planted.fa"
    fa.write_text(">p\n" + "".join("ACGT"[v] for v in fwd) + "\n")
    prefix = tmp / "planted"
    g3sa_amd.build_index(str(fa), str(prefix), 4)
    text, sa = true_sa(fwd)
    return prefix, fwd, text, Brute(text, sa)


def test_reseeding_matches_bruteforce(planted, oracle):
    prefix, fwd, text, b = planted
    reads = []
    for p in (300, 1500, 2900):
        reads.append(fwd[p - 40:p + 90])            # flank + segment + flank
        reads.append(fwd[p - 20:p + 70])
    for p in (700, 2000, 3300):
        reads.append(fwd[p - 45:p + 75])
        r = fwd[p - 45:p + 75].copy()
        r[5] = (r[5] + 1) % 4
        reads.append(r)
    n = len(text)
    reads.append(text[n - 2900 - 90: n - 2900 + 40])   # reverse strand copy
    idx = oracle.load(prefix)
    seeds2 = 0
    try:
        for q in reads:
            q = np.asarray(q, dtype=np.uint8)
            got = [(k, s, m, e) for (k, _l, s, m, e) in oracle.seed_intervals(idx, q)]
            r1 = round1(b, q)
            r2 = round2(b, q, r1)
            r3 = round3(b, q)
            seeds2 += len(r2)
            assert got == r1 + r2 + r3, q.tolist()
    finally:
        oracle.lib.orc_index_free(idx)
    assert seeds2 >= 5  # the second round really ran
