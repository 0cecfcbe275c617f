"""Extension and global alignment in the oracle.

* The extension's cell-set specification against a lane-by-lane emulation of
  the reference's 32-thread warp program (bwa/src/ksw.cu:814-1034).
* Its score and end position against a textbook full-matrix DP where the band
  cannot bind (w large, no tile or corner limits reachable).
* ksw_global2 against a textbook Gotoh global alignment, and its CIGAR replayed
  against the sequences.
"""
from __future__ import annotations

import numpy as np
import pytest

NEG = -(1 << 30)


def mutate(rng, q, rate=0.08):
    t = []
    i = 0
    while i < len(q):
        u = rng.random()
        if u < rate / 3:
            i += 1
            continue
        if u < 2 * rate / 3:
            t.append(int(rng.integers(0, 4)))
        t.append(int(q[i]) if rng.random() > rate / 3 else int(rng.integers(0, 4)))
        i += 1
    t += rng.integers(0, 4, int(rng.integers(0, 20))).tolist()
    return np.array(t, dtype=np.uint8)


def test_extension_spec_matches_warp_emulation(oracle):
    rng = np.random.default_rng(5)
    for it in range(3000):
        qlen = int(rng.integers(1, 200))
        q = rng.integers(0, 4, qlen).astype(np.uint8)
        t = mutate(rng, q)
        if rng.random() < 0.3:
            t = t[: int(rng.integers(0, len(t) + 1))]
        w = int(rng.choice([1, 4, 16, 40, 100]))
        h0 = int(rng.integers(0, 300))
        gaps = (6, 1, 6, 1) if rng.random() < 0.7 else tuple(int(x) for x in rng.integers(1, 8, 4))
        a = oracle.extend(q, t, h0, w=w, gaps=gaps)
        b = oracle.extend(q, t, h0, w=w, gaps=gaps, emulate=True)
        assert a == b, (it, qlen, len(t), w, h0, gaps, a, b)


def textbook_extend(q, t, h0, mat, od, ed, oi, ei):
    """Full-matrix local extension from h0 (BWA ksw_extend recurrences, no band,
    no z-drop): returns the best score and (row, col) of its first row's last
    maximal cell, or (-1, -1) when nothing beats h0."""
    Q, T = len(q), len(t)
    H = np.zeros((T + 1, Q + 1), dtype=np.int64)
    E = np.zeros((T + 1, Q + 1), dtype=np.int64)
    H[0, 0] = h0
    for j in range(1, Q + 1):
        H[0, j] = max(h0 - oi - ei * j, 0)
    for i in range(1, T + 1):
        H[i, 0] = max(h0 - od - ed * i, 0)
    best, bi, bj = h0, -1, -1
    for i in range(1, T + 1):
        f = 0
        for j in range(1, Q + 1):
            d = H[i - 1, j - 1]
            M = d + mat[t[i - 1] * 5 + q[j - 1]] if d else 0
            h = max(M, E[i, j], f)
            H[i, j] = h
            if i < T:
                E[i + 1, j] = max(E[i, j] - ed, M - od - ed, 0)
            f = max(f - ei, M - oi - ei, 0)
            if h > best:
                best, bi, bj = h, i - 1, j - 1
            elif h == best and best > h0 and bi == i - 1:
                bj = j - 1
    return best, bi, bj


def test_extension_against_textbook_dp(oracle):
    rng = np.random.default_rng(9)
    mat = oracle.mat()
    for _ in range(150):
        qlen = int(rng.integers(1, 60))
        q = rng.integers(0, 4, qlen).astype(np.uint8)
        t = mutate(rng, q, rate=0.1)[:60]
        h0 = int(rng.integers(5, 80))
        s, qle, tle = oracle.extend(q, t, h0, w=1000, end_bonus=1000)
        best, bi, bj = textbook_extend(q, t, h0, mat, 6, 1, 6, 1)
        assert s == best
        if best > h0:
            assert (tle - 1, qle - 1) == (bi, bj)


def textbook_global(q, t, mat, od, ed, oi, ei):
    Q, T = len(q), len(t)
    H = np.full((T + 1, Q + 1), NEG, dtype=np.int64)
    E = np.full((T + 1, Q + 1), NEG, dtype=np.int64)
    F = np.full((T + 1, Q + 1), NEG, dtype=np.int64)
    H[0, 0] = 0
    for j in range(1, Q + 1):
        H[0, j] = -(oi + ei * j)
        F[0, j] = H[0, j]
    for i in range(1, T + 1):
        H[i, 0] = -(od + ed * i)
        E[i, 0] = H[i, 0]
    for i in range(1, T + 1):
        for j in range(1, Q + 1):
            E[i, j] = max(E[i - 1, j] - ed, H[i - 1, j] - od - ed)
            F[i, j] = max(F[i, j - 1] - ei, H[i, j - 1] - oi - ei)
            H[i, j] = max(H[i - 1, j - 1] + mat[t[i - 1] * 5 + q[j - 1]], E[i, j], F[i, j])
    return int(H[T, Q])


def replay(cig, q, t, mat, od, ed, oi, ei):
    x = y = 0
    score = 0
    for c in cig:
        op, ln = int(c) & 0xF, int(c) >> 4
        if op == 0:
            score += sum(int(mat[t[y + k] * 5 + q[x + k]]) for k in range(ln))
            x += ln
            y += ln
        elif op == 1:
            score -= oi + ei * ln
            x += ln
        elif op == 2:
            score -= od + ed * ln
            y += ln
    return x, y, score


def test_global_alignment(oracle):
    rng = np.random.default_rng(21)
    mat = oracle.mat()
    for _ in range(120):
        qlen = int(rng.integers(5, 50))
        q = rng.integers(0, 4, qlen).astype(np.uint8)
        t = mutate(rng, q, rate=0.12)[: qlen + 8]
        if len(t) == 0:
            continue
        s, cig = oracle.global2(q, t, w=200)
        x, y, rs = replay(cig, q, t, mat, 6, 1, 6, 1)
        assert (x, y) == (len(q), len(t))
        assert rs == s
        # the reference's gap recurrences open from M only, so its optimum can
        # fall below Gotoh's (which also opens from E/F) but never above it
        assert s <= textbook_global(q, t, mat, 6, 1, 6, 1)
        if len(t) == len(q) and np.array_equal(q, t):
            assert s == len(q) and len(cig) == 1
