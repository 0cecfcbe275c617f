"""The oracle's region filter and region sort (filterRegions + sortRegions,
bwa/src/aux.cu:2408-2613; DR-8 ascending-index anchors) against hand-worked
known answers at every threshold and against a restatement on random
regions.

Rules (default options): regions scoring below T = 30 go; of two regions on
one contig whose reference overlap exceeds 0.95f x the shorter reference span
and whose query overlap exceeds 0.95f x the shorter query span, the lower
score goes (the later one on a tie); a survivor whose query overlap with a
higher-scoring survivor is >= 0.5f x the shorter query span is secondary and
is dropped (flag 0x8 unset); the rest are sorted by score, descending.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import g3sa_amd
from test_index import small_reference

F95 = np.float32(0.95)
F50 = np.float32(0.5)


@pytest.fixture(scope="module")
def small_idx(built, oracle, tmp_path_factory):
    tmp = tmp_path_factory.mktemp("region")
    fa, fwd = small_reference(tmp)
    prefix = tmp / "small"
    g3sa_amd.build_index(str(fa), str(prefix), 4)
    idx = oracle.load(prefix)
    yield idx
    oracle.lib.orc_index_free(idx)


def run(oracle, idx, regs):
    n = len(regs)
    cols = {k: np.array([r[k] for r in regs] or [0], dtype=np.int64 if k in ("rb", "re") else np.int32)
            for k in ("rb", "re", "qb", "qe", "score", "rid")}
    oi = np.zeros(max(n, 1), np.int32)
    os_ = np.zeros(max(n, 1), np.int32)
    k = oracle.lib.orc_region_dump(idx, C.byref(oracle.opt), n, *[cols[c].ctypes.data for c in
                                   ("rb", "re", "qb", "qe", "score", "rid")], oi.ctypes.data, os_.ctypes.data)
    return oi[:k].tolist(), os_[:k].tolist()


def restate(regs, T=30):
    kept = [r["score"] >= T for r in regs]
    n = len(regs)
    for i in range(n):
        if not kept[i]:
            continue
        for j in range(i + 1, n):
            if not kept[j]:
                continue
            a, b = regs[i], regs[j]
            if a["rid"] != b["rid"]:
                continue
            if (a["rb"] < b["rb"] and a["re"] < b["rb"]) or (b["rb"] < a["rb"] and b["re"] < a["rb"]):
                continue
            o_r = a["re"] - b["rb"] if a["rb"] < b["rb"] else b["re"] - a["rb"]
            o_q = a["qe"] - b["qb"] if a["qb"] < b["qb"] else b["qe"] - a["qb"]
            mr = min(a["re"] - a["rb"], b["re"] - b["rb"])
            mq = min(a["qe"] - a["qb"], b["qe"] - b["qb"])
            if np.float32(o_r) > F95 * np.float32(mr) and np.float32(o_q) > F95 * np.float32(mq):
                if a["score"] < b["score"]:
                    kept[i] = False
                    break
                kept[j] = False
    surv = [i for i in range(n) if kept[i]]
    sec = {i: -1 for i in surv}
    for x, i in enumerate(surv):
        for y in range(x + 1, len(surv)):
            j = surv[y]
            if sec[i] >= 0 or sec[j] >= 0:
                continue
            a, b = regs[i], regs[j]
            qb_max, qe_min = max(a["qb"], b["qb"]), min(a["qe"], b["qe"])
            if qe_min > qb_max:
                min_l = min(a["qe"] - a["qb"], b["qe"] - b["qb"])
                if np.float32(qe_min - qb_max) >= np.float32(min_l) * F50:
                    if a["score"] > b["score"]:
                        sec[j] = x
                    else:
                        sec[i] = y
    prim = [i for i in surv if sec[i] < 0]
    # stable bubble sort by score, descending (no alt contigs here)
    out = prim[:]
    for a in range(len(out) - 1):
        swapped = False
        for b in range(len(out) - 1 - a):
            if regs[out[b]]["score"] < regs[out[b + 1]]["score"]:
                out[b], out[b + 1] = out[b + 1], out[b]
                swapped = True
        if not swapped:
            break
    return out


def R(rb, re, qb, qe, score, rid=0):
    return {"rb": rb, "re": re, "qb": qb, "qe": qe, "score": score, "rid": rid}


CASES = [
    ("below T", [R(0, 100, 0, 100, 29), R(5000, 5100, 0, 100, 30)], [1]),
    # 100-long twins shifted by 5: overlap 95 is not > 95.0 (0.95f*100 = 95.0): both survive the
    # redundancy test; the second is secondary (query overlap 95 >= 50) and dropped
    ("overlap at 0.95 boundary", [R(0, 100, 0, 100, 60), R(5, 105, 0, 100, 50)], [0]),
    # same with equal scores: not redundant, so the secondary rule applies and, on a score tie,
    # makes the earlier region the secondary one: the later region is the output
    ("overlap at 0.95 boundary, tie", [R(0, 100, 0, 100, 60), R(5, 105, 0, 100, 60)], [1]),
    # shifted by 4: overlap 96 > 95: the lower score is dropped as redundant
    ("redundant lower dropped", [R(0, 100, 0, 100, 50), R(4, 104, 0, 100, 60)], [1]),
    ("redundant tie drops later", [R(0, 100, 0, 100, 60), R(4, 104, 0, 100, 60)], [0]),
    ("other contig kept", [R(0, 100, 0, 100, 50), R(4, 104, 0, 100, 60, rid=1)], [1]),
    # disjoint query halves: two primaries, sorted by score
    ("two primaries sorted", [R(0, 75, 0, 75, 40), R(9000, 9075, 75, 150, 70)], [1, 0]),
    # query overlap exactly half of the shorter: secondary
    ("secondary at 0.5", [R(0, 100, 0, 100, 80), R(9000, 9060, 70, 130, 40)], [0]),
    ("not secondary below 0.5", [R(0, 100, 0, 100, 80), R(9000, 9060, 71, 131, 40)], [0, 1]),
]


@pytest.mark.parametrize("name,regs,expect", CASES, ids=[c[0] for c in CASES])
def test_region_filter_known_answers(small_idx, oracle, name, regs, expect):
    got, _ = run(oracle, small_idx, regs)
    assert got == expect
    assert restate(regs) == expect


def test_region_filter_random(small_idx, oracle):
    rng = np.random.default_rng(3)
    for _ in range(3000):
        regs = []
        base = int(rng.integers(0, 3000))
        for _ in range(int(rng.integers(1, 9))):
            qb = int(rng.integers(0, 120))
            qe = qb + int(rng.integers(10, 151 - qb)) if qb < 141 else 150
            rb = base + int(rng.integers(-8, 8)) + qb if rng.random() < 0.7 else int(rng.integers(0, 5000))
            re_ = rb + (qe - qb) + int(rng.integers(-3, 4))
            regs.append(R(rb, max(re_, rb + 1), qb, qe, int(rng.integers(20, 150)), int(rng.integers(0, 2))))
        assert run(oracle, small_idx, regs)[0] == restate(regs), regs
