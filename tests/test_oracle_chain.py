"""The oracle's chaining, chain weight and chain filter against independent
restatements.

Chaining (BTreeChaining over klib's B-tree, bwa/src/aux.cu:1755-1825) is
recomputed with a plain sorted list in place of the B-tree: with distinct
chain positions the B-tree's lower-bound query is the largest position <= the
seed's, so any slip in the B-tree emulation (split, descent, in-order
traversal) shows up.  Chain weight (aux.cu:91-110) is recomputed as the
closed form min(|union of query intervals|, |union of reference intervals|).
The chain filter (aux.cu:1890-1988, sequential fixed point) is checked by its
characterisation: a chain is dropped iff an earlier kept chain dominates it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import g3sa_amd
from oracle_lib import OrcSeed
from test_index import small_reference

W, MAX_GAP, MASK_LEVEL, DROP_RATIO, MIN_SEED = 100, 10000, 0.5, 0.5, 19


def merge(chain, s, l_pac):
    """0: not merged, 1: contained (dropped), 2: appended."""
    first, last = chain["seeds"][0], chain["seeds"][-1]
    if s["rid"] != chain["rid"]:
        return 0
    if (s["qbeg"] >= first["qbeg"] and s["qbeg"] + s["len"] <= last["qbeg"] + last["len"]
            and s["rbeg"] >= first["rbeg"] and s["rbeg"] + s["len"] <= last["rbeg"] + last["len"]):
        return 1
    if (last["rbeg"] < l_pac or first["rbeg"] < l_pac) and s["rbeg"] >= l_pac:
        return 0
    x, y = s["qbeg"] - last["qbeg"], s["rbeg"] - last["rbeg"]
    if y >= 0 and abs(x - y) <= W and x - last["len"] < MAX_GAP and y - last["len"] < MAX_GAP:
        chain["seeds"].append(s)
        return 2
    return 0


def brute_chain(seeds, l_pac):
    chains = []  # kept sorted by pos
    owner = [None] * len(seeds)
    for i, s in enumerate(seeds):
        cands = [c for c in chains if c["pos"] <= s["rbeg"]]
        m = 0
        if cands:
            best = max(cands, key=lambda c: c["pos"])
            m = merge(best, s, l_pac)
            if m == 2:
                owner[i] = best
        if m == 0:
            if s["rid"] < 0 or s["rid"] >= 455:
                continue
            c = {"pos": s["rbeg"], "rid": s["rid"], "seeds": [s]}
            chains.append(c)
            chains.sort(key=lambda c: c["pos"])
            owner[i] = c
    ids = {id(c): k for k, c in enumerate(chains)}
    return chains, [ids[id(o)] if o is not None else -1 for o in owner]


def weight(chain):
    q, r = set(), set()
    for s in chain["seeds"]:
        q.update(range(s["qbeg"], s["qbeg"] + s["len"]))
        r.update(range(s["rbeg"], s["rbeg"] + s["len"]))
    return min(len(q), len(r))


def brute_filter(chains, weights):
    order = sorted(range(len(chains)), key=lambda c: -weights[c])  # stable
    kept = []
    for i in order:
        bi = chains[i]["seeds"][0]["qbeg"]
        ei = chains[i]["seeds"][-1]["qbeg"] + chains[i]["seeds"][-1]["len"]
        dropped = False
        for j in kept:
            bj = chains[j]["seeds"][0]["qbeg"]
            ej = chains[j]["seeds"][-1]["qbeg"] + chains[j]["seeds"][-1]["len"]
            b_max, e_min = max(bi, bj), min(ei, ej)
            if e_min > b_max:
                min_l = min(ei - bi, ej - bj)
                if (e_min - b_max >= min_l * MASK_LEVEL and min_l < MAX_GAP and weights[i] < weights[j] * DROP_RATIO
                        and weights[j] - weights[i] >= 2 * MIN_SEED):
                    dropped = True
                    break
        if not dropped:
            kept.append(i)
    return kept


@pytest.fixture(scope="module")
def small_idx(built, oracle, tmp_path_factory):
    tmp = tmp_path_factory.mktemp("chain")
    fa, fwd = small_reference(tmp)
    prefix = tmp / "small"
    g3sa_amd.build_index(str(fa), str(prefix), 4)
    idx = oracle.load(prefix)
    yield idx, len(fwd)
    oracle.lib.orc_index_free(idx)


def run_oracle(oracle, idx, seeds):
    n = len(seeds)
    arr = (OrcSeed * max(n, 1))()
    for i, s in enumerate(seeds):
        arr[i].rbeg, arr[i].qbeg, arr[i].len, arr[i].rid = s["rbeg"], s["qbeg"], s["len"], s["rid"]
    owner = np.zeros(max(n, 1), np.int32)
    w = np.zeros(max(n, 1), np.int32)
    kept = np.zeros(max(n, 1), np.int32)
    nk = C.c_int()
    nc = oracle.lib.orc_chain_dump(idx, C.byref(oracle.opt), arr, n, owner.ctypes.data, w.ctypes.data,
                                   kept.ctypes.data, C.byref(nk))
    return nc, owner[:n].tolist(), w[:nc].tolist(), kept[:nk.value].tolist()


def synthetic_seeds(rng, l_pac):
    """Seeds of one read in sort order: colinear runs around a few anchors on
    either strand, off-diagonal and cross-contig noise; distinct positions."""
    seeds, used = [], set()
    for _ in range(int(rng.integers(1, 5))):
        anchor = int(rng.integers(0, 2 * l_pac - 400))
        rid = int(rng.integers(0, 3))
        q = int(rng.integers(0, 40))
        while q < 150:
            ln = int(rng.integers(19, 45))
            shift = int(rng.integers(-3, 4)) if rng.random() < 0.3 else 0
            if rng.random() < 0.1:
                shift = int(rng.integers(-150, 150))
            seeds.append({"qbeg": q, "len": ln, "rbeg": anchor + q + shift, "rid": rid})
            q += int(rng.integers(5, 40))
    if rng.random() < 0.2:
        seeds.append({"qbeg": 3, "len": 20, "rbeg": int(rng.integers(0, l_pac)), "rid": -1})
    if rng.random() < 0.1:
        seeds.append({"qbeg": 5, "len": 25, "rbeg": int(rng.integers(0, l_pac)), "rid": 460})
    out = []
    for s in sorted(seeds, key=lambda s: (s["qbeg"], s["len"])):
        if s["rbeg"] in used or s["rbeg"] < 0:
            continue
        used.add(s["rbeg"])
        out.append(s)
    return out


def check(oracle, idx, l_pac, seeds):
    nc, owner, w, kept = run_oracle(oracle, idx, seeds)
    chains, bowner = brute_chain(seeds, l_pac)
    assert nc == len(chains)
    assert owner == bowner
    bw = [weight(c) for c in chains]
    assert w == bw
    assert kept == brute_filter(chains, bw)
    return nc, len(kept)


def test_chaining_synthetic(small_idx, oracle):
    idx, l_pac = small_idx
    rng = np.random.default_rng(21)
    total_chains = total_dropped = 0
    for _ in range(400):
        nc, nk = check(oracle, idx, l_pac, synthetic_seeds(rng, l_pac))
        total_chains += nc
        total_dropped += nc - nk
    # the cases exercise B-tree splits (>11 chains) and the filter
    assert total_chains > 1500 and total_dropped > 50


def test_chaining_many_chains(small_idx, oracle):
    """One read with 200 unrelated seeds: a B-tree of several levels."""
    idx, l_pac = small_idx
    rng = np.random.default_rng(5)
    pos = rng.choice(np.arange(0, 2 * l_pac - 100), size=200, replace=False)
    seeds = [{"qbeg": int(i % 120), "len": 19 + int(i % 7), "rbeg": int(p), "rid": int(i % 3)}
             for i, p in enumerate(pos)]
    seeds.sort(key=lambda s: (s["qbeg"], s["len"]))
    nc, _ = check(oracle, idx, l_pac, seeds)
    assert nc > 150


def test_chaining_real_reads(small_idx, oracle, tmp_path):
    """Seeds the oracle itself resolves for reads of the small reference."""
    idx, l_pac = small_idx
    rng = np.random.default_rng(8)
    fa, fwd = small_reference(tmp_path)
    text = np.concatenate([fwd, 3 - fwd[::-1]])
    buf = (OrcSeed * 4096)()
    n_checked = 0
    for _ in range(200):
        L = int(rng.integers(60, 200))
        p = int(rng.integers(0, len(text) - L))
        r = text[p:p + L].astype(np.uint8).copy()
        for _ in range(int(rng.integers(0, 5))):
            r[int(rng.integers(0, L))] = rng.integers(0, 4)
        n = oracle.lib.orc_read_seeds(idx, C.byref(oracle.opt), r.ctypes.data, L, buf, 4096)
        seeds = [{"rbeg": buf[i].rbeg, "qbeg": buf[i].qbeg, "len": buf[i].len, "rid": buf[i].rid} for i in range(n)]
        chains, _ = brute_chain(seeds, l_pac)
        if len({c["pos"] for c in chains}) != len(chains):
            continue  # chains at equal positions: the B-tree's duplicate-key order is not a sorted list's
        check(oracle, idx, l_pac, seeds)
        n_checked += 1
    assert n_checked > 150


@pytest.mark.parametrize("wj,wi,qi,expect_drop", [
    (75, 37, 10, True),    # wi < wj/2 and wj - wi == 38: dropped (>= 38)
    (74, 36, 10, True),
    (76, 38, 10, False),   # wi == wj/2: kept (strict <)
    (75, 38, 10, False),   # wj - wi == 37: kept
    (80, 30, 65, True),    # query overlap 15 == min_l/2 (min_l 30): dropped (>=)
    (80, 30, 66, False),   # overlap 14 < 15: kept
    (80, 30, 80, False),   # no overlap: kept
])
def test_chain_filter_boundaries(small_idx, oracle, wj, wi, qi, expect_drop):
    """Hand-built pairs at the filter's thresholds (aux.cu:1890-1988): a
    heavy chain j of one seed [0, wj) and a light chain i of one seed
    [qi, qi + wi) far away on the reference; expected outcomes worked out by
    hand from the rule."""
    idx, l_pac = small_idx
    seeds = sorted([{"qbeg": 0, "len": wj, "rbeg": 100, "rid": 0},
                    {"qbeg": qi, "len": wi, "rbeg": 1000, "rid": 0}], key=lambda s: (s["qbeg"], s["len"]))
    nc, _, w, kept = run_oracle(oracle, idx, seeds)
    assert nc == 2 and sorted(w) == sorted([wj, wi])
    assert len(kept) == (1 if expect_drop else 2)
    check(oracle, idx, l_pac, seeds)
