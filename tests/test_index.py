"""Index construction and the oracle's index layer.

Pins: the reference's E. coli fixtures (.pac .ann .amb .sa .counts, byte
equality) and a brute-force suffix array / Occ on small multi-contig texts.
"""
from __future__ import annotations

import numpy as np
import pytest

import g3sa_amd
from conftest import ROOT, gen

GOLD = ROOT / "tests" / "golden" / "ecoli"


def decode_pac(pac: bytes, l_pac: int) -> np.ndarray:
    raw = np.frombuffer(pac, dtype=np.uint8)
    i = np.arange(l_pac)
    return (raw[i >> 2] >> ((3 - (i & 3)) * 2)) & 3


def test_ecoli_fixtures_byte_identical(built, tmp_path):
    ann = (GOLD / "ecoli.ann").read_text().split("\n")
    l_pac = int(ann[0].split()[0])
    _, name, anno = ann[1].split(" ", 2)
    seq = decode_pac((GOLD / "ecoli.pac").read_bytes(), l_pac)
    fa = tmp_path / "ecoli.fa"
    letters = np.frombuffer(b"ACGT", dtype=np.uint8)[seq].tobytes().decode()
    with open(fa, "w") as f:
        f.write(f">{name} {anno}\n")
        for i in range(0, len(letters), 60):
            f.write(letters[i:i + 60] + "\n")
    g3sa_amd.build_index(str(fa), str(tmp_path / "ecoli"), 4)
    for ext in ["pac", "ann", "amb", "sa", "counts"]:
        assert (tmp_path / f"ecoli.{ext}").read_bytes() == (GOLD / f"ecoli.{ext}").read_bytes(), ext


def small_reference(tmp_path, seed=3, lens=(1500, 700, 2100)):
    rng = np.random.default_rng(seed)
    fa = tmp_path / "small.fa"
    seqs = []
    with open(fa, "w") as f:
        for c, L in enumerate(lens):
            s = rng.integers(0, 4, L)
            if c == 1:  # a repeat shared with contig 0
                s[100:400] = seqs[0][500:800]
            seqs.append(s)
            f.write(f">ctg{c}\n" + "".join("ACGT"[x] for x in s) + "\n")
    return fa, np.concatenate(seqs)


def true_sa(fwd: np.ndarray):
    text = np.concatenate([fwd, 3 - fwd[::-1]])
    t = bytes((text + 1).tolist()) + b"\x00"
    sa = sorted(range(len(t)), key=lambda i: t[i:])
    return text, np.array(sa, dtype=np.int64)


def test_small_index_bruteforce(built, oracle, tmp_path):
    fa, fwd = small_reference(tmp_path)
    prefix = tmp_path / "small"
    g3sa_amd.build_index(str(fa), str(prefix), 2)
    text, sa = true_sa(fwd)
    n = len(text)
    idx = oracle.load(prefix)
    try:
        got = np.array([oracle.lib.orc_sa(idx, k) for k in range(1, n + 1)])
        np.testing.assert_array_equal(got, sa[1:])
        bwt = np.where(sa == 0, 4, text[(sa - 1) % n])
        assert oracle.lib.orc_index_primary(idx) == int(np.flatnonzero(sa == 0)[0])
        for c in range(4):
            cum = np.cumsum(bwt == c)
            probe = np.arange(0, n + 1, 7)
            occ = np.array([oracle.lib.orc_occ(idx, int(p), c) for p in probe])
            np.testing.assert_array_equal(occ, cum[probe])
    finally:
        oracle.lib.orc_index_free(idx)
    # the dense sample the GPU reads holds SA[2*i]
    raw = (tmp_path / "small.g3sa").read_bytes()
    hdr = np.frombuffer(raw[:56], dtype=np.uint64)
    assert int(hdr[1]) == n + 1
    samples = np.frombuffer(raw[56:], dtype=np.uint32)
    np.testing.assert_array_equal(samples, sa[::2])


def test_generator_deterministic(built, tmp_path):
    a, b = tmp_path / "a.fa", tmp_path / "b.fa"
    gen("ref", "-o", str(a), "--len", "20000", "--contigs", "2", "--seed", "9", "--repeat-frac", "0.2",
        "--families", "5", "--fam-min", "300", "--fam-max", "900")
    gen("ref", "-o", str(b), "--len", "20000", "--contigs", "2", "--seed", "9", "--repeat-frac", "0.2",
        "--families", "5", "--fam-min", "300", "--fam-max", "900")
    assert a.read_bytes() == b.read_bytes()
    ra, rb = tmp_path / "a.fq", tmp_path / "b.fq"
    for out in (ra, rb):
        gen("reads", "-r", str(a), "-o", str(out), "--n", "50", "--paired", "--seed", "4")
    assert ra.read_bytes() == rb.read_bytes()
    lines = ra.read_text().split("\n")
    assert lines[0].startswith("@r0/1 ") and lines[4].startswith("@r0/2 ")
    assert all(not q.startswith("@") for q in lines[3::4])


def test_bad_index_arguments(built, tmp_path):
    with pytest.raises(RuntimeError):
        g3sa_amd.build_index(str(tmp_path / "missing.fa"), str(tmp_path / "x"), 4)
    fa, _ = small_reference(tmp_path)
    with pytest.raises(RuntimeError):
        g3sa_amd.build_index(str(fa), str(tmp_path / "x"), 3)  # not a power of two
