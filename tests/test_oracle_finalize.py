"""The oracle's finalize step (position, strand, contig, CIGAR, NM) and MAPQ
against what the data fixes.

Error-free reads sampled by the generator from known places of a random
reference must come back at exactly that contig, 0-based position and strand,
with CIGAR `<len>M`, NM 0 and the extension score equal to the read length;
reads with one planted substitution must report NM 1 (bwa/src/aux.cu:2952-3059
for the finalize rules).  MAPQ (mem_approx_mapq_se, aux.cu:676-700, with no
suboptimal hit) is checked against values worked out by hand.
"""
from __future__ import annotations

import numpy as np
import pytest

import g3sa_amd
from conftest import gen, oracle_threads


@pytest.fixture(scope="module")
def sampled(built, oracle, tmp_path_factory):
    tmp = tmp_path_factory.mktemp("fin")
    fa, fq, truth = tmp / "r.fa", tmp / "r.fq", tmp / "r.truth"
    gen("ref", "-o", str(fa), "--len", "399000", "--contigs", "3", "--seed", "31")
    gen("reads", "-r", str(fa), "-o", str(fq), "--n", "600", "--len-min", "60", "--len-max", "250", "--sub", "0",
        "--indel", "0", "--seed", "32", "--truth", str(truth))
    g3sa_amd.build_index(str(fa), str(tmp / "r"), 4)
    idx = oracle.load(tmp / "r")
    names, seq, off, _ = g3sa_amd.read_fastq(str(fq))
    rows = [ln.split("\t") for ln in truth.read_text().splitlines()]
    yield idx, seq, off, rows
    oracle.lib.orc_index_free(idx)


def contig_ids(oracle, idx):
    return {oracle.lib.orc_index_contig_name(idx, c).decode(): c for c in range(oracle.lib.orc_index_n_contigs(idx))}


def test_error_free_reads_land_where_sampled(sampled, oracle):
    idx, seq, off, rows = sampled
    ids = contig_ids(oracle, idx)
    ref = oracle.align(idx, seq, off, threads=oracle_threads())
    n_checked = 0
    for i, (name, ctg, start, strand) in enumerate(rows):
        a0, a1 = ref["aln_off"][i], ref["aln_off"][i + 1]
        assert a1 - a0 == 1, f"{name}: {a1 - a0} alignments"
        L = int(off[i + 1] - off[i])
        assert ref["rid"][a0] == ids[ctg]
        assert ref["pos"][a0] == int(start), name
        assert bool(ref["flag"][a0] & 0x10) == (strand == "-"), name
        assert ref["n_cigar"][a0] == 1 and ref["cigar"][ref["cigar_off"][a0]] == (L << 4), name
        assert ref["NM"][a0] == 0 and ref["score"][a0] == L
        n_checked += 1
    assert n_checked == 600


def test_one_substitution_counts_one_mismatch(sampled, oracle):
    idx, seq, off, rows = sampled
    seq = seq.copy()
    rng = np.random.default_rng(4)
    planted = []
    for i in range(len(off) - 1):
        L = int(off[i + 1] - off[i])
        p = int(rng.integers(30, max(31, L - 30)))  # away from the ends: no clipping
        seq[off[i] + p] = (seq[off[i] + p] + 1 + rng.integers(0, 3)) % 4
        planted.append(L)
    ref = oracle.align(idx, seq, off, threads=oracle_threads())
    for i, L in enumerate(planted):
        a0 = ref["aln_off"][i]
        assert ref["NM"][a0] == 1
        assert ref["score"][a0] == L - 1 - 4  # one match lost, one mismatch paid (a = 1, b = 4)


@pytest.mark.parametrize("score,l,expect", [
    (19, 30, 0),    # score <= min_seed_len * a: 0
    (25, 30, 32),   # l < 50: tmp = identity^2 = (1 - 1/30)^2 = 0.93444; 6.02 * 6 * tmp^2 = 31.54
    (30, 30, 60),   # 6.02 * 11 = 66.2, capped at 60
    (40, 100, 32),  # tmp = 3 / ln(100) * 0.88^2 = 0.50448; 6.02 * 21 * tmp^2 = 32.17
    (22, 60, 6),    # tmp = 3 / ln(60) * (1 - 38/300)^2 = 0.55886; 6.02 * 3 * tmp^2 = 5.64
])
def test_mapq_known_answers(oracle, score, l, expect):
    """mapQ_coef_fac is (int)log(50) = 3; the length factor uses logf."""
    assert oracle.lib.orc_mapq(oracle.opt, score, 0, l, 0, l) == expect
