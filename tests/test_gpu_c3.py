"""The c3 workload's shape at a tenth of its reads: a 1 Mbp uniform
reference and 200,000 reads of 70-250 bp, 5% of them at 8% divergence with
a 20-40 bp random tail (the reads whose extensions fall back to the
one-wavefront engine and whose alignments are clipped), through the plain,
batched and batched + learned-index forms, every field against the oracle;
then as one 200,000-read graph, and as its two halves on a context and a
sibling at the same time."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import g3sa_amd
from conftest import gen, oracle_threads
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

REF = ["--len", "1000000", "--contigs", "1", "--seed", "1"]
READS = ["--n", "200000", "--len-min", "70", "--len-max", "250", "--sub", "0.01", "--indel", "0.0005",
         "--hard-frac", "0.05", "--hard-div", "0.08", "--tail-min", "20", "--tail-max", "40", "--seed", "3"]


def test_c3_parity(built, oracle, tmp_path):
    fa, fq = tmp_path / "c3.fa", tmp_path / "c3.fq"
    print("[c3] reference, reads, index", flush=True)
    gen("ref", "-o", str(fa), *REF)
    gen("reads", "-r", str(fa), "-o", str(fq), *READS)
    g3sa_amd.build_index_dev(str(fa), str(tmp_path / "c3"), 4)
    prefix = tmp_path / "c3"
    names, seq, off, _ = g3sa_amd.read_fastq(str(fq))
    n = len(off) - 1
    h = n // 2
    parts = [(seq[:off[h]], off[:h + 1].astype(np.int32)), (seq[off[h]:off[n]], (off[h:] - off[h]).astype(np.int32))]
    print("[c3] oracle", flush=True)
    idx = oracle.load(prefix)
    try:
        ref = oracle.align(idx, seq, off, threads=oracle_threads())
        ref_half = [oracle.align(idx, s, o, threads=oracle_threads()) for s, o in parts]
    finally:
        oracle.lib.orc_index_free(idx)
    with g3sa_amd.Aligner(str(prefix), max_reads=n, max_bases=int(off[-1])) as al:
        for form in ("plain", "batched", "learned"):
            print(f"[c3] {n} reads, {form} form", flush=True)
            al.set_form("plain" if form == "plain" else "batched")
            if form == "learned":
                al.set_learned(1)
            got = al.align(seq, off)
            compare(got, ref, f"c3 {form}")
        assert got.n_alns > 0.9 * n
        print("[c3] as a graph", flush=True)
        al.set_graph(True)
        for rep in range(2):
            got = al.align(seq, off)
            assert al.graph_info()["graph_used"]
            compare(got, ref, f"c3 graph run {rep}")
        al.set_graph(False)
        sib = al.sibling(n - h + 16, int(off[n] - off[h]) + 1024)
        res = [None, None]

        def work(k, c):
            res[k] = c.align(*parts[k])

        print("[c3] two sibling contexts on the halves", flush=True)
        th = [threading.Thread(target=work, args=(0, al)), threading.Thread(target=work, args=(1, sib))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        sib.close()
    for k in range(2):
        compare(res[k], ref_half[k], f"c3 sibling half {k}")
