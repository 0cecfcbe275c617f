"""The 100 Mbp workload at full index size: a 5-contig reference with 15%
interspersed repeats (GC per 10 kb window ~ N(0.41, 0.05), 200 families of
300 bp-6 kbp at 5-15% divergence), its index built on the GPU and compared
byte for byte with the host build, and 100,000 of its paired 2x150 bp reads
(aligned as single-end reads) through the plain, batched and batched +
learned-index forms, every field against the oracle; then the paired-end
statistics and pair choice of its 50,000 pairs against the oracle."""
from __future__ import annotations

import pytest

import g3sa_amd
import oracle_lib
from conftest import gen, oracle_threads
from test_gpu_pairs import _check
from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

REF = ["--len", "100000000", "--contigs", "5", "--seed", "11", "--gc-mean", "0.41", "--gc-sd", "0.05",
       "--repeat-frac", "0.15", "--families", "200", "--fam-min", "300", "--fam-max", "6000",
       "--div-min", "0.05", "--div-max", "0.15"]
READS = ["--n", "100000", "--paired", "--len", "150", "--insert-mean", "400", "--insert-sd", "50",
         "--sub-start", "0.005", "--sub-end", "0.015", "--indel", "0.0005", "--seed", "12"]


def test_c2_full_index_parity(built, oracle, tmp_path):
    fa, fq = tmp_path / "c2.fa", tmp_path / "c2.fq"
    print("[c2] reference and reads", flush=True)
    gen("ref", "-o", str(fa), *REF)
    gen("reads", "-r", str(fa), "-o", str(fq), *READS)
    print("[c2] index on the GPU and on the host", flush=True)
    g3sa_amd.build_index_dev(str(fa), str(tmp_path / "dev"), 4)
    g3sa_amd.build_index(str(fa), str(tmp_path / "host"), 4)
    for ext in ["bwt", "sa", "counts", "g3sa", "pac", "ann", "amb"]:
        assert (tmp_path / f"dev.{ext}").read_bytes() == (tmp_path / f"host.{ext}").read_bytes(), ext
        (tmp_path / f"host.{ext}").unlink()
    prefix = tmp_path / "dev"
    names, seq, off, _ = g3sa_amd.read_fastq(str(fq))
    print("[c2] oracle", flush=True)
    idx = oracle.load(prefix)
    try:
        ref = oracle.align(idx, seq, off, threads=oracle_threads())
    finally:
        oracle.lib.orc_index_free(idx)
    n = len(off) - 1
    with g3sa_amd.Aligner(str(prefix), max_reads=n, max_bases=int(off[-1])) as al:
        for form in ("plain", "batched", "learned"):
            print(f"[c2] {n} reads, {form} form", flush=True)
            al.set_form("plain" if form == "plain" else "batched")
            if form == "learned":
                info = al.set_learned(1)
                assert info["max_err"] <= 256 and info["keys_over_window"] == 0, info
            got = al.align(seq, off)
            compare(got, ref, f"c2 {form}")
            assert got.work["n_seeds"] == ref["work"]["sa_reads"]
        print(f"[c2] {n // 2} pairs: insert-size statistics and pair choice", flush=True)
        l_pac, ctg = oracle_lib.read_ann(prefix)
        pes, st = _check(al, got, oracle, l_pac, ctg, 10000, 1000, "c2 pairs")
        assert pes[1]["failed"] == 0 and st["proper"] > 0.9 * (n // 2)
    assert got.n_alns > 0.95 * n
