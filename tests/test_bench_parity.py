"""bench.py's parity sample on the CPU: the oracle's SAM for the first reads of
a FASTQ against a command line's SAM for the same reads.  The command line is
stood in for by the oracle itself (there is no GPU here), so the comparison
must report identical output, and a planted difference must be reported at
its line."""
from __future__ import annotations

import stat

import bench
import g3sa_amd
from conftest import ROOT, gen


def test_parity_sample_plumbing(built, tmp_path, monkeypatch):
    fa, prefix, fq = tmp_path / "r.fa", tmp_path / "r", tmp_path / "r.fq"
    gen("ref", "-o", str(fa), "--len", "400000", "--contigs", "2", "--seed", "3")
    g3sa_amd.build_index(str(fa), str(prefix), 4)
    gen("reads", "-r", str(fa), "-o", str(fq), "--n", "1500", "--len", "150", "--seed", "4")
    # `g3a mem -S -t 16 -o OUT [--learned] PREFIX FQ` answered by the oracle
    stand_in = tmp_path / "mem.sh"
    stand_in.write_text("#!/bin/bash\nout=$6; shift 6; [ \"$1\" = --learned ] && shift\n"
                        f"exec {ROOT / 'oracle' / '_build' / 'g3a_oracle'} -t 2 -m sam -o \"$out\" \"$1\" \"$2\"\n")
    stand_in.chmod(stand_in.stat().st_mode | stat.S_IXUSR)
    monkeypatch.setattr(g3sa_amd, "CLI_PATH", stand_in)

    sam = tmp_path / "oracle.sam"
    base = bench.cpu_baseline(prefix, fq, 1000, sam)
    assert base is not None and base["kind"] == "port" and base["value"] > 0
    assert sam.read_bytes().startswith(b"@HD")
    got = bench.parity_sample(prefix, fq, 1000, sam, True)
    assert got["identical"] and got["reads"] == 1000 and got["lines"] > 1000, got
    assert (tmp_path / "r_head1000.fq").read_text().count("\n") == 4000

    lines = sam.read_bytes().split(b"\n")
    k = next(i for i, ln in enumerate(lines) if ln and not ln.startswith(b"@"))
    lines[k] = lines[k].replace(b"\t", b"\tX", 1)
    sam.write_bytes(b"\n".join(lines))
    got = bench.parity_sample(prefix, fq, 1000, sam, False)
    assert not got["identical"] and got["first_difference"]["line"] == k, got


def test_bench_fastq_parse_matches_reader(built, tmp_path):
    """bench.load_reads (vectorised) against g3sa_amd.read_fastq on names
    with and without comments, N bases and lower case."""
    fa, fq = tmp_path / "r.fa", tmp_path / "r.fq"
    gen("ref", "-o", str(fa), "--len", "100000", "--contigs", "1", "--seed", "5")
    gen("reads", "-r", str(fa), "-o", str(fq), "--n", "300", "--len-min", "70", "--len-max", "250", "--seed", "6")
    recs = fq.read_bytes().split(b"\n")
    recs[0] = b"@plain_name"                    # no comment
    recs[5] = recs[5][:10] + b"nN" + recs[5][12:]  # N bases, lower case
    fq.write_bytes(b"\n".join(recs))
    codes, off, names, name_off, quals, qual_off = bench.load_reads(fq)
    r_names, r_seq, r_off, r_quals = g3sa_amd.read_fastq(str(fq))
    assert list(off) == list(r_off) and bytes(codes) == bytes(r_seq)
    assert [names[name_off[i]:name_off[i + 1]] for i in range(len(r_names))] == r_names
    assert [quals[qual_off[i]:qual_off[i + 1]] for i in range(len(r_quals))] == r_quals
