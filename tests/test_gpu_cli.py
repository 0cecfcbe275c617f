"""The `g3a mem` command line (the reference's `bwa mem` entry) against the
oracle's command line: byte-identical output in the reference's record format
and in standard SAM, across batch boundaries and both extension forms."""
from __future__ import annotations

import os
import subprocess

import pytest

import g3sa_amd
from conftest import ROOT, oracle_threads

pytestmark = pytest.mark.gpu

ORACLE_CLI = ROOT / "oracle" / "_build" / "g3a_oracle"


def oracle_out(ds, fq, mode, tmp_path):
    out = tmp_path / f"oracle_{mode}.txt"
    subprocess.run([str(ORACLE_CLI), "-t", str(oracle_threads()), "-m", mode, "-o", str(out), str(ds.prefix),
                    str(fq)], check=True, capture_output=True)
    return out.read_bytes()


def gpu_out(ds, fq, extra, tmp_path, name, chunk=None):
    out = tmp_path / f"gpu_{name}.txt"
    env = dict(os.environ)
    if chunk:
        env["G3A_FASTQ_CHUNK"] = str(chunk)
    r = subprocess.run([str(g3sa_amd.CLI_PATH), "mem", "-o", str(out), *extra, str(ds.prefix), str(fq)],
                       capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr
    return out.read_bytes()


def first_diff(a: bytes, b: bytes) -> str:
    la, lb = a.split(b"\n"), b.split(b"\n")
    for i, (x, y) in enumerate(zip(la, lb)):
        if x != y:
            return f"line {i}: gpu {x[:200]!r} oracle {y[:200]!r}"
    return f"line counts {len(la)} vs {len(lb)}"


@pytest.mark.parametrize("mode", ["compat", "sam"])
def test_cli_byte_identical(ds_c1, tmp_path, mode):
    fq = ds_c1.reads["c3"]
    ref = oracle_out(ds_c1, fq, mode, tmp_path)
    flags = ["-S"] if mode == "sam" else []
    print(f"[cli] {mode}: one batch", flush=True)
    got = gpu_out(ds_c1, fq, flags, tmp_path, f"{mode}_one")
    assert got == ref, first_diff(got, ref)
    print(f"[cli] {mode}: batches of 700 reads", flush=True)
    got = gpu_out(ds_c1, fq, flags + ["-Z", "700"], tmp_path, f"{mode}_many")
    assert got == ref, first_diff(got, ref)
    print(f"[cli] {mode}: batches of 700 reads, launches without graphs", flush=True)
    got = gpu_out(ds_c1, fq, flags + ["-Z", "700", "--no-graph"], tmp_path, f"{mode}_many_eager")
    assert got == ref, first_diff(got, ref)
    # FASTQ read in 997-byte chunks: records cut at every possible offset
    print(f"[cli] {mode}: 997-byte input chunks", flush=True)
    got = gpu_out(ds_c1, fq, flags + ["-Z", "1500"], tmp_path, f"{mode}_chunks", chunk=997)
    assert got == ref, first_diff(got, ref)
    print(f"[cli] {mode}: learned-index form", flush=True)
    got = gpu_out(ds_c1, fq, flags + ["-Z", "1500", "--learned"], tmp_path, f"{mode}_learned")
    assert got == ref, first_diff(got, ref)


def test_cli_repeats_sam(ds_rep, tmp_path):
    fq = ds_rep.reads["main"]
    ref = oracle_out(ds_rep, fq, "sam", tmp_path)
    print("[cli] repeats: sam", flush=True)
    got = gpu_out(ds_rep, fq, ["-S", "-Z", "5000"], tmp_path, "rep_sam")
    assert got == ref, first_diff(got, ref)
    print("[cli] repeats: sam, learned-index form", flush=True)
    got_l = gpu_out(ds_rep, fq, ["-S", "-Z", "5000", "--learned"], tmp_path, "rep_sam_learned")
    assert got_l == ref, first_diff(got_l, ref)
    lines = [ln for ln in got.split(b"\n") if ln and not ln.startswith(b"@")]
    assert len(lines) >= 12000


def test_cli_long_read_alone_in_batch(ds_c1, tmp_path):
    """A read over 320 bases with -Z 1 (and in a batch of 3): the batch is not
    refused, and the read comes out unmapped as the oracle writes it."""
    src = ds_c1.reads["main"].read_bytes().split(b"\n")
    recs = [src[i:i + 4] for i in range(0, 24, 4)]
    long_seq = recs[1][1] + recs[2][1] + recs[3][1][:100]
    recs[2] = [b"@long_read x", long_seq, b"+", b"I" * len(long_seq)]
    fq = tmp_path / "long.fq"
    fq.write_bytes(b"".join(b"\n".join(r) + b"\n" for r in recs))
    for mode, flags in [("compat", []), ("sam", ["-S"])]:
        ref = oracle_out(ds_c1, fq, mode, tmp_path)
        for z in ("1", "3"):
            print(f"[cli] long read: {mode} -Z {z}", flush=True)
            got = gpu_out(ds_c1, fq, flags + ["-Z", z], tmp_path, f"long_{mode}_{z}")
            assert got == ref, first_diff(got, ref)
