"""Shared fixtures: built native code, the oracle, and small synthetic datasets.

Datasets are generated with the project's deterministic generator (g3a-gen)
into a per-session temporary directory; nothing here reads the reference
repository, so the same fixtures work on any machine with the built code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))

import g3sa_amd  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (run with -m gpu)")


@pytest.fixture(scope="session")
def built():
    """The native build must exist (tests never build it themselves)."""
    missing = [p for p in (g3sa_amd.LIB_PATH, g3sa_amd.CLI_PATH, g3sa_amd.GEN_PATH,
                           ROOT / "oracle" / "_build" / "liboracle.so",
                           ROOT / "oracle" / "_build" / "g3a_oracle") if not p.exists()]
    if missing:
        pytest.fail(f"native build missing ({missing}); run __graft_entry__.build()")
    return g3sa_amd.BUILD


def gen(*args: str) -> None:
    subprocess.run([str(g3sa_amd.GEN_PATH), *args], check=True)


class Dataset:
    def __init__(self, root: Path, name: str):
        self.root = root
        self.prefix = root / name
        self.fasta = root / f"{name}.fa"
        self.reads = {}


@pytest.fixture(scope="session")
def ds_c1(built, tmp_path_factory):
    """Config 1 exactly: 1 Mbp uniform reference (seed 1), 10,000 SE 150 bp reads
    (seed 2) at 1% substitution and 0.05% indels, default scoring."""
    root = tmp_path_factory.mktemp("c1")
    d = Dataset(root, "c1")
    gen("ref", "-o", str(d.fasta), "--len", "1000000", "--contigs", "1", "--seed", "1")
    g3sa_amd.build_index(str(d.fasta), str(d.prefix), 4)
    fq = root / "c1.fq"
    gen("reads", "-r", str(d.fasta), "-o", str(fq), "--n", "10000", "--len", "150", "--sub", "0.01",
        "--indel", "0.0005", "--seed", "2", "--truth", str(root / "c1.truth"))
    d.reads["main"] = fq
    # config-3 style reads on the same reference: 70-250 bp, hard reads with tails
    fq3 = root / "c3.fq"
    gen("reads", "-r", str(d.fasta), "-o", str(fq3), "--n", "4000", "--len-min", "70", "--len-max", "250",
        "--sub", "0.01", "--indel", "0.0005", "--hard-frac", "0.3", "--hard-div", "0.08", "--seed", "3")
    d.reads["c3"] = fq3
    return d


@pytest.fixture(scope="session")
def ds_rep(built, tmp_path_factory):
    """Repeat-rich multi-contig reference (config-2 model at 8 Mbp) with paired
    reads aligned as single-end."""
    root = tmp_path_factory.mktemp("rep")
    d = Dataset(root, "rep")
    gen("ref", "-o", str(d.fasta), "--len", "8000000", "--contigs", "5", "--seed", "11", "--gc-mean", "0.41",
        "--gc-sd", "0.05", "--repeat-frac", "0.15", "--families", "200", "--fam-min", "300", "--fam-max", "6000",
        "--div-min", "0.05", "--div-max", "0.15", "--tandem-frac", "0.03")
    g3sa_amd.build_index(str(d.fasta), str(d.prefix), 8)
    fq = root / "rep.fq"
    gen("reads", "-r", str(d.fasta), "-o", str(fq), "--n", "12000", "--paired", "--len", "150",
        "--insert-mean", "400", "--insert-sd", "50", "--sub-start", "0.005", "--sub-end", "0.015",
        "--indel", "0.0005", "--seed", "12")
    d.reads["main"] = fq
    return d


@pytest.fixture(scope="session")
def oracle(built):
    import oracle_lib
    return oracle_lib.Oracle()


def oracle_threads() -> int:
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except Exception:
        return 4
