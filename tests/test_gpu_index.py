"""Index construction on the GPU (prefix doubling, BWT and samples in HBM)
against the host SA-IS build: every file byte-identical, on texts chosen to
stress the suffix sort (long repeats, runs of one base, periodic text, tiny
and odd-length contigs, several Occ superblocks), and the reference's E. coli
fixtures (.sa and .counts) built from the fixture's own .pac."""
from __future__ import annotations

import numpy as np
import pytest

import g3sa_amd
from conftest import ROOT, gen

pytestmark = pytest.mark.gpu
GOLD = ROOT / "tests" / "golden" / "ecoli"
EXTS = ["bwt", "sa", "counts", "g3sa", "pac", "ann", "amb"]


def write_fa(path, seqs):
    with open(path, "w") as f:
        for c, s in enumerate(seqs):
            f.write(f">c{c} test\n")
            letters = s if isinstance(s, str) else "".join("ACGT"[x] for x in s)
            for i in range(0, len(letters), 70):
                f.write(letters[i:i + 70] + "\n")


def compare(tmp_path, fa, sa_intv):
    host, dev = tmp_path / "host", tmp_path / "dev"
    g3sa_amd.build_index(str(fa), str(host), sa_intv)
    g3sa_amd.build_index_dev(str(fa), str(dev), sa_intv)
    for ext in EXTS:
        a = (tmp_path / f"host.{ext}").read_bytes()
        b = (tmp_path / f"dev.{ext}").read_bytes()
        assert a == b, f".{ext} differs (host {len(a)} bytes, device {len(b)} bytes)"


def _cases():
    rng = np.random.default_rng(7)
    u = lambda n: rng.integers(0, 4, n)  # noqa: E731
    rep = u(4000)
    a = u(3000)
    a[1000:2200] = rep[:1200]  # an exact 1.2 kbp repeat, also against its own contig
    return {
        "multi_contig_repeat": ([a, u(1501), rep, u(6)], 4),
        "runs_of_one_base": (["A" * 5000 + "C" + "A" * 3000, "T" * 777, u(999)], 8),
        "periodic": (["ACGT" * 2000 + "ACG", "AC" * 1500], 1),
        "tiny_contigs": (["A", "C", "GT", u(5), u(33)], 32),
        "several_superblocks": ([u(1_300_001), u(950_003)], 4),
    }


@pytest.mark.parametrize("case", list(_cases()))
def test_device_index_matches_host(built, tmp_path, case):
    seqs, intv = _cases()[case]
    fa = tmp_path / "ref.fa"
    write_fa(fa, seqs)
    compare(tmp_path, fa, intv)


def test_device_index_repeats_and_tandems(built, tmp_path):
    fa = tmp_path / "ref.fa"
    gen("ref", "-o", str(fa), "--len", "3000000", "--contigs", "3", "--seed", "31", "--gc-mean", "0.41",
        "--gc-sd", "0.05", "--repeat-frac", "0.3", "--families", "40", "--fam-min", "300", "--fam-max", "6000",
        "--div-min", "0.0", "--div-max", "0.05", "--tandem-frac", "0.05")
    compare(tmp_path, fa, 4)


def test_device_index_ecoli_fixture(built, tmp_path):
    ann = (GOLD / "ecoli.ann").read_text().split("\n")
    l_pac = int(ann[0].split()[0])
    _, name, anno = ann[1].split(" ", 2)
    raw = np.frombuffer((GOLD / "ecoli.pac").read_bytes(), dtype=np.uint8)
    i = np.arange(l_pac)
    seq = (raw[i >> 2] >> ((3 - (i & 3)) * 2)) & 3
    fa = tmp_path / "ecoli.fa"
    with open(fa, "w") as f:
        f.write(f">{name} {anno}\n")
        letters = np.frombuffer(b"ACGT", dtype=np.uint8)[seq].tobytes().decode()
        for k in range(0, len(letters), 60):
            f.write(letters[k:k + 60] + "\n")
    g3sa_amd.build_index_dev(str(fa), str(tmp_path / "ecoli"), 4)
    for ext in ["pac", "ann", "amb", "sa", "counts"]:
        assert (tmp_path / f"ecoli.{ext}").read_bytes() == (GOLD / f"ecoli.{ext}").read_bytes(), ext
