"""The device Occ layouts against brute force: every BWT row of an index
whose text spans several 2^21-row superblocks (so the narrow form's relative
counts, derived base-3 counts and '$' discount all cross superblock edges).
The lookup code is the kernels' own, compiled for the host."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

import g3sa_amd
from conftest import ROOT


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = tmp_path_factory.mktemp("layout") / "layout_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT), "-o", str(out), str(ROOT / "tests" / "layout_check.cpp"),
                    str(ROOT / "g3sa_amd" / "csrc" / "index_io.cpp")], check=True, capture_output=True)
    return out


# (contig lengths, base composition): uniform texts over several superblocks
# and a tiny one; an A-rich text, whose first superblock holds more than 2^20
# A rows (the relative counts need all 21 bits)
@pytest.mark.parametrize("lens,comp", [((1_300_000, 950_000), None), ((3000, 700), None),
                                       ((1_300_000,), (0.94, 0.02, 0.02, 0.02))])
def test_occ_layouts_every_row(built, checker, tmp_path, lens, comp):
    rng = np.random.default_rng(sum(lens))
    fa = tmp_path / "ref.fa"
    with open(fa, "w") as f:
        for c, L in enumerate(lens):
            s = rng.choice(4, L, p=comp) if comp else rng.integers(0, 4, L)
            f.write(f">c{c}\n")
            for i in range(0, L, 60):
                f.write("".join("ACGT"[x] for x in s[i:i + 60]) + "\n")
    prefix = tmp_path / "ref"
    g3sa_amd.build_index(str(fa), str(prefix), 8)
    r = subprocess.run([str(checker), str(prefix)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    rows, n_sb = (int(v) for v in r.stdout.split()[1:3])
    assert rows == 2 * sum(lens) + 1
    assert n_sb == (rows >> 21) + 1
    if len(lens) > 1:
        assert n_sb >= 3 or sum(lens) < 1_000_000  # the rows cross two superblock edges
