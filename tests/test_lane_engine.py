"""The batched form's lane extension engine on the CPU (its own code compiled
for the host, tests/lane_check.cpp): random tasks of 1-320 query bases with
substitutions and indels, random gap penalties, narrow and wide bands and
small h0, and targets whose diagonal keeps returning to h0 (the reference's
32-residue rule then gives a cell); the 32-bit engine must equal the
oracle's banded extension, and the packed engine (two strip rows per
register, 16-bit lanes) the 32-bit engine, in score, end positions and cell
count."""
from __future__ import annotations

import subprocess

import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = tmp_path_factory.mktemp("lane") / "lane_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", str(ROOT), "-o", str(out), str(ROOT / "tests" / "lane_check.cpp"),
                    "-x", "c", str(ROOT / "oracle" / "oracle.c"), "-lm"], check=True, capture_output=True)
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_lane_engine_packed_and_oracle(checker, seed):
    r = subprocess.run([str(checker), "20000", str(seed)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    assert "bad 0" in r.stdout, r.stdout
    # both outcomes of the residue rule occurred
    kv = dict(zip(r.stdout.split()[::2], r.stdout.split()[1::2]))
    assert int(kv["h0_cell"]) > 0 and int(kv["h0_none"]) > 0, r.stdout
