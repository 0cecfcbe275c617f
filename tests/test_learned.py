"""The learned-index form on the CPU, through the kernels' own code compiled
for the host (tests/learned_check.cpp): full SA from sample walks, keys,
model training, exact windowed lookups against std::lower_bound and the
FM-index's intervals, a zeroed model (misses only, never a wrong answer),
and seeding of simulated reads with and without the learned form."""
from __future__ import annotations

import subprocess

import pytest

import g3sa_amd
from conftest import ROOT, gen


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = tmp_path_factory.mktemp("learned") / "learned_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-pthread", "-I", str(ROOT), "-o", str(out),
                    str(ROOT / "tests" / "learned_check.cpp"), str(ROOT / "g3sa_amd" / "csrc" / "index_io.cpp")],
                   check=True, capture_output=True)
    return out


def _run(checker, prefix, reads=1500):
    r = subprocess.run([str(checker), str(prefix), str(reads)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    return {k: float(v) for k, v in (t.split("=") for t in r.stdout.split()[1:])}


def test_learned_uniform(built, checker, tmp_path):
    fa, prefix = tmp_path / "u.fa", tmp_path / "u"
    gen("ref", "-o", str(fa), "--len", "1000000", "--contigs", "1", "--seed", "1")
    g3sa_amd.build_index(str(fa), str(prefix), 8)
    s = _run(checker, prefix)
    assert s["rows"] == 2_000_001
    # the leaves meet the 256-row bound (fit to 250): no window misses
    assert s["max_err"] <= 252 and s["over_window"] == 0 and s["lb_miss"] == 0 and s["range_miss"] == 0
    assert s["lookup_miss"] == 0 and s["lookups"] > 0 and s["text_runs"] > 0
    assert s["steps_learned"] < 0.8 * s["steps_plain"]
    # the zeroed model misses nearly everything, and answered nothing wrong
    assert s["zeroed_lb_miss"] > 0.9 * s["distinct"]


def test_learned_repeats(built, checker, tmp_path):
    fa, prefix = tmp_path / "r.fa", tmp_path / "r"
    gen("ref", "-o", str(fa), "--len", "2100000", "--contigs", "3", "--seed", "5", "--gc-mean", "0.41",
        "--gc-sd", "0.05", "--repeat-frac", "0.2", "--families", "50", "--fam-min", "300", "--fam-max", "6000",
        "--div-min", "0.0", "--div-max", "0.05", "--tandem-frac", "0.03")
    g3sa_amd.build_index(str(fa), str(prefix), 16)
    s = _run(checker, prefix)
    assert s["distinct"] < s["rows"]  # repeats and tandem units share keys
    assert s["lookups"] > 0 and s["text_runs"] > 0
    # clusters of near-equal keys get segments of their own: the bound holds
    assert s["max_err"] <= 252 and s["over_window"] == 0 and s["lb_miss"] == 0 and s["range_miss"] == 0


def test_learned_tiny_text(built, checker, tmp_path):
    """A text of fewer than 32 rows: every suffix is short."""
    fa, prefix = tmp_path / "t.fa", tmp_path / "t"
    fa.write_text(">t\nACGTTGCAAC\n")
    g3sa_amd.build_index(str(fa), str(prefix), 1)
    s = _run(checker, prefix, reads=50)
    assert s["rows"] == 21
