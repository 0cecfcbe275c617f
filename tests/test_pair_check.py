"""The exact repeated-add of the insert-size deviation sum (g3a::repeat_add,
g3sa_amd/csrc/g3a_pair.h) equals the loop of single rounded adds it stands
for, on the host (tests/pair_check.cpp)."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_repeat_add_equals_the_add_loop(tmp_path):
    out = tmp_path / "pair_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", str(ROOT), "-o", str(out),
                    str(ROOT / "tests" / "pair_check.cpp")], check=True)
    r = subprocess.run([str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
