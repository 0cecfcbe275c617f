"""Mutation checks of the oracle pins (CPU only).

Each case copies the oracle's C sources to a temporary directory, changes one
token in the function a pin covers, builds the copy with gcc and runs that
pin against it: the pin must fail on every mutant (and passes on the real
oracle in tests/test_oracle_pins.py).  Nothing here touches the product
library or the GPU.
"""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

import oracle_lib
from test_oracle_pairs import check_pairs
from test_oracle_pins import (chain_idx, check_chain, check_emitters, check_global, check_quirk,  # noqa: F401
                              check_reseed, emitter_case, planted_reseed)

SRC = oracle_lib.ROOT / "oracle"

# (name, file, original text, mutated text, pin)
MUTANTS = [
    # ksw_global2 (bwa/src/ksw.cu:706-812)
    ("global: E wins M ties", "oracle.c", "d = m >= e ? 0 : 1;", "d = m > e ? 0 : 1;", "global"),
    ("global: F wins ties", "oracle.c", "d = h >= f ? d : 2;", "d = h > f ? d : 2;", "global"),
    ("global: E extends on ties", "oracle.c", "d |= e > t ? 1 << 2 : 0;", "d |= e >= t ? 1 << 2 : 0;", "global"),
    ("global: first row to w-1", "oracle.c", "for (j = 1; j <= qlen && j <= w; ++j) eh_h[j]",
     "for (j = 1; j <= qlen && j < w; ++j) eh_h[j]", "global"),
    ("global: band one column short", "oracle.c", "end = i + w + 1 < qlen ? i + w + 1 : qlen;",
     "end = i + w < qlen ? i + w : qlen;", "global"),
    # klib B-tree as the chaining uses it (bwa/src/kbtree_CUDA.cu:22-119)
    ("chain: equal key not matched", "oracle.c",
     "if ((*rr = chain_cmp(kpos, b->chains[x->key[begin]].pos)) < 0) --begin;",
     "if ((*rr = chain_cmp(kpos, b->chains[x->key[begin]].pos)) <= 0) --begin;", "chain"),
    ("chain: upper bound search", "oracle.c", "if (chain_cmp(b->chains[x->key[mid]].pos, kpos) < 0) begin = mid + 1;",
     "if (chain_cmp(b->chains[x->key[mid]].pos, kpos) <= 0) begin = mid + 1;", "chain"),
    ("chain: no early return on equal key", "oracle.c", "if (i >= 0 && r == 0) return x->key[i];",
     "if (0) return x->key[i];", "chain"),
    # reseedV2 (bwa/src/seed.cu:639-913)
    ("reseed: pivot limit 95", "oracle.c", "if (pivot > 77) continue;", "if (pivot > 95) continue;", "reseed"),
    ("reseed: no collected flag", "oracle.c", "} else if (!collected && (int)ILEN(olep) >= min_seed_len) {",
     "} else if ((int)ILEN(olep) >= min_seed_len) {", "reseed"),
    ("reseed: no append after the loop", "oracle.c",
     "if (leps.n > 0 && (int)ILEN(leps.a[0]) >= min_seed_len) iv_push(&seeds, leps.a[0]);",
     "if (0) iv_push(&seeds, leps.a[0]);", "reseed"),
    # ksw_extend_warp2 quirk rows (bwa/src/ksw.cu:924-1003)
    ("extend: no quirk rows", "oracle.c", "} else if ((i & 31) == 0 && j > i + w + 1) {", "} else if (0) {",
     "quirk"),
    ("extend: quirk reads the next first-row value", "oracle.c",
     "} else if ((i & 31) == 0 && j > i + w + 1) {\n        hdiag = MAX2(h0 - oe_ins - (j - 1) * e_ins, 0);",
     "} else if ((i & 31) == 0 && j > i + w + 1) {\n        hdiag = MAX2(h0 - oe_ins - j * e_ins, 0);", "quirk"),
    # finalize, MAPQ and the emitters (bwa/src/aux.cu:2952-3059, :676-700; pipeline.cpp:287-351)
    ("finalize: clips not in genome order", "oracle.c",
     "int clip5 = is_rev ? len - g->qe : g->qb, clip3 = is_rev ? g->qb : len - g->qe;",
     "int clip5 = g->qb, clip3 = len - g->qe;", "emit"),
    ("mapq: identity squared once", "oracle.c", "tmp *= identity * identity;", "tmp *= identity;", "emit"),
    ("sam: 0-based POS", "oracle_main.c", "sputi(&s, fpos - contig_off + 1);", "sputi(&s, fpos - contig_off);",
     "emit"),
    ("sam: read-order clips on reverse hits", "oracle_main.c",
     "int clip_l = rev ? len - al->qe : al->qb, clip_r = rev ? al->qb : len - al->qe;",
     "int clip_l = al->qb, clip_r = len - al->qe;", "emit"),
    ("compat: unmapped position -1", "oracle_main.c", 'sputs(&s, " * 0 * ");', 'sputs(&s, " * -1 * ");', "emit"),
    # paired-end primitives (bwa/src/aux.cu:411-500, 1000-1070)
    ("pairs: orientation code", "oracle.c", "return (r1 == r2 ? 0 : 1) ^ (p2 > b1 ? 0 : 3);",
     "return (r1 == r2 ? 0 : 1) ^ (p2 > b1 ? 0 : 2);", "pairs"),
    ("pairs: strict overlap in cal_sub", "oracle.c", "if (e_min - b_max >= min_l * opt->mask_level) break;",
     "if (e_min - b_max > min_l * opt->mask_level) break;", "pairs"),
    ("pairs: outlier bound 1.5", "oracle.c", "#define ORC_OUTLIER_BOUND 2.0", "#define ORC_OUTLIER_BOUND 1.5",
     "pairs"),
    ("pairs: quartile index not rounded", "oracle.c", "p75 = (int)q[(int)(.75 * qn + .499)];",
     "p75 = (int)q[(int)(.75 * qn)];", "pairs"),
    ("pairs: sample deviation", "oracle.c", "r->std = sqrt(r->std / x);", "r->std = sqrt(r->std / (x - 1));",
     "pairs"),
    ("pairs: direction ratio 0.1", "oracle.c", "#define ORC_MIN_DIR_RATIO 0.05", "#define ORC_MIN_DIR_RATIO 0.1",
     "pairs"),
    ("pairs: pair score penalty sign", "oracle.c", "+ .721 * log(2. * erfc(", "- .721 * log(2. * erfc(", "pairs"),
    ("pairs: best pair counted as near-best", "oracle.c", "for (i = (int)nu - 2, *n_sub = 0; i >= 0; --i)",
     "for (i = (int)nu - 1, *n_sub = 0; i >= 0; --i)", "pairs"),
]


def build_mutant(tmp: Path, fname: str, old: str, new: str) -> tuple[Path, Path]:
    for f in ("oracle.c", "oracle.h", "oracle_main.c"):
        shutil.copy(SRC / f, tmp / f)
    text = (tmp / fname).read_text()
    assert text.count(old) == 1, f"mutation site not unique in {fname}: {old!r}"
    (tmp / fname).write_text(text.replace(old, new))
    flags = ["-O1", "-std=gnu11", "-fPIC", "-pthread", "-ffp-contract=off", "-w"]
    lib, cli = tmp / "liboracle.so", tmp / "g3a_oracle"
    subprocess.run(["gcc", *flags, "-shared", "-o", str(lib), str(tmp / "oracle.c"), "-lm"], check=True)
    subprocess.run(["gcc", *flags, "-o", str(cli), str(tmp / "oracle_main.c"), str(tmp / "oracle.c"), "-lm"],
                   check=True)
    return lib, cli


@pytest.mark.parametrize("name,fname,old,new,pin", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pin_catches_mutant(name, fname, old, new, pin, tmp_path, chain_idx, planted_reseed,  # noqa: F811
                            emitter_case):  # noqa: F811
    lib, cli = build_mutant(tmp_path, fname, old, new)
    orc = oracle_lib.Oracle(lib)
    with pytest.raises(AssertionError):
        if pin == "global":
            check_global(orc)
        elif pin == "chain":
            idx = orc.load(chain_idx)
            try:
                check_chain(orc, idx)
            finally:
                orc.lib.orc_index_free(idx)
        elif pin == "reseed":
            check_reseed(orc, planted_reseed)
        elif pin == "quirk":
            check_quirk(orc)
        elif pin == "pairs":
            check_pairs(orc)
        else:
            check_emitters(cli, emitter_case, tmp_path)
