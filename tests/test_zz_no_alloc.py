"""The allocation-free entry point (g3a_bwa_align_noalloc) against
g3a_bwa_align on the same batches through the C ABI: every alignment field,
both CIGARs, MAPQ, MD and the device-formatted SAM records identical, on two
calls in a row (the second with a smaller batch), on a repeat-rich text in
both extension forms; batches over its read limit are refused."""
from __future__ import annotations

import numpy as np
import pytest

import g3sa_amd
from test_gpu_parity import FIELDS

pytestmark = pytest.mark.gpu


def _same(a: g3sa_amd.Result, b: g3sa_amd.Result, label: str) -> None:
    np.testing.assert_array_equal(a.aln_off, b.aln_off, err_msg=label)
    for f, _ in FIELDS:
        np.testing.assert_array_equal(getattr(a, f), getattr(b, f), err_msg=f"{label}: {f}")
    for i in range(a.n_alns):
        ka = a.cigar[a.cigar_off[i]: a.cigar_off[i] + a.n_cigar[i] + a.n_core[i]]
        kb = b.cigar[b.cigar_off[i]: b.cigar_off[i] + b.n_cigar[i] + b.n_core[i]]
        assert np.array_equal(ka, kb), f"{label}: CIGARs of alignment {i}"
    np.testing.assert_array_equal(a.mapq, b.mapq, err_msg=f"{label}: mapq")
    assert a.md == b.md, f"{label}: MD"


def test_noalloc_matches_bwa_align(ds_rep):
    names, seq, off, quals = g3sa_amd.read_fastq(str(ds_rep.reads["main"]))
    batches = [(0, 3000), (5000, 5700)]  # the second call smaller than the first
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=len(off) - 1, max_bases=int(off[-1])) as al:
        for form in ("batched", "plain"):
            al.set_form(form)
            for lo, hi in batches:
                s, o = seq[off[lo]:off[hi]], (off[lo:hi + 1] - off[lo]).astype(np.int32)
                nm, ql = names[lo:hi], quals[lo:hi]
                name_off = np.concatenate([[0], np.cumsum([len(x) for x in nm])]).astype(np.int64)
                qual_off = np.concatenate([[0], np.cumsum([len(x) for x in ql])]).astype(np.int64)
                label = f"{form} reads {lo}:{hi}"
                print(f"[noalloc] {label}", flush=True)
                al.upload(s, o)
                al.upload_text(b"".join(nm), name_off, b"".join(ql), qual_off)
                al.run()
                ref = al.download()
                al.format(sam=True)
                ref_text = al.text()
                al.run_noalloc()
                got = al.download()
                _same(got, ref, label)
                assert got.n_alns > 0.9 * (hi - lo), label
                al.format(sam=True)
                assert al.text() == ref_text, f"{label}: SAM records"


def test_noalloc_refuses_batches_over_its_limit(ds_c1):
    names, seq, off, _ = g3sa_amd.read_fastq(str(ds_c1.reads["main"]))
    n = 4097
    assert len(off) - 1 >= n
    with g3sa_amd.Aligner(str(ds_c1.prefix), max_reads=n, max_bases=int(off[n])) as al:
        al.upload(seq[:off[n]], off[:n + 1])
        rc = al.lib.g3a_bwa_align_noalloc(al.ctx)
        assert rc == -1
        assert b"4096" in al.lib.g3a_last_error(al.ctx)
        # a batch within the limit still runs on the same context
        al.upload(seq[:off[100]], off[:101])
        al.run_noalloc()
        assert al.download().n_alns > 90
