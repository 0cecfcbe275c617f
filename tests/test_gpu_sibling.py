"""Sibling contexts (g3a_device_alloc_sibling) through the C ABI: two contexts
sharing one device index align the two halves of a batch at the same time
from two threads, with the same alignments, CIGARs, MAPQ, MD and SAM records
as one context aligning each half; a sibling keeps working after its source
is destroyed; the learned index is shared when built first; calls that would
rebuild a shared index are refused."""
from __future__ import annotations

import threading

import numpy as np
import pytest

import g3sa_amd
from test_gpu_graph import _batch, _run
from test_zz_no_alloc import _same

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("learned", [False, True])
def test_siblings_concurrent_halves(ds_rep, learned):
    names, seq, off, quals = g3sa_amd.read_fastq(str(ds_rep.reads["main"]))
    n = len(off) - 1
    halves = [_batch(names, seq, off, quals, 0, n // 2), _batch(names, seq, off, quals, n // 2, n)]
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=n, max_bases=int(off[-1])) as al:
        if learned:
            al.set_learned(1)
        refs = [_run(al, h) for h in halves]
    src = g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=n, max_bases=int(off[-1]))
    if learned:
        src.set_learned(1)
    sib = src.sibling(n - n // 2 + 16, int(off[-1]))
    out = [None, None]

    def work(k, al):
        out[k] = _run(al, halves[k])

    print(f"[sibling] concurrent halves, learned={learned}", flush=True)
    th = [threading.Thread(target=work, args=(0, src)), threading.Thread(target=work, args=(1, sib))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for k in range(2):
        _same(out[k][0], refs[k][0], f"half {k}")
        assert out[k][1] == refs[k][1], f"half {k}: SAM records"
    # the index outlives its source while a sibling holds it
    rc = src.lib.g3a_memcpy_index(src.ctx)
    assert rc != 0 and b"shared" in src.lib.g3a_last_error(src.ctx)
    src.close()
    print("[sibling] after the source is destroyed", flush=True)
    again = _run(sib, halves[0])
    _same(again[0], refs[0][0], "sibling alone")
    assert again[1] == refs[0][1]
    sib.close()


def test_sibling_refuses_to_build_learned(ds_c1):
    with g3sa_amd.Aligner(str(ds_c1.prefix), max_reads=1000) as src:
        sib = src.sibling(1000)
        info = g3sa_amd.LearnedInfo()
        rc = sib.lib.g3a_set_learned(sib.ctx, 1, info)
        assert rc != 0 and b"sibling" in sib.lib.g3a_last_error(sib.ctx)
        sib.close()
