"""Batches without host round trips, and as HIP graphs, through the C ABI:
the same alignments, CIGARs, MAPQ, MD and SAM records as the eager launch
sequence, over graph replays with eager work between them, after a change of
batch size (a new capture), in both extension forms, and when every
capacity starts too small (each stage's overflow stops the kernels after it
and the batch runs again with the buffer grown)."""
from __future__ import annotations

import numpy as np
import pytest

import g3sa_amd
from test_zz_no_alloc import _same

pytestmark = pytest.mark.gpu


def _batch(names, seq, off, quals, lo, hi):
    s, o = seq[off[lo]:off[hi]], (off[lo:hi + 1] - off[lo]).astype(np.int32)
    nm, ql = names[lo:hi], quals[lo:hi]
    name_off = np.concatenate([[0], np.cumsum([len(x) for x in nm])]).astype(np.int64)
    qual_off = np.concatenate([[0], np.cumsum([len(x) for x in ql])]).astype(np.int64)
    return s, o, b"".join(nm), name_off, b"".join(ql), qual_off


def _run(al, b):
    al.upload(b[0], b[1])
    al.upload_text(b[2], b[3], b[4], b[5])
    al.run()
    res = al.download()
    al.format(sam=True)
    return res, al.text()


@pytest.mark.parametrize("form", ["batched", "plain"])
def test_graph_matches_eager(ds_rep, form):
    names, seq, off, quals = g3sa_amd.read_fastq(str(ds_rep.reads["main"]))
    n = len(off) - 1
    batches = [(0, 3000), (0, 3000), (3000, 5500), (0, 3000)]
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=n, max_bases=int(off[-1])) as al:
        al.set_form(form)
        refs = {}
        for lo, hi in sorted(set(batches)):
            refs[(lo, hi)] = _run(al, _batch(names, seq, off, quals, lo, hi))
            assert not al.graph_info()["graph_used"]
        al.set_graph(True)
        for k, (lo, hi) in enumerate(batches):
            label = f"{form} graph run {k} reads {lo}:{hi}"
            print(f"[graph] {label}", flush=True)
            got, text = _run(al, _batch(names, seq, off, quals, lo, hi))
            info = al.graph_info()
            assert info["graph_used"], label
            _same(got, refs[(lo, hi)][0], label)
            assert text == refs[(lo, hi)][1], f"{label}: SAM records"
            # eager work on the same stream between replays
            al.set_graph(False)
            again, _ = _run(al, _batch(names, seq, off, quals, lo, hi))
            _same(again, refs[(lo, hi)][0], f"{label}: eager after it")
            al.set_graph(True)
        assert al.graph_info()["captures"] >= 2


@pytest.mark.parametrize("graph", [False, True])
def test_overflowing_capacities_rerun(ds_rep, graph):
    names, seq, off, quals = g3sa_amd.read_fastq(str(ds_rep.reads["main"]))
    n = len(off) - 1
    b = _batch(names, seq, off, quals, 0, 4000)
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=n, max_bases=int(off[-1])) as al:
        ref = _run(al, b)
    with g3sa_amd.Aligner(str(ds_rep.prefix), max_reads=n, max_bases=int(off[-1])) as al:
        al.set_graph(graph)
        al.set_batch_caps(0, 0, 0, 0, 0, 0, 0)
        print(f"[graph] overflow reruns, graph={graph}", flush=True)
        got = _run(al, b)
        info = al.graph_info()
        # seeds, tasks, alignments and the CIGAR and direction buffers all
        # start below what 4000 reads need
        print(f"[graph] {info}", flush=True)
        assert info["overflow_reruns"] >= 4, info
        _same(got[0], ref[0], f"graph={graph}")
        assert got[1] == ref[1], "SAM records"
        # the next batch starts from the grown capacities
        again = _run(al, b)
        assert al.graph_info()["overflow_reruns"] == info["overflow_reruns"]
        _same(again[0], ref[0], f"graph={graph}, second batch")
