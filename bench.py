#!/usr/bin/env python3
"""Benchmark: reads/s of the BWA-MEM hot path on one MI355X, SAM included.

Workload (default, the metric's configuration): a seeded 1 Gbp reference in
24 contigs (GC per 10 kb window ~ N(0.41, 0.05), 20% interspersed repeats
from 200 families of 300 bp-6 kbp at 5-15% divergence) and 5M pairs of
2x150 bp reads at 1% error aligned as 10M single-end reads with BWA default
scoring.  Everything is generated deterministically on the machine
(synthetic data); the index is built on the GPU (`g3a index -d 0`).

One step = one batch of 10M reads, resident in HBM as 2-bit codes, names and
base qualities, taken through the whole device pipeline (seeding, SA lookup,
seed sort, chaining + chain filter, banded extension, region filter, global
alignment + finalize) and formatted on the device into SAM records with
MAPQ, NM, MD and AS (the last SAM byte in HBM).  `value` is reads per second
of such steps.  The same steps through PCIe (reads in from the host, SAM out
to the host) are reported as `pcie_inclusive`; the `g3a mem` command line over
the FASTQ file as `end_to_end`.  The CPU baseline is the oracle (C, up to 16
host threads) on a bounded sample of the same reads, alignment plus record
formatting; its SAM for that sample is compared byte for byte with `g3a mem
-S` on the same reads at the workload's full index size (`parity_sample`).

  python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--config c4|c2|c3|c1]
                  [--form all|both|batched|plain|learned]

Forms (identical output): plain (one wavefront per extension task), batched
(length-binned persistent lane engine) and learned (batched plus the
learned-index seeding).  `value` is the learned form's rate when it ran (the
metric's batched + learned form), else the batched form's; the others are
reported beside it on the same definition.  The learned index is built once
per context on the device and, like the index upload, outside the timed
steps.
"""
from __future__ import annotations

import argparse
import threading
import json
import os
import re
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import g3sa_amd  # noqa: E402

ORACLE_CLI = ROOT / "oracle" / "_build" / "g3a_oracle"
HBM_PEAK_GBS = 8000.0
# k_seed's memory-side bytes per launch, from two rocprofv3 --pmc passes
# (FETCH_SIZE, WRITE_SIZE), each scaled by the same counter read on a plain
# copy kernel of known size; one file per form and workload
TRAFFIC_FILES = {("c2", "batched"): "k_seed_traffic_100Mbp_2Mx150.txt",
                 ("c2", "learned"): "k_seed_learned_traffic_100Mbp_2Mx150.txt",
                 ("c4", "batched"): "k_seed_traffic_1Gbp_10Mx150.txt",
                 ("c4", "learned"): "k_seed_learned_traffic_1Gbp_10Mx150.txt"}

CONFIGS = {
    "c2": dict(
        workload="bwa-mem SE, 100 Mbp 5-contig synthetic reference (15% repeats), 1M pairs 2x150 bp as 2M reads",
        ref=["--len", "100000000", "--contigs", "5", "--seed", "11", "--gc-mean", "0.41", "--gc-sd", "0.05",
             "--repeat-frac", "0.15", "--families", "200", "--fam-min", "300", "--fam-max", "6000",
             "--div-min", "0.05", "--div-max", "0.15"],
        reads=["--n", "2000000", "--paired", "--len", "150", "--insert-mean", "400", "--insert-sd", "50",
               "--sub-start", "0.005", "--sub-end", "0.015", "--indel", "0.0005", "--seed", "12"],
        ref_len=100_000_000, n_reads=2_000_000),
    "c4": dict(
        workload="bwa-mem SE, 1 Gbp 24-contig synthetic reference (20% repeats), 5M pairs 2x150 bp as 10M reads",
        ref=["--len", "1000000008", "--contigs", "24", "--seed", "21", "--gc-mean", "0.41", "--gc-sd", "0.05",
             "--repeat-frac", "0.2", "--families", "200", "--fam-min", "300", "--fam-max", "6000",
             "--div-min", "0.05", "--div-max", "0.15"],
        reads=["--n", "10000000", "--paired", "--len", "150", "--insert-mean", "400", "--insert-sd", "50",
               "--sub", "0.01", "--indel", "0.0005", "--seed", "22"],
        ref_len=1_000_000_008, n_reads=10_000_000),
    "c3": dict(
        workload="bwa-mem SE, 1 Mbp uniform synthetic reference, 2M reads of 70-250 bp (5% at 8% divergence "
                 "with a 20-40 bp random tail)",
        ref=["--len", "1000000", "--contigs", "1", "--seed", "1"],
        reads=["--n", "2000000", "--len-min", "70", "--len-max", "250", "--sub", "0.01", "--indel", "0.0005",
               "--hard-frac", "0.05", "--hard-div", "0.08", "--tail-min", "20", "--tail-max", "40", "--seed", "3"],
        ref_len=1_000_000, n_reads=2_000_000),
    "c1": dict(
        workload="bwa-mem SE, 1 Mbp uniform synthetic reference, 10k x 150 bp reads",
        ref=["--len", "1000000", "--contigs", "1", "--seed", "1"],
        reads=["--n", "10000", "--len", "150", "--sub", "0.01", "--indel", "0.0005", "--seed", "2"],
        ref_len=1_000_000, n_reads=10_000),
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def prepare(cfg_name: str, data_dir: Path) -> tuple[Path, Path, dict]:
    cfg = CONFIGS[cfg_name]
    data_dir.mkdir(parents=True, exist_ok=True)
    fa, prefix, fq = data_dir / f"{cfg_name}.fa", data_dir / cfg_name, data_dir / f"{cfg_name}.fq"
    info = {}
    if not fa.exists():
        log(f"generating reference ({cfg['ref_len']} bp)")
        t0 = time.time()
        subprocess.run([str(g3sa_amd.GEN_PATH), "ref", "-o", str(fa), *cfg["ref"]], check=True)
        info["reference_gen_s"] = round(time.time() - t0, 1)
    if not (data_dir / f"{cfg_name}.g3sa").exists():
        log("building index on the GPU (g3a index -d 0)")
        t0 = time.time()
        subprocess.run([str(g3sa_amd.CLI_PATH), "index", "-d", "0", "-s", "4", str(fa), str(prefix)], check=True)
        info["index_build_s"] = round(time.time() - t0, 1)
        log(f"index built in {info['index_build_s']} s")
    if not fq.exists():
        log(f"generating {cfg['n_reads']} reads")
        t0 = time.time()
        subprocess.run([str(g3sa_amd.GEN_PATH), "reads", "-r", str(fa), "-o", str(fq), *cfg["reads"]], check=True)
        info["reads_gen_s"] = round(time.time() - t0, 1)
    return prefix, fq, info


def _ranges(buf: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """The bytes buf[lo[i]:hi[i]] concatenated, and their int64 offsets."""
    lens = (hi - lo).astype(np.int64)
    off = np.zeros(len(lens) + 1, dtype=np.int64)
    np.cumsum(lens, out=off[1:])
    idx = np.repeat(lo.astype(np.int64) - off[:-1], lens)
    idx += np.arange(off[-1], dtype=np.int64)
    return buf[idx], off


def load_reads(fq: Path):
    """Vectorised FASTQ parse (4-line records) with the reference's rules:
    2-bit codes and int32 offsets; names (up to the first space) and
    qualities with int64 offsets."""
    buf = np.fromfile(fq, dtype=np.uint8)
    nl = np.flatnonzero(buf == 10)
    starts = np.concatenate([[0], nl[:-1] + 1])
    h_lo, h_hi = starts[0::4] + 1, nl[0::4]
    sp = np.flatnonzero(buf == 32)
    k = np.searchsorted(sp, h_lo)
    cand = np.where(k < len(sp), sp[np.minimum(k, max(len(sp) - 1, 0))] if len(sp) else h_hi, h_hi)
    n_hi = np.minimum(cand, h_hi)
    names, name_off = _ranges(buf, h_lo, n_hi)
    quals, qual_off = _ranges(buf, starts[3::4], nl[3::4])
    codes, off = _ranges(buf, starts[1::4], nl[1::4])
    table = np.full(256, 4, dtype=np.uint8)
    for i, ch in enumerate(b"ACGT"):
        table[ch] = i
        table[ch + 32] = i
    return table[codes], off.astype(np.int32), names.tobytes(), name_off, quals.tobytes(), qual_off


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prefix: Path, fq: Path, sample: int, sam_out: Path | None = None) -> dict | None:
    """The oracle on the first `sample` reads: alignment on up to 16 host
    threads (those this process may use), then SAM record formatting; writing
    its SAM to `sam_out` (for the parity sample) is not timed."""
    if not ORACLE_CLI.exists():
        return None
    try:
        cores = len(os.sched_getaffinity(0))
    except Exception:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    log(f"CPU baseline: oracle on {sample} reads, {cores} threads")
    r = subprocess.run([str(ORACLE_CLI), "-t", str(cores), "-n", str(sample), "-m", "sam", "-o",
                        str(sam_out) if sam_out else "/dev/null", str(prefix), str(fq)], capture_output=True, text=True)
    m = re.search(r"oracle_reads_per_s ([0-9.]+) threads (\d+) reads (\d+) seconds ([0-9.]+)", r.stderr)
    f = re.search(r"oracle_format_seconds ([0-9.]+)", r.stderr)
    if r.returncode != 0 or not m or not f:
        log(f"oracle failed: {r.stderr[-500:]}")
        return None
    reads, t_aln, t_fmt = int(m.group(3)), float(m.group(4)), float(f.group(1))
    return {"value": round(reads / (t_aln + t_fmt), 1), "unit": "reads/s", "cores": int(m.group(2)), "kind": "port",
            "cpu": cpu_model(),
            "sample": f"first {reads} reads of the same FASTQ: alignment on {m.group(2)} threads ({t_aln:.1f} s) "
                      f"+ SAM record formatting on one ({t_fmt:.2f} s)"}


def parity_sample(prefix: Path, fq: Path, sample: int, oracle_sam: Path, learned: bool) -> dict:
    """Byte comparison, at the workload's full index size, of `g3a mem -S` on
    the first `sample` reads against the oracle's SAM for the same reads."""
    head = fq.with_name(f"{fq.stem}_head{sample}.fq")
    if not head.exists():
        with open(fq, "rb") as src, open(head, "wb") as dst:
            for i, line in enumerate(src):
                if i >= 4 * sample:
                    break
                dst.write(line)
    got = oracle_sam.with_name(oracle_sam.stem + "_g3a.sam")
    cmd = [str(g3sa_amd.CLI_PATH), "mem", "-S", "-t", "16", "-o", str(got)] + (["--learned"] if learned else [])
    r = subprocess.run(cmd + [str(prefix), str(head)], capture_output=True, text=True)
    out = {"reads": sample, "what": "g3a mem -S" + (" --learned" if learned else "") + " vs the oracle's SAM"}
    if r.returncode != 0:
        log(f"g3a mem failed on the parity sample: {r.stderr[-300:]}")
        return {**out, "identical": False, "error": "g3a mem failed"}
    a, b = got.read_bytes(), oracle_sam.read_bytes()
    out.update(identical=a == b, bytes=len(b), lines=b.count(b"\n"))
    if a != b:
        la, lb = a.split(b"\n"), b.split(b"\n")
        i = next((k for k in range(min(len(la), len(lb))) if la[k] != lb[k]), min(len(la), len(lb)))
        out["first_difference"] = {"line": i, "g3a": la[i][:200].decode(errors="replace") if i < len(la) else None,
                                   "oracle": lb[i][:200].decode(errors="replace") if i < len(lb) else None}
        log(f"PARITY SAMPLE DIFFERS at line {i}")
    got.unlink()
    return out


def run_form(al, n: int, steps: int, warmup: int, form: str, autotune: bool = True) -> dict:
    """Steps of align + device SAM formatting on the resident batch."""
    for i in range(warmup):
        al.run()
        al.format(sam=True)
        log(f"[{form}] warmup step {i + 1}/{warmup} done")
    tuned = None
    if form in ("batched", "learned") and autotune:
        if warmup == 0:
            al.run()
        tuned = al.autotune()
        log(f"[{form}] autotuned: {tuned}")
    stage_ms, kern_ms, fmt_s = [], [], []
    st, text_bytes = None, 0
    t0 = time.perf_counter()
    for i in range(steps):
        al.run()  # returns after the stream is synchronised
        tf = time.perf_counter()
        text_bytes = al.format(sam=True)  # records in HBM; returns after the stream is synchronised
        fmt_s.append(time.perf_counter() - tf)
        st = al.stats()
        stage_ms.append(st["ms"])
        kern_ms.append(st["kernel_ms"])
        log(f"[{form}] step {i + 1}/{steps}: {st['ms']['total']:.1f} ms device pipeline, "
            f"{1e3 * fmt_s[-1]:.1f} ms SAM formatting")
    elapsed = time.perf_counter() - t0
    dev_ms = float(np.mean([s["total"] for s in stage_ms]))
    return {
        "reads_per_s": round(n * steps / elapsed, 1),
        "ms_per_step": round(1000.0 * elapsed / steps, 3),
        "device_pipeline_reads_per_s": round(n / (dev_ms * 1e-3), 1),
        "sam_format_ms": round(1e3 * float(np.mean(fmt_s)), 3),
        "sam_bytes": text_bytes,
        "stages_ms": {k: round(float(np.mean([s[k] for s in stage_ms])), 3) for k in stage_ms[0]},
        "kernel_ms": {k: round(float(np.mean([s[k] for s in kern_ms])), 3) for k in kern_ms[0]},
        "work": st["work"],
        "tuning": tuned,
    }


def paired_end_times(al, n: int, reps: int = 3) -> dict:
    """Insert-size statistics and best-pair choice over the batch the context
    aligned last (its reads are pairs 2p, 2p+1), after the timed steps: wall
    time of each call, which ends with its results on the host (4 orientation
    records; six int32 per pair)."""
    if n % 2:
        return None
    ts, tp = [], []
    pes, rows = None, None
    for _ in range(reps):
        t = time.perf_counter()
        pes = al.mem_pestat(10000)
        ts.append(time.perf_counter() - t)
        t = time.perf_counter()
        rows = al.mem_pair(pes)
        tp.append(time.perf_counter() - t)
    fr = pes[1]
    return {"pairs": n // 2, "pestat_ms": round(1e3 * min(ts), 3), "pair_ms": round(1e3 * min(tp), 3),
            "pair_rows_bytes": 24 * (n // 2),
            "fr": {"avg": round(fr["avg"], 3), "std": round(fr["std"], 3), "low": fr["low"], "high": fr["high"]},
            "failed": [p["failed"] for p in pes],
            "paired_frac": round(float((rows[:, 0] > 0).mean()), 4),
            "proper_frac": round(float((rows[:, 5] > 0).mean()), 4)}


def _part(batch, a: int, b: int):
    """Reads [a, b) of a loaded batch, with offsets starting at 0."""
    seq, off, names, name_off, quals, qual_off = batch
    return (seq[int(off[a]):int(off[b])], (off[a:b + 1] - off[a]).astype(np.int32),
            names[int(name_off[a]):int(name_off[b])], name_off[a:b + 1] - name_off[a],
            quals[int(qual_off[a]):int(qual_off[b])], qual_off[a:b + 1] - qual_off[a])


def two_context_steps(al, batch, steps: int, warmup: int) -> dict:
    """The step with the batch split in halves over `al` and a sibling context
    that shares its device index (g3a_device_alloc_sibling), each half
    aligned and formatted from its own host thread at the same time; the
    resident batch of `al` is restored afterwards."""
    off = batch[1]
    n = len(off) - 1
    h = n // 2
    parts = [_part(batch, 0, h), _part(batch, h, n)]
    sib = al.sibling(n - h + 16, int(off[n] - off[h]) + 1024)
    ctxs = [al, sib]
    for c, p in zip(ctxs, parts):
        c.upload(p[0], p[1])
        c.upload_text(*p[2:])

    def both():
        def one(c):
            c.run()
            c.format(sam=True)
        th = [threading.Thread(target=one, args=(c,)) for c in ctxs]
        for t in th:
            t.start()
        for t in th:
            t.join()

    for _ in range(max(warmup, 1)):
        both()
    t0 = time.perf_counter()
    for i in range(steps):
        both()
        log(f"two contexts: step {i + 1}/{steps} done")
    el = time.perf_counter() - t0
    sib.close()
    al.upload(batch[0], batch[1])
    al.upload_text(*batch[2:])
    return {"reads_per_s": round(n * steps / el, 1), "ms_per_step": round(1e3 * el / steps, 3)}


def pcie_inclusive(al, batch, steps: int) -> dict:
    """The same step with the batch coming from host memory and the SAM
    records going back to it (reads in over PCIe, text out)."""
    seq, off, names, name_off, quals, qual_off = batch
    n = len(off) - 1
    t0 = time.perf_counter()
    for _ in range(steps):
        al.upload(seq, off)
        al.upload_text(names, name_off, quals, qual_off)
        al.run()
        al.format(sam=True)
        al.text()
    el = time.perf_counter() - t0
    return {"reads_per_s": round(n * steps / el, 1), "ms_per_step": round(1e3 * el / steps, 3),
            "what": "host codes, names and qualities in, align, SAM records formatted on the device, SAM text out"}


def graph_ab(data_dir: Path, batch_reads: int = 262144, steps: int = 10) -> dict:
    """One batch of the 70-250 bp reads over the 1 Mbp reference (the c3
    workload), aligned repeatedly with the launches issued one by one and as
    a captured HIP graph; the two outputs must be identical."""
    prefix, fq, _ = prepare("c3", data_dir)
    seq, off = load_reads(fq)[:2]
    n = min(batch_reads, len(off) - 1)
    s, o = seq[:int(off[n])], off[:n + 1]
    out, res = {}, {"workload": CONFIGS["c3"]["workload"], "batch_reads": n, "steps": steps}
    with g3sa_amd.Aligner(str(prefix), max_reads=n, max_bases=int(o[-1])) as al:
        al.upload(s, o)
        for on in (False, True):
            al.set_graph(on)
            for _ in range(2):  # warm-up (the first graph run captures)
                al.run()
            t0 = time.perf_counter()
            for _ in range(steps):
                al.run()
            res[f"{'graph' if on else 'eager'}_ms_per_batch"] = round(1e3 * (time.perf_counter() - t0) / steps, 3)
            out[on] = al.download()
        res["graph_info"] = al.graph_info()
    a, b = out[False], out[True]
    res["identical"] = bool(np.array_equal(a.aln_off, b.aln_off) and np.array_equal(a.pos, b.pos)
                            and np.array_equal(a.score, b.score) and np.array_equal(a.flag, b.flag)
                            and np.array_equal(a.cigar, b.cigar) and np.array_equal(a.mapq, b.mapq) and a.md == b.md)
    res["speedup"] = round(res["eager_ms_per_batch"] / res["graph_ms_per_batch"], 3)
    return res


def gather_ceiling() -> dict | None:
    """Random-access ceiling for seeding: g3a-gather reads random 16-byte and
    64-byte rows of a 1 GiB table (the narrow Occ lines of the 1 Gbp index
    are 16-byte rows of a 1 GB table) and reports useful bytes per second."""
    exe = g3sa_amd.BUILD / "g3a-gather"
    if not exe.exists():
        return None
    r = subprocess.run([str(exe), "1024"], capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        log(f"g3a-gather failed: {r.stderr[-300:]}")
        return None
    return json.loads(r.stdout.strip().splitlines()[-1])


def seed_traffic(config: str, form: str) -> int | None:
    """Bytes per k_seed launch recorded from the counter passes, or None."""
    name = TRAFFIC_FILES.get((config, form))
    f = ROOT / "profiles" / name if name else None
    if f is None or not f.exists():
        return None
    kv = {}
    for line in f.read_text().splitlines():
        k, sep, v = line.partition(":")
        if sep:
            kv[k.strip()] = v.strip()
    try:
        return int(float(kv["traffic_bytes_per_launch"]))
    except (KeyError, ValueError):
        return None


def end_to_end(prefix: Path, fq: Path, learned: bool) -> dict | None:
    """`g3a mem` over the whole FASTQ (parse, align, format, write to /dev/null),
    timed by the driver from the first read parsed to the last line written."""
    cmd = [str(g3sa_amd.CLI_PATH), "mem", "-S", "-t", "16", "-o", "/dev/null"] + (["--learned"] if learned else [])
    r = subprocess.run(cmd + [str(prefix), str(fq)], capture_output=True, text=True)
    m = re.search(r"(\d+) reads in ([0-9.]+) s \(([0-9.]+) reads/s end to end\)", r.stderr)
    if r.returncode != 0 or not m:
        log(f"g3a mem failed: {r.stderr[-300:]}")
        return None
    out = {"reads_per_s": float(m.group(3)), "reads": int(m.group(1)), "seconds": float(m.group(2)),
           "what": "g3a mem -S" + (" --learned" if learned else "") + ": FASTQ parsed on the host, SAM formatted on "
                   "the device, written to /dev/null, batches of 250k reads, index load (and learned-index build) "
                   "excluded"}
    b = re.search(r"busy time: load ([0-9.]+) s, align\+format ([0-9.]+) s, write ([0-9.]+) s", r.stderr)
    if b:
        out["stage_busy_s"] = {"load": float(b.group(1)), "align_format": float(b.group(2)),
                               "write": float(b.group(3))}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c4", choices=sorted(CONFIGS))
    ap.add_argument("--data", default=os.environ.get("G3A_DATA", "/tmp/g3a_bench"))
    ap.add_argument("--cpu-sample", type=int, default=200000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e2e", action="store_true", help="skip the end-to-end command-line run")
    ap.add_argument("--tuning", default=None, metavar="ROWS,LANE_BLOCKS",
                    help="fixed batched-form tuning instead of the autotuner (e.g. 16,16)")
    ap.add_argument("--seed-blocks", type=int, default=None, metavar="N",
                    help="fixed persistent seeding blocks per CU instead of the autotuner (0: occupancy maximum)")
    ap.add_argument("--seed-block", type=int, default=1024, metavar="T",
                    help="with --seed-blocks: seeding threads per block (a multiple of 64, at most 1024)")
    ap.add_argument("--seed-order", type=int, default=1, choices=[0, 1],
                    help="with --seed-blocks: seed in minimizer order (1) or input order (0)")
    ap.add_argument("--form", default="all", choices=["all", "both", "batched", "plain", "learned"],
                    help="form(s) to time: all = plain, batched, learned; both = plain, batched")
    ap.add_argument("--learned-mode", type=int, default=1, choices=[1, 2],
                    help="learned form: 1 = full SA and model lookups, 2 = full SA only")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run an elementwise copy of MIB MiB (torch) to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 4
        x = torch.ones(n, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.add(x, 0, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(reads are independent: a multi-GPU run would give each GPU its share of a batch's reads, "
              "each with its own copy of the index)", file=sys.stderr)
        return 2
    cfg = CONFIGS[args.config]
    log("random-gather ceiling")
    ceiling = gather_ceiling()
    prefix, fq, prep = prepare(args.config, Path(args.data))
    log("loading reads")
    batch = load_reads(fq)
    seq, off = batch[0], batch[1]
    n = len(off) - 1
    log(f"{n} reads, {int(off[-1])} bases; creating device context")
    t0 = time.time()
    al = g3sa_amd.Aligner(str(prefix), max_reads=n, max_bases=int(off[-1]))
    prep["index_load_upload_s"] = round(time.time() - t0, 1)
    al.upload(seq, off)
    al.upload_text(*batch[2:])
    # seeding grid: fixed, or the autotuner's pick on one aligned pass
    if args.seed_blocks is not None:
        al.set_seed_grid(args.seed_block, args.seed_blocks)
        al.set_seed_order(args.seed_order)
        seed_tuning = {"seed_threads_per_block": args.seed_block, "seed_blocks_per_cu": args.seed_blocks,
                       "seed_ordered": args.seed_order}
    else:
        al.run()
        seed_tuning = al.autotune_seed()
        log(f"seeding autotuned: {seed_tuning}")
    forms = {"all": ["plain", "batched", "learned"], "both": ["plain", "batched"]}.get(args.form, [args.form])
    per_form = {}
    learned_info, learned_tuning = None, None
    for form in forms:
        al.set_form("plain" if form == "plain" else "batched")
        if form != "plain" and args.tuning:
            rows, lb = (int(v) for v in args.tuning.split(","))
            al.set_tuning(rows, lb, 8)
        if form == "learned":
            learned_info = al.set_learned(args.learned_mode)
            log(f"learned index built: {learned_info}")
            if args.seed_blocks is None:
                al.run()
                learned_tuning = al.autotune_seed()
                log(f"seeding autotuned (learned form): {learned_tuning}")
        per_form[form] = run_form(al, n, args.steps, args.warmup, form,
                                  autotune=not args.tuning and (form == "batched" or "batched" not in per_form))
        per_form[form]["paired_end"] = paired_end_times(al, n)
        log(f"[{form}] paired-end primitives: {per_form[form]['paired_end']}")
        # the headline step: halves on two contexts at the same time
        one = {k: per_form[form][k] for k in ("reads_per_s", "ms_per_step")}
        two = two_context_steps(al, batch, args.steps, args.warmup)
        log(f"[{form}] one context {one['ms_per_step']:.1f} ms, two contexts {two['ms_per_step']:.1f} ms per step")
        per_form[form].update(two)
        per_form[form]["one_context"] = one
        if form == "learned":
            per_form[form]["learned_work"] = al.learned_stats()
    main_form = "learned" if "learned" in per_form else ("batched" if "batched" in per_form else forms[0])
    m = per_form[main_form]
    log("PCIe-inclusive pass")
    pcie = pcie_inclusive(al, batch, 1)
    mean, kms, work = m["stages_ms"], m["kernel_ms"], m["work"]
    # dominant kernel and its roofline (k_seed: HBM; extension: integer VALU)
    ext_ms = kms["ext_lane"] + kms["ext_wave"]
    # algorithmic bytes of k_seed: each FM step reads the Occ lines of rows
    # k-1 and k+s-1: two 16-byte lines below 2^32 text rows (two 32-byte
    # lines above), plus the read's bases once; in the learned form also one
    # 4-byte SA entry per text comparison and 2 bits per base it covers, and
    # per model lookup 280 bytes of root and leaf parameters and 16
    # eight-byte keys of window search
    narrow = 2 * cfg["ref_len"] + 1 < (1 << 32)
    step_bytes = 32 if narrow else 64
    lw = m.get("learned_work") or {}
    seed_bytes = (step_bytes * work["fm_steps"] + int(off[-1]) + 4 * lw.get("text_runs", 0)
                  + lw.get("text_run_bases", 0) // 4 + (280 + 128) * lw.get("lookups", 0))
    seed_gbs = seed_bytes / (kms["k_seed"] * 1e-3) / 1e9 if kms["k_seed"] > 0 else 0.0
    dominant = "k_seed" if kms["k_seed"] >= ext_ms else "extension"
    roof = {"bound": "hbm", "kernel": "k_seed", "achieved": round(seed_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(seed_gbs / HBM_PEAK_GBS, 4), "traffic": seed_traffic(args.config, main_form),
            "algorithmic_bytes": seed_bytes, "bytes_per_fm_step": step_bytes, "launch_ms": round(kms["k_seed"], 3)}
    if dominant != "k_seed":
        roof["dominant"] = dominant
    if ceiling:
        roof["random_gather_ceiling_GBps"] = ceiling
    # cells of the reference's extension cell set; the batched form counts
    # only the cells it computes (pruning skips the rest), so take the plain
    # form's count when it ran
    ref_cells = per_form["plain"]["work"]["ext_cells"] if "plain" in per_form else work["ext_cells"]
    gcups = ref_cells / (ext_ms * 1e-3) / 1e9 if ext_ms > 0 else 0.0
    result = {
        "metric": "reads/s (byte-identical SAM)",
        "value": m["reads_per_s"],
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": cfg["workload"], "reads": n, "bases": int(off[-1]),
                   "form": "batched+learned" if main_form == "learned" else main_form,
                   "contexts": 2,
                   "step": "resident reads -> alignments -> SAM records with MAPQ/NM/MD/AS in HBM; the batch's "
                           "halves on two contexts sharing one device index, at the same time"},
        "device_pipeline_reads_per_s": m["device_pipeline_reads_per_s"],
        "sam_format_ms": m["sam_format_ms"],
        "sam_bytes": m["sam_bytes"],
        "pcie_inclusive": pcie,
        "paired_end": m["paired_end"],
        "stages_ms": mean,
        "kernel_ms": kms,
        "work": work,
        "extension_gcups": round(gcups, 2),
        "extension_cells": {"reference_cell_set": ref_cells, "computed": work["ext_cells"]},
        "tuning": m.get("tuning"),
        "seed_tuning": seed_tuning,
        "forms": {f: {"reads_per_s": v["reads_per_s"], "ms_per_step": v["ms_per_step"],
                      "one_context": v["one_context"],
                      "device_pipeline_reads_per_s": v["device_pipeline_reads_per_s"],
                      "sam_format_ms": v["sam_format_ms"], "kernel_ms": v["kernel_ms"],
                      "stages_ms": v["stages_ms"], "fm_steps": v["work"]["fm_steps"]} for f, v in per_form.items()},
        "roofline": roof,
        "prepare": prep,
    }
    if learned_info is not None:
        result["learned_index"] = {**learned_info, "seed_tuning": learned_tuning,
                                   "work": per_form["learned"]["learned_work"]}
        if "plain" in per_form:
            result["learned_vs_plain"] = round(per_form["learned"]["reads_per_s"] / per_form["plain"]["reads_per_s"], 3)
        if "batched" in per_form:
            result["learned_vs_batched"] = round(per_form["learned"]["reads_per_s"] /
                                                 per_form["batched"]["reads_per_s"], 3)
    al.close()
    del batch, seq, off
    log("graph on/off: one 256k-read batch of the c3 workload")
    result["graph_ab"] = graph_ab(Path(args.data))
    if not args.no_e2e:
        log("end-to-end: g3a mem")
        result["end_to_end"] = end_to_end(prefix, fq, main_form == "learned")
    if not args.no_cpu:
        sample = min(args.cpu_sample, n)
        oracle_sam = Path(args.data) / f"{args.config}_oracle_{sample}.sam"
        result["cpu_baseline"] = cpu_baseline(prefix, fq, sample, oracle_sam)
        if result["cpu_baseline"] is not None:
            log(f"parity sample: g3a mem vs the oracle on {sample} reads")
            result["parity_sample"] = parity_sample(prefix, fq, sample, oracle_sam, main_form == "learned")
            oracle_sam.unlink()
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
