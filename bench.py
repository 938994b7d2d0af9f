#!/usr/bin/env python3
"""Mapping throughput of the gfx950 per-read mapping path on one MI355X.

Workload (default ``genome``): a 3.1 Gbp synthetic reference of 24 contigs
(50-250 Mbp) with 8% satellite arrays (171 bp units, 1-3% copy divergence),
92,000 ONT-like reads (lognormal mean 24 kb / sd 12 kb, 7% errors split
40/30/30 sub/ins/del, both strands), k=25 r=0.01 t=0.4 -d 0.075 -o 0.3,
Containment.  The reference index is built on the GPU and stays resident; the
reads are uploaded to HBM before the timed region.  One step maps every read
(batched) and leaves the per-read records in page-locked host memory.  The
timed steps are taken by --mappers mappers (default 3) in turn, each on its
own HIP stream and host thread, so one batch's host waits and kernel tails
overlap another's kernels (the serving loop's workers run the same way); the
single-mapper rate is reported beside it (one_mapper).

Also reported, outside the timed region: the plain one-read-per-launch form on
a sample (with and without its captured graph), per-stage kernel times with the
dominant kernel's roofline, the other configurations (plain_small with its
one-read-per-launch form, batched_uneven, dense: `other_configs`), the serving
loop at a sweep of offered rates, and the CPU baseline (the C oracle on the
host cores, -@1 and all cores) on a bounded sample whose records are checked
bit-for-bit against the GPU's.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import importlib.util
import json
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


G = _load("shmap_amd", os.path.join(ROOT, "shmap-rs_amd", "shmap_amd.py"))

HBM_PEAK_GBS = 8000.0
ENGINE_CLOCK_GHZ = 2.4  # MI355X peak engine clock
PROBE_COPY_BYTES = 1 << 30  # calibration copy: reads and writes exactly this many bytes

# main kernel of each timed stage (for the counter passes)
STAGE_KERNEL = {"sketch": "k_sketch<0>(", "sketch_compact": "k_sketch_compact(", "seed_wave": "k_seed_wave(",
                "seed_wg": "k_seed_wg(",
                "contrib": "k_contrib(", "extract": "k_extract(",
                "prune": "k_prune(", "refine": "k_refine(", "final": "k_final("}


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo, as lscpu prints it)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.lower().startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# workloads (shmap-rs_amd/workloads.py: the five configurations, seeded)
# ---------------------------------------------------------------------------

sys.path.insert(0, os.path.join(ROOT, "shmap-rs_amd"))
import workloads as W  # noqa: E402

WORKLOADS = W.WORKLOADS


def make_workload(name, seed=1, n_reads=None):
    return W.make(name, seed, n_reads, log=lambda m: log("[bench] " + m))


# ---------------------------------------------------------------------------
# device buffers straight from the HIP runtime (inputs resident in HBM)
# ---------------------------------------------------------------------------

class Hip:
    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]

    def upload(self, arr: np.ndarray) -> int:
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(arr.nbytes, 64)) == 0
        assert self.L.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        return p.value

    def free(self, p: int) -> None:
        self.L.hipFree(C.c_void_p(p))

    def sync(self) -> None:
        assert self.L.hipDeviceSynchronize() == 0


# ---------------------------------------------------------------------------
# roofline: algorithmic bytes of each stage
# ---------------------------------------------------------------------------

def stage_bytes(res: np.ndarray, bases_total: int) -> dict:
    m = res["m"].astype(np.int64).sum()
    u = res["u"].astype(np.int64).sum()
    E = res["seed_matches"].astype(np.int64).clip(min=0).sum()
    return {
        "sketch": bases_total + 13 * m,                    # every base read once, (h, r, strand) per kept k-mer
        "sketch_compact": (16 + 13) * m,                   # one 16 B staged record in, (h, r, strand) out
        "seed_wave": 13 * m + 16 * u + 37 * u,             # sketch in, one table slot per probe, seed rows out
        "contrib": 12 * E,                                 # one packed hit record per examined hit
    }


def _pmc_pass(counter: str, args, out_dir: str) -> dict:
    """One rocprofv3 --pmc pass over a probe run; {kernel name: [values per dispatch]}
    (KiB for the size counters, instructions for SQ_INSTS_VALU)."""
    import csv
    import subprocess
    cmd = ["timeout", "-s", "KILL", "240", "rocprofv3", "--pmc", counter, "--output-format", "csv",
           "-d", out_dir, "-o", counter.lower(), "--", sys.executable, os.path.abspath(__file__), "--pmc-probe",
           "--workload", args.workload, "--metric", args.metric]
    if args.reads:
        cmd += ["--reads", str(args.reads)]
    log(f"[bench] counter pass {counter} ...")
    r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True, cwd="/tmp",
                       env=dict(os.environ, TMPDIR="/tmp"))
    if r.returncode != 0:
        raise RuntimeError(f"rocprofv3 --pmc {counter} exited {r.returncode}: {r.stderr[-400:]}")
    found = []
    for dp, _, fs in os.walk(out_dir):
        found += [os.path.join(dp, f) for f in fs if f.startswith(counter.lower()) and
                  f.endswith("counter_collection.csv")]
    if not found:
        raise RuntimeError(f"no counter_collection.csv for {counter}")
    vals = {}
    for f in found:
        for row in csv.DictReader(open(f)):
            if row.get("Counter_Name") != counter:
                continue
            vals.setdefault(row["Kernel_Name"], []).append((int(row["Dispatch_Id"]), float(row["Counter_Value"])))
    log(f"[bench] counter pass {counter} done")
    return {k: [v for _, v in sorted(x)] for k, x in vals.items()}


def measure_traffic(args, stage: str) -> dict:
    """HBM bytes of the stage's kernel in one mapping step, from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each scaled by what the same counter reads
    for a copy of PROBE_COPY_BYTES."""
    import tempfile
    want = STAGE_KERNEL[stage]
    res = {"kernel": want.rstrip("("), "counters": "FETCH_SIZE, WRITE_SIZE"}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        for counter, key in (("FETCH_SIZE", "fetch"), ("WRITE_SIZE", "write")):
            vals = _pmc_pass(counter, args, os.path.join(td, counter.lower()))
            copy = [v for k, x in vals.items() if "k_probe_copy(" in k for v in x]
            kern = [x for k, x in vals.items() if want in k]
            if not copy or not kern:
                raise RuntimeError(f"{counter}: dispatches not found")
            scale = PROBE_COPY_BYTES / (copy[-1] * 1024.0)
            res[key + "_kib"] = kern[0][-1]  # the mapping step's dispatch (the index build's comes first)
            res[key + "_scale"] = round(scale, 4)
            res[key + "_bytes"] = int(kern[0][-1] * 1024.0 * scale)
    res["bytes"] = res["fetch_bytes"] + res["write_bytes"]
    return res


def measure_issue(args, stage: str, kernel_ms: float) -> dict:
    """VALU issue of the stage's kernel in one mapping step: SQ_INSTS_VALU (wave
    instructions; a wave64 VALU instruction takes 4 cycles of a 16-lane SIMD)
    against what the device's SIMDs issue in the kernel's time at the peak
    engine clock."""
    import tempfile
    want = STAGE_KERNEL[stage]
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        vals = _pmc_pass("SQ_INSTS_VALU", args, os.path.join(td, "valu"))
    kern = [x for k, x in vals.items() if want in k]
    if not kern:
        raise RuntimeError("SQ_INSTS_VALU: dispatch not found")
    insts = kern[0][-1]
    simds = 4 * G.device_cus()
    frac = insts * 4.0 / (kernel_ms * 1e-3 * ENGINE_CLOCK_GHZ * 1e9 * simds)
    return {"valu_wave_insts": int(insts), "simds": simds, "engine_clock_ghz": ENGINE_CLOCK_GHZ,
            "valu_issue_frac": round(frac, 4)}


def run_serving(idx, prm, lens, ref, args) -> dict:
    """Config 5: Poisson arrivals of reads (lognormal 1-100 kb, mean 15 kb, an
    even mix of 0.1% and 7% error profiles) into the serving loop over the
    resident index; per-read latency from scheduled arrival to result.  A
    warm-up pass runs every offered rate of the sweep once, then a second pass
    is measured; the headline figures are the run at --serve-rate."""
    import threading
    n = args.serve_reads
    bases, off, _ = W.serving_reads(lens, ref, n)
    hb = G.host_array(len(bases), np.uint8)  # page-locked: a batch of consecutive arrivals is one upload
    hb[:] = bases
    lg = C.CDLL(os.path.join(ROOT, "shmap-rs_amd", "lib", "libshm_probe.so"))
    lg.shm_loadgen_replay.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_double]
    srv = G.Server(idx, prm, workers=args.serve_workers, max_batch_reads=args.serve_batch,
                   max_batch_bases=1 << 31)
    rates = [args.serve_rate] + [r for r in args.serve_sweep if r != args.serve_rate]

    def one_run(rate):
        arrival = np.cumsum(np.random.default_rng(11).exponential(1.0 / rate, n))
        b0 = srv.batch_stats()
        t_begin = srv.now() + 0.005
        rc = []
        th = threading.Thread(target=lambda: rc.append(lg.shm_loadgen_replay(
            srv._p, hb.ctypes.data, off.ctypes.data, arrival.ctypes.data, n, t_begin)))
        th.start()
        got, have = [], 0
        while have < n:
            g = srv.poll(cap=65536, min_results=1, timeout_ms=60_000)
            if not len(g) and not th.is_alive() and srv.counts() == (0, 0, 0):
                break
            got.append(g)
            have += len(g)
        th.join()
        if rc and rc[0] != 0 or have < n:
            raise RuntimeError(f"serving run failed: rc={rc} got={have}/{n} {srv.error()}")
        s = np.concatenate(got)
        s = s[np.argsort(s["tag"])]
        lat = (s["done_s"] - (t_begin + arrival)) * 1e3
        span = float(s["done_s"].max() - t_begin)
        keys = s["worker"].astype(np.float64) * 1e6 + s["start_s"]
        _, first = np.unique(keys, return_index=True)
        bt = (s["done_s"][first] - s["start_s"][first]) * 1e3
        stats = {"offered_reads_per_s": rate, "reads_per_s": round(n / span, 1),
                 "gbp_per_s": round(float(off[-1]) / span / 1e9, 3),
                 "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                 "max_ms": round(float(lat.max()), 3), "batches": int(len(first)),
                 "mean_batch_reads": round(float(s["batch_reads"].mean()), 1),
                 "batch_ms_p50": round(float(np.percentile(bt, 50)), 3),
                 "queue_wait_p50_ms": round(float(np.percentile((s["start_s"] - t_begin - arrival) * 1e3, 50)), 3),
                 "slowest_batches": [[int(s["batch_reads"][first][i]), round(float(bt[i]), 2)]
                                     for i in np.argsort(-bt)[:3]],
                 "batch_events": {k: v - b0[k] for k, v in srv.batch_stats().items()}}
        log(f"[bench] serving at {rate:.0f} reads/s offered: {stats}")
        return s, stats

    # a first pass over every rate sizes the workers' buffers for the batches
    # those rates form and captures their read classes (as a long-running
    # server would have); the second pass is measured
    for r in sorted(rates):
        one_run(r)
    runs = {r: one_run(r) for r in sorted(rates)}
    srv.close()
    s, head = runs[args.serve_rate]
    # the same reads in one device-resident batch: results must not depend on batching
    mp = G.Mapper(idx, prm)
    whole = mp.map_batch((hb, off))
    same = all(np.array_equal(whole[f].view(np.int64) if whole.dtype[f] == np.float64 else whole[f],
                              s["r"][f].view(np.int64) if whole.dtype[f] == np.float64 else s["r"][f])
               for f in whole.dtype.names)
    out = {"reads": n, **head, "workers": args.serve_workers, "mapped_reads": int(s["r"]["mapped"].sum()),
           "same_as_one_batch": same,
           "workload": "lognormal 1-100 kb (mean 15 kb), 50/50 0.1% and 7% errors, Poisson arrivals",
           "batches_as": "speculative, graph-captured per read class",
           "sweep": [runs[r][1] for r in sorted(runs)]}
    return out


def run_cli(args, segs, bases, off, res, prm, idx, h_frac) -> dict:
    """The shmap command line end to end on the same workload: the reference
    and the reads written as FASTA (to /dev/shm when present), then one run
    that builds the index from the FASTA and saves it, and one that maps from
    the saved index; wall time of each, and the CLI's own timings of reading,
    indexing, mapping and output (its --profile-log).  The PAF must equal the
    library's records formatted."""
    import shutil
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "shmap-rs_amd", "lib", "shmap")
    base = "/dev/shm" if os.path.isdir("/dev/shm") else "/tmp"
    td = tempfile.mkdtemp(dir=base, prefix="shmap_cli_")
    threads = min(len(os.sched_getaffinity(0)), 16)
    try:
        t0 = time.time()
        with open(os.path.join(td, "ref.fa"), "wb") as f:
            for name, seq in segs:
                f.write(b">" + name.encode() + b"\n")
                f.write(np.ascontiguousarray(seq).tobytes())
                f.write(b"\n")
        n = len(off) - 1
        with open(os.path.join(td, "reads.fa"), "wb") as f:
            for i0 in range(0, n, 4096):
                parts = []
                for i in range(i0, min(n, i0 + 4096)):
                    parts.append(b">r%d\n" % i)
                    parts.append(bases[off[i]:off[i + 1]].tobytes())
                    parts.append(b"\n")
                f.write(b"".join(parts))
        log(f"[bench] CLI inputs written in {time.time() - t0:.1f}s")
        opts = ["-k", "25", "-r", str(h_frac), "-t", str(W.PARAMS["theta"]), "-d", str(W.PARAMS["min_diff"]),
                "-o", str(W.PARAMS["max_overlap"]), "-@", str(threads), "-x"]
        runs = {}
        for tag, src, extra in (("from_fasta", "ref.fa", ["--save-index", "ref.idx"]), ("from_index", "ref.idx", [])):
            t0 = time.time()
            r = subprocess.run([exe, "-s", src, "-p", "reads.fa", *opts, *extra, "--profile-log", tag + ".json"],
                               cwd=td, capture_output=True, timeout=600)
            wall = time.time() - t0
            if r.returncode != 0:
                raise RuntimeError(f"shmap {tag} exited {r.returncode}: {r.stderr.decode()[-500:]}")
            prof = json.load(open(os.path.join(td, tag + ".json")))
            runs[tag] = (wall, prof, r.stdout.decode())
            log(f"[bench] CLI {tag}: {wall:.2f}s wall, {prof}")
        strip = lambda t: ["\t".join(x for x in l.split("\t") if not x.startswith("t:f:")) for l in t.splitlines() if l]
        want = [G.format_paf(idx, prm, f"r{i}", res[i]) for i in np.nonzero(res["mapped"] == 1)[0]]
        wall, prof, out = runs["from_index"]
        return {"reads_per_s": round(n / wall, 1), "wall_s": round(wall, 3), "threads": threads,
                "read_s": prof["read_s"], "parse_gb_per_s": round(prof["input_bytes"] / max(prof["read_s"], 1e-9) / 1e9, 3),
                "index_load_s": prof["index_s"], "mapping_s": prof["mapping_s"], "output_s": prof["output_s"],
                "from_fasta": {"wall_s": round(runs["from_fasta"][0], 3), "index_s": runs["from_fasta"][1]["index_s"]},
                "same_paf": strip(out) == strip("\n".join(want)) and strip(runs["from_fasta"][2]) == strip(out),
                "mode": "shmap -s <saved index> -p reads.fa: read, map on the GPU, write PAF; files in " + base}
    finally:
        shutil.rmtree(td, ignore_errors=True)


def run_other_workload(name, args, hip) -> dict:
    """Another configuration of the same path, timed the same way as the
    headline (index resident, reads in HBM, tuning from a sample drawn with
    another seed, three mappers taking the steps in turn; one mapper beside)."""
    t0 = time.time()
    lens, ref, bases, off, truth, w = make_workload(name)
    n = len(off) - 1
    k, h_frac = W.PARAMS["k"], w["r"]
    idx = G.SketchIndex(G.FracMinHash(k, h_frac), W.segments(lens, ref), None, device=0)
    hip.sync()
    t_index = time.time() - t0
    prm = G.make_params(k=k, h_frac=h_frac, theta=W.PARAMS["theta"], min_diff=W.PARAMS["min_diff"],
                        max_overlap=W.PARAMS["max_overlap"], metric=args.metric)
    # the sample: half the batch's size (a few heavy reads would set a small
    # sample's time and steer the tuning away from what a full batch needs)
    tb, to, _ = W.reads(name, lens, ref, seed=7, n_reads=max(min(n, 20_000), n // 2))
    mps = [G.Mapper(idx, prm) for _ in range(3)]
    tuning = mps[0].autotune((tb, to))
    del tb, to
    for m_ in mps[1:]:
        m_.set_tuning(**tuning)
    d_b, d_o = hip.upload(bases), hip.upload(off)
    outs = [G.host_array(n, G.RESULT_DTYPE) for _ in mps]
    for m_, o_ in zip(mps, outs):
        m_.map_device(d_b, d_o, n, out=o_)
    hip.sync()
    t1 = time.time()
    for _ in range(2):
        mps[0].map_device(d_b, d_o, n, out=outs[0])
    one_ms = (time.time() - t1) / 2 * 1e3
    steps = 6
    import threading

    def worker(wi):
        for _ in range(wi, steps, len(mps)):
            mps[wi].map_device(d_b, d_o, n, out=outs[wi])
            log(f"[bench] {name} step (mapper {wi}) done")

    t2 = time.time()
    ths = [threading.Thread(target=worker, args=(wi,)) for wi in range(len(mps))]
    for t_ in ths:
        t_.start()
    for t_ in ths:
        t_.join()
    hip.sync()
    el = time.time() - t2
    same = all(np.array_equal(o_, outs[0]) for o_ in outs[1:])
    mps[0].set_timing(True)
    mps[0].map_device(d_b, d_o, n, out=outs[0])
    _, stage = mps[0].last_timing()
    res = outs[0]
    mk = res["mapped"] == 1
    hit = mk & (res["segm_id"] == truth[:n, 0]) & (res["t_l"] < truth[:n, 2]) & (res["t_r"] > truth[:n, 1])
    plain = None
    if name == "plain_small":  # config 1's own form: one read per launch (a captured one-read batch)
        m1 = mps[1]
        m1.map_batch([bases[off[0]:off[1]].tobytes()])
        tp = time.time()
        recs = [m1.map_batch([bases[off[i]:off[i + 1]].tobytes()]) for i in range(n)]
        dt = time.time() - tp
        plain = {"reads_per_s": round(n / dt, 1), "same_records": bool(np.array_equal(np.concatenate(recs), res))}
    hip.free(d_b)
    hip.free(d_o)
    out = {"reads": n, "read_gbp": round(float(off[-1]) / 1e9, 3), "r": h_frac, "plain_form": plain,
           "reads_per_s": round(n * steps / el, 1), "ms_per_step": round(el / steps * 1e3, 3),
           "one_mapper": {"reads_per_s": round(n / one_ms * 1e3, 1), "ms_per_step": round(one_ms, 3)},
           "gbp_per_s": round(float(off[-1]) * steps / el / 1e9, 2), "mapped_reads": int(mk.sum()),
           "overlapping_true_origin": int(hit.sum()), "mappers_agree": bool(same), "tuning": tuning,
           "stage_ms": {s_: round(v, 3) for s_, v in stage.items()}, "index_build_s": round(t_index, 2)}
    log(f"[bench] {name}: {out}")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="genome", choices=sorted(WORKLOADS))
    ap.add_argument("--metric", default="Containment")
    ap.add_argument("--reads", type=int, default=None, help="override the read count")
    ap.add_argument("--plain-reads", type=int, default=200)
    ap.add_argument("--mappers", type=int, default=3,
                    help="mappers mapping whole batches concurrently in the timed steps (each its own HIP stream, "
                         "host thread and page-locked output), as the serving loop's workers do")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving-loop run")
    ap.add_argument("--no-cli", action="store_true", help="skip the command-line end-to-end run")
    ap.add_argument("--no-others", action="store_true",
                    help="skip the other configurations (plain_small, batched_uneven, dense)")
    ap.add_argument("--no-metrics", action="store_true", help="skip the other scoring metrics' timings")
    ap.add_argument("--all-metrics", action="store_true",
                    help="time the other scoring metrics on any workload (default: genome only)")
    ap.add_argument("--serve-reads", type=int, default=200_000)
    ap.add_argument("--serve-rate", type=float, default=1_000_000.0, help="offered reads/s (Poisson arrivals)")
    ap.add_argument("--serve-sweep", type=lambda v: [float(x) for x in v.split(",") if x], default=[
        500_000.0, 1_000_000.0, 2_000_000.0, 4_000_000.0], help="offered rates of the serving sweep (comma list)")
    ap.add_argument("--serve-workers", type=int, default=3)
    ap.add_argument("--serve-batch", type=int, default=16384, help="admission budget: reads per batch")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the built-in batching choices instead of tuning them on the batch first")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc passes that measure the dominant kernel's HBM traffic")
    ap.add_argument("--pmc-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--probe-tuning", default="", help=argparse.SUPPRESS)  # field=value,... for --pmc-probe
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU runs are not part "
              f"of this build: reads shard across GPUs with a replicated index, see DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    lens, ref, bases, off, truth, w = make_workload(args.workload, n_reads=args.reads)
    n_reads = len(off) - 1
    k, h_frac, theta, d, o = W.PARAMS["k"], w["r"], W.PARAMS["theta"], W.PARAMS["min_diff"], W.PARAMS["max_overlap"]
    segs = W.segments(lens, ref)

    hip = Hip()
    t0 = time.time()
    sk = G.FracMinHash(k, h_frac)
    idx = G.SketchIndex(sk, segs, None, device=0)
    hip.sync()
    t_index = time.time() - t0
    st = idx.stats()
    log(f"[bench] GPU index: {st['indexed_kmers']} kmers, {st['indexed_hits']} hits in {t_index:.2f}s")

    if args.pmc_probe:
        # one calibration copy, then one mapping step: run under rocprofv3 --pmc
        probe = C.CDLL(os.path.join(ROOT, "shmap-rs_amd", "lib", "libshm_probe.so"))
        probe.shm_probe_copy.argtypes = [C.c_uint64]
        assert probe.shm_probe_copy(PROBE_COPY_BYTES) == 0
        prm = G.make_params(k=k, h_frac=h_frac, theta=theta, min_diff=d, max_overlap=o, metric=args.metric)
        mp = G.Mapper(idx, prm)
        if args.probe_tuning:
            mp.set_tuning(**{kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.probe_tuning.split(",")})
        d_bases, d_off = hip.upload(bases), hip.upload(off)
        mp.map_device(d_bases, d_off, n_reads)
        hip.sync()
        log("[bench] probe step done")
        return

    prm = G.make_params(k=k, h_frac=h_frac, theta=theta, min_diff=d, max_overlap=o, metric=args.metric)
    mp = G.Mapper(idx, prm)
    d_bases = hip.upload(bases)
    d_off = hip.upload(off)

    # batching choices tuned (shm_mapper_autotune, untimed) on a sample of the
    # same workload drawn with another seed, not on the timed batch
    t_tune = time.time()
    if args.no_autotune:
        tuning = mp.tuning()
    else:
        tb, to, _ = W.reads(args.workload, lens, ref, seed=7, n_reads=min(n_reads, 20_000))
        tuning = mp.autotune((tb, to))
        del tb, to
    tune_s = time.time() - t_tune
    log(f"[bench] tuning {tuning} ({'built-in' if args.no_autotune else f'autotuned in {tune_s:.1f}s'})")

    out = G.host_array(n_reads, G.RESULT_DTYPE)  # page-locked: results come back at link speed
    for i in range(args.warmup):
        res = mp.map_device(d_bases, d_off, n_reads, out=out)
        log(f"[bench] warmup step {i} done")
    hip.sync()
    # one mapper, one step at a time
    t_one = time.time()
    for _ in range(2):
        mp.map_device(d_bases, d_off, n_reads, out=out)
    one_ms = (time.time() - t_one) / 2 * 1e3
    log(f"[bench] one mapper: {one_ms:.2f} ms per step")
    # the timed steps: M mappers (threads, streams) take whole batches in turn,
    # so one batch's host waits and kernel tails overlap another's kernels
    M = max(1, args.mappers)
    mps = [mp] + [G.Mapper(idx, prm) for _ in range(M - 1)]
    for m_ in mps[1:]:
        m_.set_tuning(**tuning)
    outs = [out] + [G.host_array(n_reads, G.RESULT_DTYPE) for _ in range(M - 1)]
    for m_, o_ in zip(mps[1:], outs[1:]):
        m_.map_device(d_bases, d_off, n_reads, out=o_)
    hip.sync()
    times = []
    lock = threading.Lock()

    def worker(wi):
        for i in range(wi, args.steps, M):
            t1 = time.time()
            mps[wi].map_device(d_bases, d_off, n_reads, out=outs[wi])
            dt = time.time() - t1
            with lock:
                times.append(dt)
            log(f"[bench] step {i} (mapper {wi}): {dt * 1e3:.1f} ms")

    t_start = time.time()
    ths = [threading.Thread(target=worker, args=(wi,)) for wi in range(M)]
    for t_ in ths:
        t_.start()
    for t_ in ths:
        t_.join()
    hip.sync()
    elapsed = time.time() - t_start
    value = n_reads * args.steps / elapsed
    res = out
    for o_ in outs[1:]:
        assert np.array_equal(o_, out), "concurrent mappers disagree"
    mps = outs = m_ = o_ = None  # the extra mappers (and their streams) go now
    n_mapped = int(res["mapped"].sum())
    gbp = float(off[-1]) / 1e9

    # per-stage kernel times (separate pass: timing adds syncs between stages)
    mp.set_timing(True)
    res_t = mp.map_device(d_bases, d_off, n_reads)
    mp.set_timing(False)
    total_ms, stage_ms = mp.last_timing()
    counts = mp.last_counts()
    log(f"[bench] work counts: {counts}")
    assert np.array_equal(res_t["mapped"], res["mapped"])
    sb = stage_bytes(res, int(off[-1]))
    dom = max(stage_ms, key=lambda s: stage_ms[s])
    dom_bytes = sb.get(dom)
    roof = {"kernel": dom, "bound": "hbm", "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
            "kernel_ms": round(stage_ms[dom], 4)}
    if dom_bytes:
        ach = dom_bytes / (stage_ms[dom] * 1e-3) / 1e9
        roof.update(achieved=round(ach, 2), frac=round(ach / HBM_PEAK_GBS, 5), algorithmic_bytes=int(dom_bytes))
    else:
        roof.update(achieved=None, frac=None, algorithmic_bytes=None)
    if not args.no_traffic and dom in STAGE_KERNEL:
        try:
            tr = measure_traffic(args, dom)
            roof["traffic"] = tr["bytes"]
            roof["traffic_detail"] = tr
        except Exception as e:  # the counters are a report, not the measurement
            roof["traffic_error"] = str(e)[:300]
            log(f"[bench] traffic: {e}")
        try:
            roof["issue"] = measure_issue(args, dom, stage_ms[dom])
            log(f"[bench] issue: {roof['issue']}")
        except Exception as e:
            roof["issue_error"] = str(e)[:300]
            log(f"[bench] issue: {e}")
    stage_rates = {s: (round(sb[s] / (stage_ms[s] * 1e-3) / 1e9, 2) if s in sb and stage_ms[s] > 0 else None)
                   for s in stage_ms}
    log("[bench] stage ms: " + ", ".join(f"{s}={v:.2f}" for s, v in stage_ms.items()))

    # the same batch under each of the other scoring metrics (two timed steps each)
    # (one mapper, one step at a time)
    per_metric = {args.metric: {"reads_per_s": round(n_reads / one_ms * 1e3, 2), "ms_per_step": round(one_ms, 3),
                                "mapped_reads": n_mapped}}
    # (the genome workload's configuration names all four; elsewhere on request)
    if not args.no_metrics and (args.workload == "genome" or args.all_metrics):
        for mname in ("Containment", "Jaccard", "bucket_SH", "bucket_LCS"):
            if mname == args.metric:
                continue
            pm = G.Mapper(idx, G.make_params(k=k, h_frac=h_frac, theta=theta, min_diff=d, max_overlap=o,
                                             metric=mname))
            pm.set_tuning(**tuning)
            rm = pm.map_device(d_bases, d_off, n_reads, out=out)
            hip.sync()
            log(f"[bench] {mname}: warmup step done")
            tm = time.time()
            for _ in range(2):
                rm = pm.map_device(d_bases, d_off, n_reads, out=out)
                log(f"[bench] {mname}: step done")
            hip.sync()
            ms = (time.time() - tm) / 2 * 1e3
            per_metric[mname] = {"reads_per_s": round(n_reads / ms * 1e3, 2), "ms_per_step": round(ms, 3),
                                 "mapped_reads": int(rm["mapped"].sum())}
            del pm
        log(f"[bench] per metric: {per_metric}")
        res = mp.map_device(d_bases, d_off, n_reads, out=out)  # the headline metric's records again

    # placement: mapped reads against the generator's true origin (contig, start, end)
    mk = res["mapped"] == 1
    hit = mk & (res["segm_id"] == truth[:n_reads, 0]) & (res["t_l"] < truth[:n_reads, 2]) & \
        (res["t_r"] > truth[:n_reads, 1])
    placement = {"mapped": int(mk.sum()), "overlapping_true_origin": int(hit.sum()),
                 "frac": round(float(hit.sum()) / max(int(mk.sum()), 1), 5)}
    log(f"[bench] placement: {placement}")

    # plain form: one read per launch (a speculative batch of one read, its
    # stream work captured once and replayed), then without the graph
    n_plain = min(args.plain_reads, n_reads)
    plain_reads = [bases[off[i]:off[i + 1]].tobytes() for i in range(n_plain)]
    plain_lat = {}
    plain_rate = None
    plain_same = True
    for graphs in (True, False):
        mp.set_graphs(graphs)
        if n_plain:
            mp.map_batch([plain_reads[0]])  # captures the one-read class
        recs, lat_p = [], []
        tp = time.time()
        for r in plain_reads:
            t1 = time.time()
            recs.append(mp.map_batch([r]))
            lat_p.append((time.time() - t1) * 1e3)
        rate = n_plain / (time.time() - tp) if n_plain else None
        plain = np.concatenate(recs) if n_plain else res[:0]
        plain_same = plain_same and bool(np.array_equal(plain, res[:n_plain]))
        key = "graph" if graphs else "no_graph"
        plain_lat[key] = {"reads_per_s": round(rate, 1) if rate else None,
                          "p50_ms": round(float(np.percentile(lat_p, 50)), 4) if n_plain else None,
                          "p99_ms": round(float(np.percentile(lat_p, 99)), 4) if n_plain else None}
        if graphs:
            plain_rate = rate
    mp.set_graphs(True)
    log(f"[bench] plain form: {plain_lat}")

    # End to end from host memory: the reads start in page-locked host memory,
    # 2-bit packed as the reader delivers them (shm_pack_reads; its rate is
    # reported beside), and every step uploads its batch and maps it; M mappers
    # take the steps in turn, so one batch's upload overlaps another's kernels.
    t_pack = time.time()
    pk = G.PackedReads((bases, off), threads=min(len(os.sched_getaffinity(0)), 16))
    pack_s = time.time() - t_pack
    log(f"[bench] packed {pk.bases / 1e9:.3f} Gbp into {len(pk.packed) / 1e9:.3f} GB in {pack_s:.2f}s "
        f"({pk.n_exc} runs of other bytes)")
    mps = [mp] + [G.Mapper(idx, prm) for _ in range(M - 1)]
    for m_ in mps[1:]:
        m_.set_tuning(**tuning)
    outs = [G.host_array(n_reads, G.RESULT_DTYPE) for _ in range(M)]
    # two untimed steps per mapper: the first sizes its buffers for the packed
    # path's padded read class, the second settles the capacities it speculates
    # from (a buffer that still grew inside the timed steps would wait, in
    # hipFree, for every other mapper's upload)
    for _ in range(2):
        for m_, o_ in zip(mps, outs):
            m_.map_packed(pk, out=o_)
    assert np.array_equal(outs[0], res), "packed reads map differently"
    e2e_steps = max(args.steps, M)

    def e2e_worker(wi):
        for i in range(wi, e2e_steps, M):
            mps[wi].map_packed(pk, out=outs[wi])
            log(f"[bench] end-to-end step {i} (mapper {wi}) done")

    hip.sync()
    tq = time.time()
    ths = [threading.Thread(target=e2e_worker, args=(wi,)) for wi in range(M)]
    for t_ in ths:
        t_.start()
    for t_ in ths:
        t_.join()
    hip.sync()
    e2e_s = time.time() - tq
    for o_ in outs:
        assert np.array_equal(o_, res), "concurrent packed mappers disagree"
    mps = outs = m_ = o_ = None
    end_to_end = {"reads_per_s": round(n_reads * e2e_steps / e2e_s, 2),
                  "ms_per_step": round(e2e_s / e2e_steps * 1e3, 3), "steps": e2e_steps, "mappers": M,
                  "upload_gb_per_step": round(len(pk.packed) / 1e9, 3),
                  "gbp_per_s": round(pk.bases * e2e_steps / e2e_s / 1e9, 3),
                  "pack_gbp_per_s": round(pk.bases / pack_s / 1e9, 3),
                  "mode": "reads in page-locked host memory, 2-bit packed, uploaded and unpacked inside each step, "
                          "upload overlapped across mappers"}
    log(f"[bench] end to end: {end_to_end}")
    del pk
    # the same with the reads as bytes (four times the upload), one mapper
    h_bases = G.host_array(len(bases), np.uint8)
    h_bases[:] = bases
    hip.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.sync()
    tq = time.time()
    for _ in range(2):
        assert hip.L.hipMemcpy(C.c_void_p(d_bases), C.c_void_p(h_bases.ctypes.data), len(bases), 1) == 0
        assert hip.L.hipMemcpy(C.c_void_p(d_off), C.c_void_p(off.ctypes.data), off.nbytes, 1) == 0
        mp.map_device(d_bases, d_off, n_reads, out=out)
    pcie_ms = (time.time() - tq) / 2 * 1e3
    del h_bases
    pcie = {"reads_per_s": round(n_reads / pcie_ms * 1e3, 2), "ms_per_step": round(pcie_ms, 3),
            "upload_gb": round(len(bases) / 1e9, 3), "mode": "reads as bytes, one mapper, upload then map"}
    log(f"[bench] bytes upload: {pcie}")

    # whole-path algorithmic bytes per step: bases, one 16 B probe per distinct
    # k-mer, 8 B per examined hit, 2m reference k-mers of 12 B per refined
    # bucket, 16 B per sketched k-mer, and 8 B per hit record the prune pass
    # reads (P, counted by the oracle on the CPU sample below, scaled to the
    # batch by reads)
    Lr = (off[1:] - off[:-1]).astype(np.int64)
    pb = int((Lr + 16 * res["u"].astype(np.int64) + 8 * res["seed_matches"].astype(np.int64).clip(min=0)
              + 24 * res["refined_buckets"].astype(np.int64) * res["m"].astype(np.int64)
              + 16 * res["m"].astype(np.int64)).sum())

    def path_roofline(prune_hits):
        b = pb + 8 * (prune_hits or 0)
        gbs = b * args.steps / elapsed / 1e9
        return {"bytes_per_step": b, "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(gbs / HBM_PEAK_GBS, 5), "prune_hits": prune_hits,
                "prune_hits_by": "oracle on the CPU sample, scaled to the batch" if prune_hits is not None
                else "not counted (no CPU sample)"}

    path_roof = path_roofline(None)

    cli = None
    if not args.no_cli:
        try:
            cli = run_cli(args, segs, bases, off, res, prm, idx, h_frac)
        except Exception as e:  # a report beside the measurement
            cli = {"error": str(e)[:300]}
        log(f"[bench] CLI: {cli}")

    # the other configurations of the same path (after the genome's own legs)
    others = None
    if not args.no_others and args.workload == "genome":
        others = {}
        for name in ("plain_small", "batched_uneven", "dense"):
            others[name] = run_other_workload(name, args, hip)

    serving = None
    if not args.no_serve:
        mp = None  # the serving workers get the device's hardware queues to themselves
        serving = run_serving(idx, prm, lens, ref, args)
        log(f"[bench] serving: {serving}")

    # latency: every read of a step completes with the step
    lat = np.array(times) * 1e3

    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_ffi as O  # the CPU oracle: baseline and checker only
        ncores = len(os.sched_getaffinity(0))
        tc = time.time()
        oidx = O.OracleIndex(k, h_frac, segs, None, threads=ncores)
        log(f"[bench] oracle index built in {time.time() - tc:.1f}s")
        op = O.params(k=k, h_frac=h_frac, theta=theta, min_diff=d, max_overlap=o, metric=args.metric)
        # bounded sample: grow until ~cpu_seconds of single-thread work
        n_s, spent, ores = 64, 0.0, None
        while True:
            n_s = min(n_s, n_reads)
            sample = (bases, off[:n_s + 1])
            ores, spent = oidx.map_reads_timed(op, sample, threads=1)
            if spent >= args.cpu_seconds / 4 or n_s == n_reads:
                break
            n_s = int(n_s * min(8.0, max(2.0, args.cpu_seconds / 4 / max(spent, 1e-3))))
        cpu1 = n_s / spent
        ores_all, spent_all = oidx.map_reads_timed(op, sample, threads=ncores)
        cpun = n_s / spent_all
        path_roof = path_roofline(int(round(oidx.last_prune_hits * n_reads / n_s)))
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import parity_util  # tests/parity_util.py: field-by-field record diff
        bad = parity_util.diff_results(ores, res[:n_s])
        bad_all = parity_util.diff_results(ores_all, res[:n_s])
        parity = {"reads_compared": n_s, "mismatches": len(bad) + len(bad_all), "first": (bad + bad_all)[:3]}
        cpu = {"value": round(cpu1, 2), "unit": "reads/s", "cores": 1, "kind": "port",
               "sample": f"first {n_s} reads of the same workload, oracle -@1 ({spent:.1f}s)",
               "all_cores": {"value": round(cpun, 2), "cores": ncores}}
        log(f"[bench] CPU oracle: {cpu1:.1f} reads/s at -@1, {cpun:.1f} reads/s at -@{ncores}; parity {parity}")
        cpu["cpu_model"] = cpu_model()
        if args.workload == "genome":
            # sanity anchor: the oracle's -@1 mean per-read time on a B02-like
            # input (the 3.1 Gbp reference, 24 kb reads with 0.5% errors, r=0.01)
            # against the reference's published 0.184 ms/read on a Xeon Gold 5218
            # (RESULTS.md:345, per-read-B02-Containment.tsv)
            ab, ao, _ = G.synth_reads(17, ref, lens, 2000, (24_000, 0), lognormal=False, clip=(24_000, 24_000),
                                      errors=((0.005, 0.6, 0.2, 0.2),), both_strands=True)
            ares, aspent = oidx.map_reads_timed(op, (ab, ao), threads=1)
            ms = aspent / (len(ao) - 1) * 1e3
            cpu["anchor"] = {"input": "B02-like: 2,000 reads of 24 kb, 0.5% errors (60/20/20), genome reference",
                             "oracle_ms_per_read_at_1": round(ms, 4), "published_ms_per_read": 0.184,
                             "ratio": round(ms / 0.184, 3), "mapped": int(ares["mapped"].sum())}
            log(f"[bench] CPU anchor: {cpu['anchor']}")

    out = {
        "metric": "mapped reads/s, 3.1 Gbp ref r=0.01 (batched vs plain GPU, bit-exact PAF); p99 ms",
        "value": round(value, 2),
        "unit": "reads/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator, shmap-rs_amd/csrc/synth.c)",
        "config": {"workload": args.workload, "reference_gbp": round(float(lens.sum()) / 1e9, 3),
                   "contigs": int(len(lens)), "reads": n_reads, "read_gbp": round(gbp, 3), "k": k, "r": h_frac,
                   "t": theta, "d": d, "o": o, "metric": args.metric, "mode": "batched, device-resident",
                   "mappers": M, "tuning": tuning,
                   "tuning_by": "built-in rule" if args.no_autotune else
                   "shm_mapper_autotune on 20,000 reads of the same workload drawn with another seed"},
        "one_mapper": {"reads_per_s": round(n_reads / one_ms * 1e3, 2), "ms_per_step": round(one_ms, 3)},
        "gbp_per_s": round(gbp * args.steps / elapsed, 3),
        "mapped_reads": n_mapped,
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "plain_gpu": {"reads_per_s": round(plain_rate, 2) if plain_rate else None, "reads": n_plain,
                      "same_records": plain_same, **plain_lat},
        "batched_over_plain": round(value / plain_rate, 1) if plain_rate else None,
        "end_to_end": end_to_end,
        "pcie_inclusive_bytes": pcie,
        "per_metric": per_metric,
        "placement": placement,
        "serving": serving,
        "other_configs": others,
        "cli": cli,
        "path_roofline": path_roof,
        "index_build_s": round(t_index, 3),
        "stage_ms": {s: round(v, 3) for s, v in stage_ms.items()},
        "work_counts": counts,
        "stage_gbs": stage_rates,
        "roofline": roof,
        "cpu_baseline": cpu,
        "parity_sample": parity,
    }
    if cpu:
        out["gpu_over_cpu_all_cores"] = round(value / cpu["all_cores"]["value"], 1)
    hip.free(d_bases)
    hip.free(d_off)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
