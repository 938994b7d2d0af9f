"""The five benchmark workloads, mapped on the GPU and by the oracle at their
real parameters (k=25, r=0.01 or 0.10, t=0.4, -d 0.075 -o 0.3), compared
record by record: every counter and every PAF field of every read bit for
bit, and the PAF text of every mapped read.  The reference gates its own
benchmark on whole-run PAF equality (benchmarks/data/suite.toml:428-431).

  plain_small     full size (10 Mbp, 2,000 x 10 kb), batched and one read per
                  call
  batched_uneven  full size (100 Mbp with 20% tandem arrays, 20,000 HiFi-like
                  reads), one batch under a tuning chosen by the autotuner on
                  reads of another seed
  genome          the 3.1 Gbp reference, all 92,000 ONT-like reads, under each
                  of the four metrics
  serving         the genome index, all 200,000 reads of the serving mix
                  submitted over time (Poisson arrivals) to the serving loop
  dense           the 3.1 Gbp reference at r=0.10 (310 M indexed k-mers), all
                  125,000 HiFi-like reads

Slow by design (the 3.1 Gbp reference, two oracle indexes); every phase
prints a progress line."""
import os
import sys
import threading
import time

import numpy as np
import pytest

import oracle_ffi as O
import parity_util as P
import shmap_amd as G
import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

T0 = time.time()


def say(*a):
    print(f"[configs {time.time() - T0:6.1f}s]", *a, file=sys.stderr, flush=True)


def threads():
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def compare(name, segs_gpu_idx, prm, o_idx, op, names, ores, gres):
    """Every field of every read, then the PAF text of the mapped reads."""
    bad = P.diff_results(ores, gres)
    assert not bad, (name, bad)
    mapped = np.nonzero(ores["mapped"] == 1)[0]
    for i in mapped:
        g = G.format_paf(segs_gpu_idx, prm, names[i], gres[i])
        o = o_idx.format_paf(op, names[i], ores[i])
        assert g == o, (name, i, g, o)
    return len(mapped)


def run_workload(name, lens, ref, metric="Containment", n_reads=None, o_idx=None, g_idx=None, tune_on=None,
                 check_one_per_call=False):
    w = W.WORKLOADS[name]
    r = w["r"]
    bases, off, truth = W.reads(name, lens, ref, 2, n_reads)
    n = len(off) - 1
    names = [f"S{i}_1" for i in range(n)]
    segs = W.segments(lens, ref)
    if g_idx is None:
        g_idx = G.SketchIndex(G.FracMinHash(25, r), segs, None)
        say(f"{name}: GPU index built")
    if o_idx is None:
        o_idx = O.OracleIndex(25, r, segs, None, threads=threads())
        say(f"{name}: oracle index built")
    prm = G.make_params(h_frac=r, metric=metric, **W.PARAMS)
    op = O.params(h_frac=r, metric=metric, **W.PARAMS)
    mp = G.Mapper(g_idx, prm)
    if tune_on is not None:
        tb, to, _ = tune_on
        say(f"{name}: autotune -> {mp.autotune((tb, to))}")
    gres = mp.map_batch((bases, off))
    say(f"{name} {metric}: GPU mapped {n} reads ({int(gres['mapped'].sum())} mapped)")
    ores = o_idx.map_reads(op, (bases, off), threads=threads())
    say(f"{name} {metric}: oracle mapped {n} reads ({int(ores['mapped'].sum())} mapped)")
    nm = compare(f"{name}/{metric}", g_idx, prm, o_idx, op, names, ores, gres)
    if check_one_per_call:
        one = np.concatenate([mp.map_batch([bases[off[i]:off[i + 1]].tobytes()]) for i in range(n)])
        bad = P.diff_results(gres, one)
        assert not bad, bad
        say(f"{name}: {n} reads one per call equal the batch")
    return gres, truth, nm, g_idx, o_idx


def test_plain_small_full():
    lens, ref = W.reference("plain_small")
    gres, truth, nm, _, _ = run_workload("plain_small", lens, ref, check_one_per_call=True)
    # 0.5% substitutions keep 88% of the 25-mers: nearly every read maps
    assert nm > 1900


def test_batched_uneven_full():
    lens, ref = W.reference("batched_uneven")
    tune = W.reads("batched_uneven", lens, ref, 7, 2000)  # another seed's reads
    gres, truth, nm, _, _ = run_workload("batched_uneven", lens, ref, tune_on=tune)
    assert nm > 10_000


@pytest.fixture(scope="module")
def genome():
    say("genome: generating the 3.1 Gbp reference")
    lens, ref = W.reference("genome")
    segs = W.segments(lens, ref)
    g_idx = G.SketchIndex(G.FracMinHash(25, 0.01), segs, None)
    say("genome: GPU index built")
    o_idx = O.OracleIndex(25, 0.01, segs, None, threads=threads())
    say("genome: oracle index built")
    st_g, st_o = g_idx.stats(), o_idx.stats()
    for key in ("indexed_kmers", "indexed_hits", "blacklisted_kmers"):
        assert st_g[key] == st_o[key], key
    yield lens, ref, g_idx, o_idx
    del g_idx, o_idx


@pytest.mark.parametrize("metric", ["Containment", "Jaccard", "bucket_SH", "bucket_LCS"])
def test_genome_all_reads(genome, metric):
    lens, ref, g_idx, o_idx = genome
    gres, truth, nm, _, _ = run_workload("genome", lens, ref, metric=metric, g_idx=g_idx, o_idx=o_idx)
    assert len(gres) == 92_000


def test_serving_mix_matches_oracle(genome):
    lens, ref, g_idx, o_idx = genome
    n = 200_000
    bases, off, _ = W.serving_reads(lens, ref, n)
    prm = G.make_params(h_frac=0.01, **W.PARAMS)
    op = O.params(h_frac=0.01, **W.PARAMS)
    hb = G.host_array(len(bases), np.uint8)
    hb[:] = bases
    srv = G.Server(g_idx, prm, workers=3, max_batch_reads=16384)
    rng = np.random.default_rng(13)
    arrival = np.cumsum(rng.exponential(1.0 / 1_000_000.0, n))  # Poisson, 1 M reads/s offered
    reads = [hb[off[i]:off[i + 1]] for i in range(n)]

    def submit():
        t0 = time.perf_counter()
        i = 0
        while i < n:
            now = time.perf_counter() - t0
            j = int(np.searchsorted(arrival, now, side="right"))
            if j > i:
                srv.submit_many(np.arange(i, j, dtype=np.uint64), [bytes(r) for r in reads[i:j]])
                i = j
            else:
                time.sleep(0.0002)

    th = threading.Thread(target=submit)
    th.start()
    got = []
    t_end = time.time() + 120
    while sum(len(g) for g in got) < n:
        # an empty poll is normal while nothing is queued yet
        g = srv.poll(cap=65536, min_results=1, timeout_ms=1000)
        assert len(g) or time.time() < t_end, "serving loop stalled"
        got.append(g)
    th.join()
    s = np.concatenate(got)
    s = s[np.argsort(s["tag"])]
    srv.close()
    assert np.array_equal(s["tag"], np.arange(n))
    assert np.all(s["r"]["status"] == 0)
    say(f"serving: {n} reads served, {len(np.unique(s['start_s']))} batches, "
        f"{int(s['r']['mapped'].sum())} mapped")
    ores = o_idx.map_reads(op, (bases, off), threads=threads())
    names = [f"q{i}" for i in range(n)]
    nm = compare("serving", g_idx, prm, o_idx, op, names, ores, s["r"])
    assert nm > n // 4


def test_dense_all_reads(genome):
    lens, ref, _, _ = genome
    gres, truth, nm, g_idx, o_idx = run_workload("dense", lens, ref)
    # HiFi-like reads at r=0.10: every mapped read lies on its true origin
    mk = gres["mapped"] == 1
    on = mk & (gres["segm_id"] == truth[:, 0]) & (gres["t_l"] < truth[:, 2]) & (gres["t_r"] > truth[:, 1])
    say(f"dense: {int(mk.sum())} mapped, {int(on.sum())} on their true origin")
    assert mk.sum() > 120_000 and on.sum() >= 0.99 * mk.sum()
    del g_idx, o_idx


def test_dense_bucket_lcs_sample(genome):
    """bucket_LCS at r=0.10: sketches of m ~ 2,300 (most above the wave's
    register tails, so the patience tails continue in LDS) on the first 2,000
    dense reads."""
    lens, ref, _, _ = genome
    gres, truth, nm, g_idx, o_idx = run_workload("dense", lens, ref, metric="bucket_LCS", n_reads=2000)
    assert nm > 1500 and (gres["m"] > 2048).sum() > 500
    del g_idx, o_idx
