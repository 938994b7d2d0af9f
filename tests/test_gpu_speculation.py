"""Speculative batches: once a mapper has seen a batch, it sizes the next
one's device buffers from the shapes seen so far and synchronises with the
host once per batch.  Its records are those of the synchronised path and of
the oracle -- also when a batch outgrows the learned sizes (device checks stop
it and it is mapped again), when it is smaller, and when it holds empty and
short reads."""
import numpy as np
import pytest

import parity_util as P
import shmap_amd as G
import test_gpu_parity as T

pytestmark = pytest.mark.gpu

PRM = dict(theta=0.4, min_diff=0.075, max_overlap=0.3)


def _repeat_workload(seed=5, unit=6000, copies=24, unique=240_000, n_reads=40, read_len=2500):
    """One unique contig and `copies` near-identical copies of one unit, with
    reads from each: a read from the unit has ~copies hits per seed, so its
    bucket tables and buckets per k-mer far exceed a unique read's."""
    rng = np.random.default_rng(seed)
    acgt = np.frombuffer(b"ACGT", dtype=np.uint8)
    uniq = acgt[rng.integers(0, 4, unique)]
    base = acgt[rng.integers(0, 4, unit)]
    segs = [("uniq", uniq.tobytes())]
    for c in range(copies):
        u = base.copy()
        pos = rng.integers(0, unit, unit // 100)  # 1% divergence per copy
        u[pos] = acgt[rng.integers(0, 4, len(pos))]
        segs.append((f"rep{c}", u.tobytes()))

    def reads_from(src, n):
        out = []
        for _ in range(n):
            a = int(rng.integers(0, len(src) - read_len))
            r = src[a:a + read_len].copy()
            pos = rng.integers(0, read_len, read_len // 200)
            r[pos] = acgt[rng.integers(0, 4, len(pos))]
            out.append(r.tobytes())
        return out

    return segs, reads_from(uniq, n_reads), reads_from(base, n_reads)


@pytest.mark.parametrize("metric", ["Containment", "bucket_LCS"])
def test_speculative_batches_give_the_synchronised_records(metric):
    segs, reads = T._workload(**T.WORKLOADS["repeats"])
    prm = dict(PRM, metric=metric)
    oi, op, ores = P.oracle_map(segs, reads, 15, 0.1, **prm)
    sk = G.FracMinHash(15, 0.1)
    idx = G.SketchIndex(sk, segs)
    p = G.make_params(k=15, h_frac=0.1, **prm)
    spec = G.Mapper(idx, p)
    sync = G.Mapper(idx, p)
    sync.set_speculation(False)
    ref = sync.map_batch(reads)
    assert sync.last_counts()["host_syncs"] >= 3
    assert not P.diff_results(ores, ref)
    first = spec.map_batch(reads)  # nothing seen yet: synchronised
    assert spec.last_counts()["host_syncs"] >= 3 and spec.last_counts()["redone"] == 0
    assert not P.diff_results(ores, first)
    n = len(reads)
    for lo, hi in ((0, n), (0, 40), (7, 9), (0, n)):  # the same shape, smaller ones, the same again
        res = spec.map_batch(reads[lo:hi])
        c = spec.last_counts()
        # the synchronised batch left buffers for batches like it
        assert c["host_syncs"] == 1 and c["redone"] == 0 and c["allocations"] == 0, c
        bad = P.diff_results(ores[lo:hi], res)
        assert not bad, bad
    # counts other than the synchronisations are the synchronised path's
    spec.map_batch(reads)
    sync.map_batch(reads)
    a, b = spec.last_counts(), sync.last_counts()
    for k in G.COUNT_NAMES:
        if k not in ("host_syncs", "redone", "allocations", "graph", "captured"):
            assert a[k] == b[k], (k, a[k], b[k])
    # results straight into page-locked memory
    out = G.host_array(len(reads), G.RESULT_DTYPE)
    pk = G.PackedReads(reads, threads=2)
    res = spec.map_packed(pk, out=out)
    assert spec.last_counts()["host_syncs"] == 1
    assert not P.diff_results(ores, res)


def test_a_batch_that_outgrows_the_learned_sizes_is_mapped_again():
    segs, uniq_reads, rep_reads = _repeat_workload()
    prm = dict(PRM, metric="Jaccard")
    oi, op, ores = P.oracle_map(segs, uniq_reads + rep_reads, 15, 0.1, **prm)
    sk = G.FracMinHash(15, 0.1)
    idx = G.SketchIndex(sk, segs)
    p = G.make_params(k=15, h_frac=0.1, **prm)
    mp = G.Mapper(idx, p)
    res_u = mp.map_batch(uniq_reads)
    assert not P.diff_results(ores[:len(uniq_reads)], res_u)
    res_u = mp.map_batch(uniq_reads)
    assert mp.last_counts()["host_syncs"] == 1
    assert not P.diff_results(ores[:len(uniq_reads)], res_u)
    # the repeat reads have many times the table slots and buckets per k-mer
    res_r = mp.map_batch(rep_reads)
    c = mp.last_counts()
    assert c["redone"] == 1 and c["host_syncs"] >= 4, c
    bad = P.diff_results(ores[len(uniq_reads):], res_r)
    assert not bad, bad
    # learned: the same batch again goes through in one synchronisation
    res_r = mp.map_batch(rep_reads)
    c = mp.last_counts()
    assert c["redone"] == 0 and c["host_syncs"] == 1, c
    assert not P.diff_results(ores[len(uniq_reads):], res_r)
    # and a mixed batch, then reads with no k-mers at all
    mixed = uniq_reads[:10] + rep_reads[:10]
    res = mp.map_batch(mixed)
    assert not P.diff_results(np.concatenate([ores[:10], ores[len(uniq_reads):len(uniq_reads) + 10]]), res)
    res = mp.map_batch([b"", b"ACGT", b"A" * 14])
    assert list(res["mapped"]) == [0, 0, 0] and list(res["m"]) == [0, 0, 0]
    res = mp.map_batch(rep_reads[:3])
    assert not P.diff_results(ores[len(uniq_reads):len(uniq_reads) + 3], res)


@pytest.mark.parametrize("metric", ["Jaccard", "bucket_LCS"])
def test_graph_replays_give_the_eager_records(metric):
    """Batches captured into a graph per read class (after a run of batches
    of the class) and replayed, with eager work of other mappers and
    synchronisations between the replays: every batch's records equal those of
    the same batch mapped without graphs (and the oracle's), for classes the
    batch fills exactly and ones it is padded to, down to a single read."""
    segs, reads = T._workload(**T.WORKLOADS["uniform"])
    prm = dict(PRM, metric=metric)
    oi, op, ores = P.oracle_map(segs, reads, 15, 0.1, **prm)
    idx = G.SketchIndex(G.FracMinHash(15, 0.1), segs)
    p = G.make_params(k=15, h_frac=0.1, **prm)
    gm = G.Mapper(idx, p)
    eager = G.Mapper(idx, p)
    eager.set_graphs(False)
    sync = G.Mapper(idx, p)
    sync.set_speculation(False)
    gm.map_batch(reads)  # learns the shape: synchronised
    eager.map_batch(reads)
    ranges = [(0, 96), (10, 106), (0, 1), (5, 6), (0, 5), (20, 25), (0, 150), (3, 99), (7, 8)]
    want = {r: eager.map_batch(reads[r[0]:r[1]]) for r in ranges}
    for rep in range(2):
        for lo, hi in ranges:
            # a class is captured after 8 batches of it in a row, then replayed
            flags = []
            for i in range(10):
                got = gm.map_batch(reads[lo:hi])
                c = gm.last_counts()
                assert c["host_syncs"] == 1 and c["redone"] == 0, (lo, hi, c)
                flags.append((c["graph"], c["captured"]))
                bad = P.diff_results(want[(lo, hi)], got)
                assert not bad, (lo, hi, i, bad)
                bad = P.diff_results(ores[lo:hi], got)
                assert not bad, (lo, hi, i, bad)
                sync.map_batch(reads[hi:hi + 7])  # eager work and synchronisations between replays
            # captured at the latest by the 8th batch since the capacities
            # last changed, then replayed
            assert flags[-1] == (1, 0) and flags[-2][0] == 1, (lo, hi, flags)
            if rep == 1:  # the second round finds every class's buffers in place
                assert c["allocations"] == 0
    # graphs off on the same mapper: the same records again
    gm.set_graphs(False)
    got = gm.map_batch(reads[:96])
    assert gm.last_counts()["graph"] == 0
    assert not P.diff_results(ores[:96], got)
