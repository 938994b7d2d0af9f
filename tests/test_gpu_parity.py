"""HIP path against the CPU oracle, through the C ABI, on the same seeded
inputs.  Bit-exact: every counter and every PAF field of every read."""
import os

import numpy as np
import pytest

import fastx
import oracle_ffi as O
import parity_util as P
import shmap_amd as G

pytestmark = pytest.mark.gpu


def _rng_seq(seed, n, alphabet=b"ACGT"):
    r = np.random.default_rng(seed)
    return np.frombuffer(alphabet, dtype=np.uint8)[r.integers(0, len(alphabet), n)].tobytes()


def test_golden_paf_through_the_gpu(golden_dir):
    """tests/golden_paf.rs:26-53 -- -k 8 -r 1.0 -t 0.1 on the tiny fixture."""
    ref = fastx.read_fasta(os.path.join(golden_dir, "tiny_ref.fa"))
    reads = fastx.read_fasta(os.path.join(golden_dir, "tiny_reads.fa"))
    idx, p, mp, res = P.gpu_map(ref, [s for _, s in reads], 8, 1.0, theta=0.1)
    lines = [G.format_paf(idx, p, n, r) for (n, _), r in zip(reads, res) if r["mapped"]]
    want = fastx.strip_time(open(os.path.join(golden_dir, "tiny.golden.paf")).read())
    assert fastx.strip_time("\n".join(lines)) == want


@pytest.mark.parametrize("k,h_frac", [(4, 1.0), (8, 1.0), (15, 0.5), (25, 0.01), (31, 0.1), (63, 0.2), (70, 0.3)])
def test_sketch_matches_oracle(k, h_frac):
    sk = G.FracMinHash(k, h_frac)
    for seed, n in [(1, 0), (2, k - 1), (3, k), (4, k + 1), (5, 5000), (6, 100_003)]:
        s = _rng_seq(seed, n, b"ACGTacgtNRY")
        for off in (0, 17):
            g = sk.sketch_slice_into(s, off)
            o = O.sketch(k, h_frac, s, off)
            assert len(g) == len(o), (k, h_frac, n)
            assert np.array_equal(g["r"], o["r"]) and np.array_equal(g["h"], o["h"])
            assert np.array_equal(g["strand"], o["strand"])


@pytest.mark.parametrize("k,h_frac", [(25, 0.01), (31, 0.1), (15, 0.5)])
def test_sketch_lane_overflow_matches_oracle(k, h_frac):
    """Runs of N: a window of k N's hashes to 0 and is always kept, so a run
    longer than k + 127 keeps every window of a lane, far above the expected
    keep rate.  The rolling pass's per-lane staging overflows there and those
    chunks are written by the re-roll pass."""
    flank = _rng_seq(11, 20_000)
    s = (flank + b"N" * 5000 + flank[:9_999] + b"n" * 300 + flank[:1234] + b"N" * (k - 1) + flank[:77]
         + b"N" * k + flank[:500] + b"N" * (k + 128) + flank[:3000])
    o = O.sketch(k, h_frac, s, 0)
    lanes = np.bincount((o["r"] - k + 1) // 128)
    assert lanes.max() == 128  # at least one lane keeps all of its windows
    g = G.FracMinHash(k, h_frac).sketch_slice_into(s, 0)
    assert len(g) == len(o)
    assert np.array_equal(g["r"], o["r"]) and np.array_equal(g["h"], o["h"])
    assert np.array_equal(g["strand"], o["strand"])
    # reads with such runs through the mapper
    ref = [("chr1", flank + b"N" * 400 + flank[::-1])]
    reads = [s[1000:26_000], b"N" * 9000, s[15_000:32_000], flank[5000:15000]]
    prm = dict(theta=0.4, min_diff=0.075, max_overlap=0.3)
    _, _, ores = P.oracle_map(ref, reads, k, h_frac, **prm)
    _, _, _, gres = P.gpu_map(ref, reads, k, h_frac, **prm)
    bad = P.diff_results(ores, gres)
    assert not bad, bad


@pytest.mark.parametrize("max_matches", [None, 1, 3])
def test_index_counts_and_hits_match_oracle(max_matches):
    unit = _rng_seq(9, 171)
    segs = [("a", _rng_seq(10, 20_000) + unit * 30 + _rng_seq(11, 5000)),
            ("b", unit * 12 + _rng_seq(12, 8000)), ("c", b"ACG"), ("d", _rng_seq(13, 3000))]
    k, hf = 12, 0.3
    g = G.SketchIndex(G.FracMinHash(k, hf), segs, max_matches)
    o = O.OracleIndex(k, hf, segs, max_matches)
    hashes = set()
    for _, s in segs:
        hashes.update(int(h) for h in O.sketch(k, hf, s)["h"])
    hashes = sorted(hashes)[:4000] + [12345, 2**63 + 5]
    for h in hashes:
        assert g.count(h) == o.count(h), h
        if o.count(h):
            assert g.hits(h) == o.hits(h), h
    st_g, st_o = g.stats(), o.stats()
    for key in ("indexed_kmers", "indexed_hits", "blacklisted_kmers"):
        assert st_g[key] == st_o[key], key
    assert [s[2] for s in g.segments()] == [len(O.sketch(k, hf, s)) for _, s in segs]


@pytest.mark.parametrize("k,h_frac", [(25, 1.0), (25, 0.3), (130, 1.0), (130, 0.3)])
def test_many_short_segments_sketch_matches_oracle(k, h_frac):
    """One batch sketch over ~100 segments whose lengths sit at the unit edges
    (under k, k, k + 127, k + 128, ... windows in units of 128), so each wave
    packs lanes of several segments, each segment ends with its halo units
    (one for k = 25, two for k = 130), and the last one ends at the end of the
    set.  At h_frac 1.0 every lane keeps all of its windows, so the re-roll
    pass writes them.  Every occurrence of every hash must match the oracle."""
    edge = [0, 1, k - 1, k, k + 1, k + 126, k + 127, k + 128, k + 129, k + 255, k + 256, k + 300]
    r = np.random.default_rng(77)
    lens = edge * 6 + [int(x) for x in r.integers(1, 700, 30)] + [9000, k + 127]
    segs = [(f"s{i}", _rng_seq(1000 + i, n, b"ACGTacgtN")) for i, n in enumerate(lens)]
    g = G.SketchIndex(G.FracMinHash(k, h_frac), segs, None)
    o = O.OracleIndex(k, h_frac, segs, None)
    assert [s[2] for s in g.segments()] == [len(O.sketch(k, h_frac, s)) for _, s in segs]
    hashes = set()
    for _, s in segs:
        hashes.update(int(h) for h in O.sketch(k, h_frac, s)["h"])
    assert len(hashes) > 1000
    for h in sorted(hashes):
        assert g.count(h) == o.count(h), h
        assert g.hits(h) == o.hits(h), h


def _workload(seed, ref_len, n_reads, read_len, err, repeat=0.0, lognormal=False, sd=0.0, n_contigs=3):
    lens = [ref_len // n_contigs] * n_contigs
    ref = G.synth_reference(seed, lens, repeat_frac=repeat, unit_len=171, div_min=0.01, div_max=0.03,
                            n_units=4, array_len=20_000)
    bases, off, truth = G.synth_reads(seed + 1, ref, lens, n_reads, (read_len, sd), lognormal=lognormal,
                                      clip=(200, 6 * read_len), errors=err)
    return P.contigs(ref, lens), P.split_reads(bases, off)


WORKLOADS = {
    "uniform": dict(seed=11, ref_len=600_000, n_reads=150, read_len=3000, err=((0.01, 0.6, 0.2, 0.2),)),
    "repeats": dict(seed=21, ref_len=600_000, n_reads=150, read_len=4000, err=((0.002, 0.6, 0.2, 0.2),),
                    repeat=0.25, lognormal=True, sd=2000),
    "noisy": dict(seed=31, ref_len=400_000, n_reads=120, read_len=3000, err=((0.07, 0.4, 0.3, 0.3),)),
}


@pytest.mark.parametrize("wl", sorted(WORKLOADS))
@pytest.mark.parametrize("metric", ["Containment", "Jaccard", "bucket_SH", "bucket_LCS"])
def test_map_matches_oracle(wl, metric):
    segs, reads = _workload(**WORKLOADS[wl])
    prm = dict(theta=0.4, min_diff=0.075, max_overlap=0.3, metric=metric)
    oi, op, ores = P.oracle_map(segs, reads, 15, 0.1, **prm)
    gi, gp, mp, gres = P.gpu_map(segs, reads, 15, 0.1, **prm)
    assert ores["mapped"].sum() > 0 or wl == "noisy"
    bad = P.diff_results(ores, gres)
    assert not bad, bad
    for i in range(len(reads)):
        if ores[i]["mapped"]:
            a = fastx.strip_time(oi.format_paf(op, f"r{i}", ores[i]))
            b = fastx.strip_time(G.format_paf(gi, gp, f"r{i}", gres[i]))
            assert a == b


@pytest.mark.parametrize("h_frac", [0.05, 0.25])
def test_lcs_long_sketches_match_oracle(h_frac):
    """bucket_LCS on 20 kb reads: value sequences of about 1,000 (r=0.05) and
    5,000 (r=0.25) per bucket, beyond the register-resident tails and, at
    r=0.25, beyond the LDS slice."""
    segs, reads = _workload(seed=41, ref_len=300_000, n_reads=16, read_len=20_000, err=((0.001, 0.6, 0.2, 0.2),))
    prm = dict(theta=0.4, min_diff=0.075, max_overlap=0.3, metric="bucket_LCS")
    oi, op, ores = P.oracle_map(segs, reads, 15, h_frac, **prm)
    gi, gp, mp, gres = P.gpu_map(segs, reads, 15, h_frac, **prm)
    assert ores["mapped"].sum() > 0
    assert int(ores["m"].max()) > (4096 if h_frac > 0.1 else 512)
    bad = P.diff_results(ores, gres)
    assert not bad, bad


def test_lcs_repeat_runs_match_oracle():
    """bucket_LCS inside tandem arrays: HiFi reads of 15 kb in arrays of one
    171 bp unit (0.1-0.2% copy divergence), so a seed occurs 15-25 times in a
    read on average and the pmatches sequence is made of long runs
    (decreasing forward, increasing reversed): the runs the wave applies at
    once from the tails' lower bounds."""
    lens = [200_000, 200_000]
    ref = G.synth_reference(51, lens, repeat_frac=0.7, unit_len=171, div_min=0.001, div_max=0.002, n_units=1,
                            array_len=60_000)
    bases, off, _ = G.synth_reads(52, ref, lens, 14, (15_000, 0.0), errors=((0.001, 0.6, 0.2, 0.2),))
    segs, reads = P.contigs(ref, lens), P.split_reads(bases, off)
    prm = dict(theta=0.4, min_diff=0.075, max_overlap=0.3, metric="bucket_LCS")
    oi, op, ores = P.oracle_map(segs, reads, 15, 0.1, **prm)
    gi, gp, mp, gres = P.gpu_map(segs, reads, 15, 0.1, **prm)
    assert ores["mapped"].sum() > 0
    # reads whose k-mers repeat at least 16 times on average, within one wave's tails
    rep = (gres["u"] > 0) & (gres["m"] >= 16 * gres["u"]) & (gres["m"] <= 2048)  # noqa: E501
    assert rep.sum() >= 3, (gres["m"], gres["u"])
    bad = P.diff_results(ores, gres)
    assert not bad, bad


@pytest.mark.parametrize("variant", [dict(no_bucket_pruning=True), dict(one_sweep=True),
                                     dict(no_bucket_pruning=True, one_sweep=True), dict(abs_pos=True)])
def test_variants_match_oracle(variant):
    segs, reads = _workload(**WORKLOADS["uniform"])
    reads = reads[:60]
    prm = dict(theta=0.4, min_diff=0.075, max_overlap=0.3, **variant)
    oi, op, ores = P.oracle_map(segs, reads, 15, 0.1, **prm)
    gi, gp, mp, gres = P.gpu_map(segs, reads, 15, 0.1, **prm)
    bad = P.diff_results(ores, gres)
    assert not bad, bad


@pytest.mark.parametrize("metric", ["Containment", "Jaccard", "bucket_SH", "bucket_LCS"])
@pytest.mark.parametrize("max_matches", [1, 4])
def test_max_matches_maps_as_the_oracle(metric, max_matches):
    """-M erases frequent hashes from the index (src/index.rs:264-291,
    360-376): their seeds find no hits, but refinement still matches them in
    the reference sketch through the read's own hash table, as
    best_fixed_length and lcs_get_ppos_in_t do (src/shmap/scoring.rs:114-135,
    189-291).  Every field and the PAF text against the oracle."""
    segs, reads = _workload(**WORKLOADS["repeats"])
    prm = dict(theta=0.4, min_diff=0.075, max_overlap=0.3, metric=metric)
    oi, op, ores = P.oracle_map(segs, reads, 15, 0.1, max_matches=max_matches, **prm)
    gi, gp, mp, gres = P.gpu_map(segs, reads, 15, 0.1, max_matches=max_matches, **prm)
    bad = P.diff_results(ores, gres)
    assert not bad, bad
    assert ores["mapped"].sum() > 0
    for i in np.nonzero(ores["mapped"] == 1)[0]:
        assert G.format_paf(gi, gp, f"r{i}", gres[i]) == oi.format_paf(op, f"r{i}", ores[i])
    # the erasure changes the outcome somewhere, so the case is exercised
    _, _, plain = P.oracle_map(segs, reads, 15, 0.1, **prm)
    assert P.diff_results(plain, ores)


def test_short_read_floor():
    """tests/short_read_sketch_floor.rs: sketches below MIN_HALFLEN never map."""
    x, seq = 0x2545F4914F6CDD1D, []
    for _ in range(4000):
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        seq.append(b"ACGT"[x % 4])
    ref = bytes(seq)
    K = 8
    lens = [K, K + 1, K + 2, K + 3, K + 4, K + 12, K + 40, K + 20, K, K + 20]
    reads = [ref[100 + i * 7:100 + i * 7 + l] for i, l in enumerate(lens)]
    idx, p, mp, res = P.gpu_map([("ref", ref)], reads, K, 1.0, theta=0.4)
    assert list(res["mapped"]) == [0, 0, 0, 0, 1, 1, 1, 1, 0, 1]
    for r in res[res["mapped"] == 1]:
        assert 0 <= r["p_start"] < r["p_end"] <= r["read_len"]
    oi, op, ores = P.oracle_map([("ref", ref)], reads, K, 1.0, theta=0.4)
    assert not P.diff_results(ores, res)


def test_batch_equals_one_read_per_call():
    """The plain form (one read per launch) and the batched form agree."""
    segs, reads = _workload(**WORKLOADS["repeats"])
    reads = reads[:40]
    idx, p, mp, batch = P.gpu_map(segs, reads, 15, 0.1, theta=0.4, min_diff=0.075, max_overlap=0.3)
    single = np.concatenate([mp.map_batch([r]) for r in reads])
    assert not P.diff_results(batch, single)


def test_empty_and_degenerate_batches():
    segs, reads = _workload(**WORKLOADS["uniform"])
    idx, p, mp, res = P.gpu_map(segs, [], 15, 0.1, theta=0.4)
    assert len(res) == 0
    res = mp.map_batch([b"", b"ACGT", b"N" * 5000, reads[0]])
    assert list(res["mapped"][:3]) == [0, 0, 0]
    assert res["unmappable_short"][0] == 1
    # every read shorter than k: the sketch has no units at all
    res = mp.map_batch([b"", b"ACGT", b"A" * 14])
    assert list(res["mapped"]) == [0, 0, 0] and list(res["m"]) == [0, 0, 0]
    # and the mapper still works afterwards
    res = mp.map_batch([reads[0]])
    assert res["m"][0] > 0


def test_table_budget_groups_give_the_same_results():
    """A bucket-table budget far below the batch's need splits the reads into
    many groups on the host; every record stays the same."""
    segs, reads = _workload(**WORKLOADS["repeats"])
    prm = dict(theta=0.4, min_diff=0.075, max_overlap=0.3)
    gi, gp, mp, whole = P.gpu_map(segs, reads, 15, 0.1, **prm)
    old = os.environ.get("SHM_TAB_BUDGET")
    os.environ["SHM_TAB_BUDGET"] = "4096"
    try:
        small = G.Mapper(gi, gp)
    finally:
        if old is None:
            del os.environ["SHM_TAB_BUDGET"]
        else:
            os.environ["SHM_TAB_BUDGET"] = old
    grouped = small.map_batch(reads)
    assert small.last_counts()["groups"] > 5
    bad = P.diff_results(whole, grouped)
    assert not bad, bad

