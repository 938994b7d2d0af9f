"""A mapper's batching choices (shm_tuning) change where per-read work runs --
whole reads by a wave or a workgroup, or chunks through global tables, for the
bucket counts, tiny or long chunks, threads or waves for the prune, and how
many hash bits the seeding sort's one-word keys keep (with one or eight bits
nearly every read's k-mers need the ordering pass after the sort) -- and
never the records.  Extreme settings force every read down one path; each run
must equal the built-in choice and the oracle bit for bit.  The autotuner
keeps a valid tuning with the same records."""
import numpy as np
import pytest

import parity_util as P
import shmap_amd as G

pytestmark = pytest.mark.gpu

K, H_FRAC = 15, 0.1
PRM = dict(theta=0.4, min_diff=0.075, max_overlap=0.3)


@pytest.fixture(scope="module")
def setup():
    lens = [260_000, 140_000]
    ref = G.synth_reference(91, lens, repeat_frac=0.25, unit_len=171, div_min=0.01, div_max=0.03, n_units=4,
                            array_len=20_000)
    bases, off, _ = G.synth_reads(92, ref, lens, 160, (4000, 2000), lognormal=True, clip=(100, 20_000),
                                  errors=((0.001, 0.6, 0.2, 0.2), (0.06, 0.4, 0.3, 0.3)))
    segs = P.contigs(ref, lens)
    reads = P.split_reads(bases, off)
    idx = G.SketchIndex(G.FracMinHash(K, H_FRAC), segs, None)
    mp = G.Mapper(idx, G.make_params(k=K, h_frac=H_FRAC, **PRM))
    return segs, reads, mp


TUNINGS = [
    # chunks of one window through global tables, waves; one hash bit in the
    # seeding keys (nearly every read is ordered after its sort)
    dict(contrib_wave_hits=1, contrib_chunk_hits=64, prune_thread_seeds=1, contrib_wg_hits=1, seed_key_bits=1),
    # whole reads by one wave, threads
    dict(contrib_wave_hits=100_000, contrib_chunk_hits=4096, prune_thread_seeds=100_000, contrib_wg_hits=0,
         seed_key_bits=8),
    # every read with hits by a workgroup
    dict(contrib_wave_hits=1, contrib_chunk_hits=0, prune_thread_seeds=0, contrib_wg_hits=1_000_000,
         seed_key_bits=55),
    dict(contrib_wave_hits=200, contrib_chunk_hits=128, prune_thread_seeds=3, contrib_wg_hits=300, seed_key_bits=20),
]
BUILT_IN = dict(contrib_wave_hits=0, contrib_chunk_hits=0, prune_thread_seeds=0, contrib_wg_hits=0, seed_key_bits=0)


def test_tunings_give_the_same_records(setup):
    segs, reads, mp = setup
    mp.set_tuning(**BUILT_IN)
    want = mp.map_batch(reads)
    assert want["mapped"].sum() > 0
    counts = mp.last_counts()
    for t in TUNINGS:
        mp.set_tuning(**t)
        got = mp.map_batch(reads)
        bad = P.diff_results(want, got)
        assert not bad, (t, bad)
        if t["contrib_wg_hits"] == 1:
            # every read with hits took the chunked path
            assert mp.last_counts()["global_table_reads"] >= int((got["seed_matches"] > 1).sum())
        if t["contrib_wg_hits"] == 1_000_000:
            # every read took a workgroup's table (none has 896 buckets)
            assert mp.last_counts()["global_table_reads"] == 0
    mp.set_tuning(**BUILT_IN)
    assert counts["examined_hits"] > 0
    _, _, ores = P.oracle_map(segs, reads, K, H_FRAC, **PRM)
    bad = P.diff_results(ores, want)
    assert not bad, bad


def test_tuning_rejects_out_of_range_values(setup):
    _, _, mp = setup
    for bad in (dict(contrib_chunk_hits=100), dict(contrib_chunk_hits=8192), dict(prune_thread_seeds=-1),
                dict(seed_key_bits=56), dict(seed_key_bits=-1)):
        with pytest.raises(G.ShmError, match="tuning out of range"):
            mp.set_tuning(**bad)
    mp.set_tuning(**BUILT_IN)


def test_autotune_keeps_a_valid_tuning(setup):
    _, reads, mp = setup
    want = mp.map_batch(reads)
    t = mp.autotune(reads[:80])
    assert t == mp.tuning()
    assert t["prune_thread_seeds"] in (0, 4, 8, 16, 32, 64, 128)
    assert t["contrib_wave_hits"] in (0, 512, 1024, 2048, 4096, 8192)
    assert t["contrib_chunk_hits"] in (0, 512, 1024, 2048, 4096)
    assert t["contrib_wg_hits"] in (0, 1, 8192, 32768, 1 << 20)
    assert t["seed_key_bits"] == 0  # not searched: the built-in keys are the fastest
    got = mp.map_batch(reads)
    bad = P.diff_results(want, got)
    assert not bad, bad
    mp.set_tuning(**BUILT_IN)


def test_satellite_reads_take_every_counting_path():
    """Reads inside tandem arrays examine 10^4 - 10^5 hits each: with the
    workgroup bound raised, reads up to 16,384 examined hits are counted by
    four-wave workgroups and the longer ones by sixteen-wave workgroups (whose
    waves fill one shared table without reservations); with it at 1 they all
    go through chunks and global tables.  Every run equals the oracle."""
    lens = [300_000, 300_000]
    ref = G.synth_reference(71, lens, repeat_frac=0.7, unit_len=171, div_min=0.005, div_max=0.01, n_units=1,
                            array_len=100_000)
    bases, off, _ = G.synth_reads(72, ref, lens, 40, (12_000, 5_000), lognormal=True, clip=(1_000, 30_000),
                                  errors=((0.001, 0.6, 0.2, 0.2),))
    segs, reads = P.contigs(ref, lens), P.split_reads(bases, off)
    _, _, ores = P.oracle_map(segs, reads, K, H_FRAC, **PRM)
    E = ores["seed_matches"].astype(np.int64)
    assert (E > 16_384).sum() >= 3 and ((E > 1_024) & (E <= 16_384)).sum() >= 3, np.sort(E)
    idx = G.SketchIndex(G.FracMinHash(K, H_FRAC), segs, None)
    mp = G.Mapper(idx, G.make_params(k=K, h_frac=H_FRAC, **PRM))
    for t in (dict(contrib_wave_hits=1024, contrib_chunk_hits=0, prune_thread_seeds=0, contrib_wg_hits=1_000_000),
              dict(contrib_wave_hits=1, contrib_chunk_hits=512, prune_thread_seeds=0, contrib_wg_hits=1),
              dict(contrib_wave_hits=0, contrib_chunk_hits=0, prune_thread_seeds=0, contrib_wg_hits=0)):
        mp.set_tuning(**t)
        got = mp.map_batch(reads)
        bad = P.diff_results(ores, got)
        assert not bad, (t, bad)
    assert ores["mapped"].sum() > 0
