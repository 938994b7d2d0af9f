"""Serving loop (shm_server_*): reads submitted while other batches are on the
GPU are admitted as workers free up, and each read's result is the same as when
it is mapped alone -- bit-exact, whatever batch and worker it lands in."""
import threading
import time

import numpy as np
import pytest

import parity_util as P
import shmap_amd as G

pytestmark = pytest.mark.gpu

K, H_FRAC = 15, 0.1
PRM = dict(theta=0.4, min_diff=0.075, max_overlap=0.3)


@pytest.fixture(scope="module")
def served_setup():
    lens = [300_000, 200_000]
    ref = G.synth_reference(41, lens, repeat_frac=0.2, unit_len=171, div_min=0.01, div_max=0.03, n_units=4,
                            array_len=20_000)
    # an even mix of HiFi-like and ONT-like error profiles, 300 .. 20,000 bases
    bases, off, _ = G.synth_reads(42, ref, lens, 240, (4000, 3000), lognormal=True, clip=(300, 20_000),
                                  errors=((0.001, 0.6, 0.2, 0.2), (0.07, 0.4, 0.3, 0.3)))
    segs = P.contigs(ref, lens)
    reads = P.split_reads(bases, off)
    reads[17] = b""                 # an empty read and a read below the sketch floor ride along
    reads[18] = reads[18][:20]
    sk = G.FracMinHash(K, H_FRAC)
    idx = G.SketchIndex(sk, segs, None)
    p = G.make_params(k=K, h_frac=H_FRAC, **PRM)
    mp = G.Mapper(idx, p)
    alone = np.concatenate([mp.map_batch([r]) for r in reads])
    return segs, reads, idx, p, alone


def _serve(idx, p, reads, submitters=3, **cfg):
    srv = G.Server(idx, p, **cfg)
    rng = np.random.default_rng(7)
    order = rng.permutation(len(reads))
    parts = np.array_split(order, submitters)

    def submit(part, seed):
        r = np.random.default_rng(seed)
        i = 0
        while i < len(part):
            step = int(r.integers(1, 9))
            chunk = part[i:i + step]
            srv.submit_many(chunk, [reads[j] for j in chunk])
            i += step
            time.sleep(float(r.uniform(0, 0.002)))

    th = [threading.Thread(target=submit, args=(part, s)) for s, part in enumerate(parts)]
    for t in th:
        t.start()
    got = []
    while sum(len(g) for g in got) < len(reads):
        got.append(srv.poll(cap=4096, min_results=1, timeout_ms=20_000))
        assert len(got[-1]) or any(t.is_alive() for t in th), "server stalled"
    for t in th:
        t.join()
    assert srv.error() == ""
    srv.close()
    served = np.concatenate(got)
    served = served[np.argsort(served["tag"])]
    assert np.array_equal(served["tag"], np.arange(len(reads)))
    return served


def test_results_alone_equal_results_in_the_mix(served_setup):
    segs, reads, idx, p, alone = served_setup
    served = _serve(idx, p, reads, workers=2, max_batch_reads=7)
    assert (served["r"]["status"] == 0).all()
    bad = P.diff_results(alone, served["r"])
    assert not bad, bad
    # the reads really were mixed: several batch sizes, both workers
    assert len(np.unique(served["batch_reads"])) > 1
    assert set(np.unique(served["worker"])) == {0, 1}
    # and the alone results are the oracle's
    _, _, ores = P.oracle_map(segs, reads, K, H_FRAC, **PRM)
    bad = P.diff_results(ores, alone)
    assert not bad, bad


def test_admission_budget_bounds_each_batch(served_setup):
    _, reads, idx, p, alone = served_setup
    budget = 30_000
    served = _serve(idx, p, reads, submitters=2, workers=3, max_batch_bases=budget)
    bad = P.diff_results(alone, served["r"])
    assert not bad, bad
    lens = np.array([len(r) for r in reads])
    # a batch is identified by (worker, start time); over budget only as a single read
    keys = {}
    for i, s in enumerate(served):
        keys.setdefault((int(s["worker"]), float(s["start_s"])), []).append(i)
    for members in keys.values():
        assert len(members) == served["batch_reads"][members[0]]
        tot = int(lens[members].sum())
        assert tot <= budget or len(members) == 1


def test_latency_fields_are_ordered(served_setup):
    _, reads, idx, p, _ = served_setup
    served = _serve(idx, p, reads[:60], submitters=1, workers=1, max_batch_reads=5)
    assert (served["submit_s"] <= served["start_s"]).all()
    assert (served["start_s"] <= served["done_s"]).all()
    assert (served["batch_reads"] <= 5).all()
