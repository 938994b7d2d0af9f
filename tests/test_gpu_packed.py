"""shm_map_batch_packed: reads packed to 2 bits on the host (shm_pack_reads)
map exactly as the same reads given as bytes, and as the oracle maps them --
with runs of N, IUPAC codes and lower case among the bases, empty and short
reads, and a read over the mapper's length limit failing alone."""
import numpy as np
import pytest

import parity_util as P
import shmap_amd as G
import test_gpu_parity as T

pytestmark = pytest.mark.gpu


def _reads():
    segs, reads = T._workload(**T.WORKLOADS["repeats"])
    rng = np.random.default_rng(3)
    out = []
    for i, r in enumerate(reads):
        r = bytearray(r)
        if i % 4 == 1 and len(r) > 400:  # a run of N, an IUPAC code, lower case
            a, ln = int(rng.integers(0, len(r) - 300)), int(rng.integers(1, 250))
            r[a:a + ln] = b"N" * ln
            r[int(rng.integers(0, len(r)))] = ord("R")
            b = int(rng.integers(0, len(r) - 100))
            r[b:b + 100] = bytes(r[b:b + 100]).lower()
        out.append(bytes(r))
    out[3] = b""
    out[9] = out[9][:20]
    return segs, out


@pytest.mark.parametrize("metric", ["Containment", "bucket_LCS"])
def test_packed_reads_map_as_bytes_and_as_the_oracle(metric):
    segs, reads = _reads()
    prm = dict(theta=0.4, min_diff=0.075, max_overlap=0.3, metric=metric)
    oi, op, ores = P.oracle_map(segs, reads, 15, 0.1, **prm)
    gi, gp, mp, gres = P.gpu_map(segs, reads, 15, 0.1, **prm)
    pk = G.PackedReads(reads, threads=4)
    assert pk.n_exc > 10
    res = mp.map_packed(pk)
    assert not P.diff_results(gres, res)
    bad = P.diff_results(ores, res)
    assert not bad, bad
    assert ores["mapped"].sum() > 50
    # twice in a row on the same mapper, the second batch smaller
    small = G.PackedReads(reads[:40], threads=2)
    assert not P.diff_results(ores[:40], mp.map_packed(small))
    out = G.host_array(len(reads), G.RESULT_DTYPE)  # results straight into page-locked memory
    assert not P.diff_results(ores, mp.map_packed(pk, out=out))


def test_packed_read_over_the_limit_fails_alone():
    segs, reads = _reads()
    prm = dict(theta=0.4, min_diff=0.075, max_overlap=0.3)
    oi, op, ores = P.oracle_map(segs, reads, 15, 0.1, **prm)
    gi, gp, mp, gres = P.gpu_map(segs, reads, 15, 0.1, **prm)
    lens = np.array([len(r) for r in reads])
    limit = int(np.sort(lens)[-5])  # the four longest reads are over it
    mp.set_read_limit(limit)
    pk = G.PackedReads(reads, threads=3)
    with pytest.raises(G.ShmError, match="failed"):
        mp.map_packed(pk)
    out = np.zeros(len(reads), dtype=G.RESULT_DTYPE)
    rc = G.lib().shm_map_batch_packed(mp._p, pk.packed.ctypes.data, pk.off.ctypes.data, pk.n, pk.exc.ctypes.data,
                                       pk.n_exc, out.ctypes.data)
    over = lens > limit
    assert rc == 4 and over.sum() == 4
    assert np.all(out["status"][over] == 4) and np.all(out["status"][~over] == 0)
    assert not P.diff_results(ores[~over], out[~over])
