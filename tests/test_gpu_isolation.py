"""Failure isolation and index-file checks that need the device.

A read that fails does so alone, as a read that panics does in the
reference (src/shmap/mod.rs:219-238, 597-612): its record carries the status,
every other read of the batch is mapped exactly as the oracle maps it, in
shm_map_batch and in the serving loop.  An index file whose hit records
disagree with its sketch, or whose runs are out of order, is refused by the
device-side checks of shm_index_load."""
import ctypes as C
import time

import numpy as np
import pytest

import parity_util as P
import shmap_amd as G

pytestmark = pytest.mark.gpu

K, H_FRAC = 15, 0.1
PRM = dict(theta=0.4, min_diff=0.075, max_overlap=0.3)


@pytest.fixture(scope="module")
def setup():
    lens = [300_000, 200_000]
    ref = G.synth_reference(51, lens, repeat_frac=0.15, unit_len=171, div_min=0.01, div_max=0.03, n_units=4,
                            array_len=20_000)
    bases, off, _ = G.synth_reads(52, ref, lens, 120, (3000, 1500), lognormal=True, clip=(300, 9_000),
                                  errors=((0.002, 0.6, 0.2, 0.2), (0.07, 0.4, 0.3, 0.3)))
    segs = P.contigs(ref, lens)
    reads = P.split_reads(bases, off)
    # three reads over the 10 kb limit, among the others
    for i, at in ((5, 1000), (60, 150_000), (119, 250_000)):
        reads[i] = bytes(segs[0][1][at:at + 12_000 + i])
    _, _, ores = P.oracle_map(segs, reads, K, H_FRAC, **PRM)
    idx = G.SketchIndex(G.FracMinHash(K, H_FRAC), segs, None)
    return segs, reads, ores, idx


LIMIT = 10_000


def test_read_over_the_limit_fails_alone(setup):
    segs, reads, ores, idx = setup
    mp = G.Mapper(idx, G.make_params(k=K, h_frac=H_FRAC, **PRM))
    mp.set_read_limit(LIMIT)
    with pytest.raises(G.ShmError, match="read 5 failed"):
        mp.map_batch(reads)
    res = mp.map_batch(reads, allow_failed_reads=True)
    long_ = np.array([len(r) > LIMIT for r in reads])
    assert long_.sum() == 3
    assert np.all(res["status"][long_] == 4) and np.all(res["mapped"][long_] == 0)
    assert np.all(res["read_len"][long_] == [len(r) for r in reads if len(r) > LIMIT])
    assert np.all(res["status"][~long_] == 0)
    bad = P.diff_results(ores[~long_], res[~long_])
    assert not bad, bad
    assert res["mapped"][~long_].sum() > 0
    # the limit lifted, the same mapper maps them all as the oracle does
    mp.set_read_limit(0)
    bad = P.diff_results(ores, mp.map_batch(reads))
    assert not bad, bad


def test_malformed_read_fails_alone(setup):
    segs, reads, ores, idx = setup
    L = G.lib()
    mp = G.Mapper(idx, G.make_params(k=K, h_frac=H_FRAC, **PRM))
    sub = [r for r in reads if len(r) <= LIMIT][:40]
    arr, n, keep = G._reads(sub)
    arr[7].len = -3
    arr[20].seq, arr[20].len = None, 100
    out = np.zeros(n, dtype=G.RESULT_DTYPE)
    rc = L.shm_map_batch(mp._p, arr, n, out.ctypes.data)
    assert rc == 1 and "read 7 failed" in L.shm_last_error().decode()
    assert out["status"][7] == 1 and out["status"][20] == 1
    ok = np.ones(n, bool)
    ok[[7, 20]] = False
    _, _, o2 = P.oracle_map(segs, [r for i, r in enumerate(sub) if ok[i]], K, H_FRAC, **PRM)
    bad = P.diff_results(o2, out[ok])
    assert not bad, bad
    del keep


def test_serving_oversized_read_fails_alone(setup):
    segs, reads, ores, idx = setup
    srv = G.Server(idx, G.make_params(k=K, h_frac=H_FRAC, **PRM), workers=2, max_read_len=LIMIT)
    tags = np.arange(len(reads), dtype=np.uint64)
    for i in range(0, len(reads), 8):
        srv.submit_many(tags[i:i + 8], reads[i:i + 8])
        time.sleep(0.001)
    got = []
    while sum(len(g) for g in got) < len(reads):
        g = srv.poll(cap=4096, min_results=1, timeout_ms=20_000)
        assert len(g), "serving loop stalled"
        got.append(g)
    s = np.concatenate(got)
    s = s[np.argsort(s["tag"])]
    srv.close()
    r = s["r"]
    long_ = np.array([len(x) > LIMIT for x in reads])
    assert np.all(r["status"][long_] == 4) and np.all(r["status"][~long_] == 0)
    bad = P.diff_results(ores[~long_], r[~long_])
    assert not bad, bad


def _arrays(path):
    """File offsets and contents of a saved index's hit arrays."""
    raw = np.fromfile(path, dtype=np.uint8)
    hdr = raw[:192].tobytes()
    import struct
    nseg = struct.unpack_from("<I", hdr, 12)[0]
    n = struct.unpack_from("<q", hdr, 32)[0]
    name_bytes = struct.unpack_from("<Q", hdr, 184)[0]
    at = 192 + name_bytes + nseg * 4 + 2 * (nseg + 1) * 8
    o = {"t_h": at, "t_r": at + 8 * n, "t_s": at + 12 * n, "hit_g": at + 13 * n, "hit_r": at + 17 * n,
         "hit_seg": at + 21 * n}
    return raw, n, o


def test_index_file_with_inconsistent_hits_is_refused(setup, tmp_path):
    segs, reads, ores, idx = setup
    path = tmp_path / "ok.idx"
    idx.save(path)
    G.SketchIndex.load(path)  # the saved file passes every check
    raw, n, o = _arrays(path)
    view = lambda name, dt: raw[o[name]:o[name] + n * np.dtype(dt).itemsize].view(dt)
    # a hit whose position is not its k-mer's (strand unchanged: the host's
    # strand-prefix check still passes, the device check must refuse it)
    bad = raw.copy()
    bad[o["hit_r"]:o["hit_r"] + 4 * n].view(np.uint32)[10] ^= 1
    (tmp_path / "pos.idx").write_bytes(bad.tobytes())
    with pytest.raises(G.ShmError, match="inconsistent"):
        G.SketchIndex.load(tmp_path / "pos.idx")
    # two hits of one run swapped (same strand, so the host's checks pass)
    hg, hr, hs, th = view("hit_g", np.uint32), view("hit_r", np.uint32), view("hit_seg", np.uint32), view("t_h", np.uint64)
    same = np.nonzero((th[hg[:-1]] == th[hg[1:]]) & ((hr[:-1] >> 31) == (hr[1:] >> 31)))[0]
    assert len(same)
    i = int(same[0])
    bad = raw.copy()
    for name in ("hit_g", "hit_r", "hit_seg"):
        a = bad[o[name]:o[name] + 4 * n].view(np.uint32)
        a[i], a[i + 1] = a[i + 1], a[i]
    (tmp_path / "order.idx").write_bytes(bad.tobytes())
    with pytest.raises(G.ShmError, match="inconsistent"):
        G.SketchIndex.load(tmp_path / "order.idx")
