"""shm_map_batch_device: reads already in device memory, results into a
pageable array (through the library's device buffer) or into page-locked
memory from host_array (written by the final kernel directly).  Both must
equal shm_map_batch's records and the oracle's, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import parity_util as P
import shmap_amd as G

pytestmark = pytest.mark.gpu

K, H_FRAC = 15, 0.1
PRM = dict(theta=0.4, min_diff=0.075, max_overlap=0.3)


class _Dev:
    """Device copies of host arrays through the HIP runtime."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def upload(self, a: np.ndarray) -> int:
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 64)) == 0
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0
        self.ptrs.append(p)
        return p.value

    def free(self):
        for p in self.ptrs:
            self.L.hipFree(p)


@pytest.fixture(scope="module")
def setup():
    lens = [250_000, 150_000]
    ref = G.synth_reference(61, lens, repeat_frac=0.15, unit_len=171, div_min=0.01, div_max=0.03, n_units=4,
                            array_len=20_000)
    bases, off, _ = G.synth_reads(62, ref, lens, 180, (3500, 1500), lognormal=True, clip=(200, 15_000),
                                  errors=((0.002, 0.6, 0.2, 0.2), (0.06, 0.4, 0.3, 0.3)))
    segs = P.contigs(ref, lens)
    idx = G.SketchIndex(G.FracMinHash(K, H_FRAC), segs, None)
    mp = G.Mapper(idx, G.make_params(k=K, h_frac=H_FRAC, **PRM))
    return segs, bases, off, idx, mp


def test_device_entry_matches_batch_and_oracle(setup):
    segs, bases, off, idx, mp = setup
    n = len(off) - 1
    reads = P.split_reads(bases, off)
    want = mp.map_batch(reads)
    dev = _Dev()
    try:
        d_bases, d_off = dev.upload(bases), dev.upload(off)
        pageable = mp.map_device(d_bases, d_off, n).copy()
        pinned = G.host_array(n, G.RESULT_DTYPE)
        got = mp.map_device(d_bases, d_off, n, out=pinned).copy()
        # the same page-locked buffer again, for a prefix of the reads
        n2 = n // 3
        off2 = off[:n2 + 1].copy()
        d_off2 = dev.upload(off2)
        pinned[:] = np.zeros(1, dtype=G.RESULT_DTYPE)
        got2 = mp.map_device(d_bases, d_off2, n2, out=pinned).copy()
    finally:
        dev.free()
    for rec in (pageable, got):
        bad = P.diff_results(want, rec)
        assert not bad, bad
    bad = P.diff_results(want[:n2], got2)
    assert not bad, bad
    assert want["mapped"].sum() > 0
    _, _, ores = P.oracle_map(segs, reads, K, H_FRAC, **PRM)
    bad = P.diff_results(ores, got)
    assert not bad, bad


def test_device_entry_reads_past_2gib(setup):
    """Reads whose bases start beyond 2 GiB into the device buffer, at an
    unaligned byte: every offset the sketch forms is then >= 2^31 (a 32-bit
    sign extension of one would send its loads astray)."""
    segs, bases, off, idx, mp = setup
    n = len(off) - 1
    want = mp.map_batch(P.split_reads(bases, off))
    start = (1 << 31) + 12_345
    dev = _Dev()
    try:
        p = C.c_void_p()
        assert dev.L.hipMalloc(C.byref(p), start + bases.nbytes + 64) == 0
        dev.ptrs.append(p)
        assert dev.L.hipMemcpy(C.c_void_p(p.value + start), bases.ctypes.data, bases.nbytes, 1) == 0
        d_off = dev.upload((off + start).astype(np.int64))
        got = mp.map_device(p.value, d_off, n).copy()
    finally:
        dev.free()
    bad = P.diff_results(want, got)
    assert not bad, bad
    assert want["mapped"].sum() > 0
