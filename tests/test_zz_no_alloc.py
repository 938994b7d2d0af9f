"""shm_sketch_noalloc, the allocation-free variant of shm_sketch, through the
C ABI: on two calls in a row (the second with a smaller input, so the fixed
scratch array still holds the first call's larger contents) its k-mers equal
shm_sketch's bit for bit, for several k and hash ratios, with non-ACGT bytes,
runs of N that overflow the per-lane staging, an offset, the null stream and
a stream of the caller's; a too-small output and a sequence beyond the array
are refused with SHM_ERR_CAPACITY."""
import ctypes as C

import numpy as np
import pytest

import shmap_amd as G

pytestmark = pytest.mark.gpu

KMER = np.dtype([("h", "<u8"), ("r", "<i4"), ("strand", "u1"), ("pad", "u1", 3)])


def _seq(seed, n, alphabet=b"ACGTacgtNRY"):
    r = np.random.default_rng(seed)
    return np.frombuffer(alphabet, dtype=np.uint8)[r.integers(0, len(alphabet), n)].tobytes()


def _with_runs(seed, n):
    s = bytearray(_seq(seed, n, b"ACGTACGTACGTacgtN"))
    for at in range(1000, n - 6000, n // 7):
        s[at:at + 5000] = b"N" * 5000  # whole lanes of kept windows: the staging overflows
    return bytes(s)


def reference(sk, seq, offset):
    """shm_sketch, the existing entry point, as raw shm_kmer records."""
    a = np.frombuffer(seq, dtype=np.uint8)
    cap = max(len(a), 1)
    out = np.zeros(cap, dtype=KMER)
    n = C.c_int64()
    rc = G.lib().shm_sketch(sk._p, a.ctypes.data, len(a), offset, out.ctypes.data, cap, C.byref(n))
    assert rc == 0, G.lib().shm_last_error()
    return out[:n.value]


def noalloc(sk, seq, offset, stream=None, cap=None):
    a = np.frombuffer(seq, dtype=np.uint8)
    cap = max(len(a), 1) if cap is None else cap
    out = np.zeros(max(cap, 1), dtype=KMER)
    out["h"] = 0xDEADBEEF  # anything the call leaves unwritten shows up
    n = C.c_int64()
    rc = G.lib().shm_sketch_noalloc(sk._p, a.ctypes.data, len(a), offset, out.ctypes.data, cap, C.byref(n), stream)
    return rc, out[:n.value] if rc == 0 else None, n.value


@pytest.mark.parametrize("k,h_frac", [(25, 0.01), (15, 0.1), (31, 0.5), (21, 1.0)])
def test_noalloc_matches_shm_sketch_twice(k, h_frac):
    sk = G.FracMinHash(k, h_frac)
    big = _with_runs(10 + k, 1_000_000)
    small = _seq(20 + k, 150_003)
    for seq, off in ((big, 0), (small, 17)):  # the second call is smaller
        want = reference(sk, seq, off)
        rc, got, n = noalloc(sk, seq, off)
        assert rc == 0, G.lib().shm_last_error()
        assert n == len(want) > 0
        assert got.tobytes() == want.tobytes()


def test_noalloc_on_a_caller_stream_and_edge_sizes():
    hip = C.CDLL("libamdhip64.so")
    st = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(st)) == 0
    try:
        sk = G.FracMinHash(25, 0.1)
        for n in (0, 24, 25, 26, 152, 153, 154, 8191, 8192 + 25):
            seq = _seq(n, n)
            want = reference(sk, seq, 5)
            rc, got, m = noalloc(sk, seq, 5, stream=st)
            assert rc == 0 and m == len(want), (n, G.lib().shm_last_error())
            assert got.tobytes() == want.tobytes(), n
    finally:
        hip.hipStreamDestroy(st)


def test_noalloc_refuses_what_does_not_fit():
    sk = G.FracMinHash(25, 0.1)
    seq = _seq(3, 200_000)
    want = reference(sk, seq, 0)
    rc, got, n = noalloc(sk, seq, 0, cap=len(want) - 1)
    assert rc == 4 and n == len(want)
    assert b"buffer holds" in G.lib().shm_last_error()
    # after a refused call the next one is whole
    rc, got, n = noalloc(sk, seq, 0)
    assert rc == 0 and got.tobytes() == want.tobytes()
    huge = bytes(30_000_000)
    assert G.lib().shm_sketch_noalloc_bytes(sk._p, len(huge), len(huge)) > (64 << 20)
    rc, _, _ = noalloc(sk, huge, 0)
    assert rc == 4 and b"fixed array holds" in G.lib().shm_last_error()
