"""ctypes binding of the CPU oracle (oracle/lib/libshmap_oracle.so).

TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.  Only
tests/, the smoke check and the benchmark's CPU-baseline leg load this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libshmap_oracle.so")
CLI = os.path.join(HERE, "lib", "shmap_oracle")

METRICS = {"Containment": 0, "Jaccard": 1, "bucket_SH": 2, "bucket_LCS": 3}


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


class OkKmer(C.Structure):
    _fields_ = [("r", C.c_int32), ("h", C.c_uint64), ("strand", C.c_uint8)]


class OkParams(C.Structure):
    _fields_ = [("k", C.c_int32), ("h_frac", C.c_double), ("theta", C.c_double), ("min_diff", C.c_double),
                ("max_overlap", C.c_double), ("metric", C.c_int32), ("no_bucket_pruning", C.c_int32),
                ("one_sweep", C.c_int32), ("abs_pos", C.c_int32), ("max_matches", C.c_int32)]


_I32 = ["mapped", "unmappable_short", "read_len", "m", "u", "S"]
_I64 = ["possible_matches", "seed_matches", "total_matches", "max_seed_matches", "seeded_buckets",
        "refined_buckets", "refine_memo_hits", "final_buckets"]
_I32B = ["p_start", "p_end", "strand", "segm_id", "t_l", "t_r", "mapq", "s_sz", "intersection",
         "intersection2", "same_strand", "b_segm", "b_b", "b2_segm", "b2_b"]
_F64 = ["J", "J2", "sh", "sigmas", "match_inefficiency"]


class OkResult(C.Structure):
    _fields_ = ([(n, C.c_int32) for n in _I32] + [(n, C.c_int64) for n in _I64] +
                [(n, C.c_int32) for n in _I32B] + [(n, C.c_double) for n in _F64])


RESULT_DTYPE = np.dtype([(n, np.int32) for n in _I32] + [(n, np.int64) for n in _I64] +
                        [(n, np.int32) for n in _I32B] + [(n, np.float64) for n in _F64], align=True)
assert RESULT_DTYPE.itemsize == C.sizeof(OkResult)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.ok_threshold.restype = C.c_uint64
        L.ok_threshold.argtypes = [C.c_double]
        L.ok_sketch.restype = C.c_int64
        L.ok_sketch.argtypes = [C.c_int32, C.c_double, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int64]
        L.ok_index_build.restype = C.c_void_p
        L.ok_index_build.argtypes = [C.c_int32, C.c_double, C.c_int32, C.c_int32, C.POINTER(C.c_char_p),
                                     C.POINTER(C.c_void_p), C.POINTER(C.c_int64), C.c_int32]
        L.ok_index_free.argtypes = [C.c_void_p]
        L.ok_index_count.restype = C.c_int32
        L.ok_index_count.argtypes = [C.c_void_p, C.c_uint64]
        L.ok_index_hits.restype = C.c_int32
        L.ok_index_hits.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int32]
        L.ok_index_n_segments.restype = C.c_int32
        L.ok_index_n_segments.argtypes = [C.c_void_p]
        L.ok_index_segment_kmers.restype = C.c_int64
        L.ok_index_segment_kmers.argtypes = [C.c_void_p, C.c_int32]
        for f in ("ok_index_indexed_kmers", "ok_index_indexed_hits", "ok_index_blacklisted_kmers"):
            getattr(L, f).restype = C.c_int64
            getattr(L, f).argtypes = [C.c_void_p]
        L.ok_map_read.argtypes = [C.c_void_p, C.POINTER(OkParams), C.c_void_p, C.c_int32, C.POINTER(OkResult)]
        L.ok_map_reads.argtypes = [C.c_void_p, C.POINTER(OkParams), C.c_int32, C.POINTER(C.c_void_p),
                                   C.POINTER(C.c_int32), C.c_void_p, C.c_int32]
        L.ok_map_reads_counted.argtypes = [C.c_void_p, C.POINTER(OkParams), C.c_int32, C.POINTER(C.c_void_p),
                                           C.POINTER(C.c_int32), C.c_void_p, C.c_int32, C.POINTER(C.c_int64)]
        L.ok_format_paf.argtypes = [C.c_void_p, C.POINTER(OkParams), C.c_char_p, C.POINTER(OkResult), C.c_double,
                                    C.c_char_p, C.c_size_t]
        L.ok_hseed.restype = C.c_double
        L.ok_hseed.argtypes = [C.c_int32, C.c_int32, C.c_int32]
        L.ok_lis_len.restype = C.c_int32
        L.ok_lis_len.argtypes = [C.c_void_p, C.c_int32]
        L.ok_calc_mapq.restype = C.c_int32
        L.ok_calc_mapq.argtypes = [C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                   C.POINTER(C.c_double)]
        L.ok_unique_pmatches.restype = C.c_int32
        L.ok_unique_pmatches.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ok_overlap.restype = C.c_double
        L.ok_overlap.argtypes = [C.c_int32] * 6
        _lib = L
    return _lib


def params(k=15, h_frac=0.05, theta=0.9, min_diff=0.02, max_overlap=0.5, metric="Containment",
           no_bucket_pruning=False, one_sweep=False, abs_pos=False, max_matches=None) -> OkParams:
    return OkParams(k, h_frac, theta, min_diff, max_overlap, METRICS[metric] if isinstance(metric, str) else metric,
                    int(no_bucket_pruning), int(one_sweep), int(abs_pos), max_matches or -1)


def sketch(k: int, h_frac: float, seq, offset: int = 0) -> np.ndarray:
    a = np.frombuffer(seq, dtype=np.uint8) if isinstance(seq, (bytes, bytearray)) else seq
    n = lib().ok_sketch(k, h_frac, a.ctypes.data if a.size else None, len(a), offset, None, 0)
    out = (OkKmer * max(n, 1))()
    lib().ok_sketch(k, h_frac, a.ctypes.data if a.size else None, len(a), offset, out, n)
    res = np.empty(n, dtype=[("r", "<i4"), ("h", "<u8"), ("strand", "u1")])
    for i in range(n):
        res[i] = (out[i].r, out[i].h, out[i].strand)
    return res


class OracleIndex:
    def __init__(self, k: int, h_frac: float, segments, max_matches: int | None = None, threads: int = 1):
        n = len(segments)
        names = (C.c_char_p * max(n, 1))()
        seqs = (C.c_void_p * max(n, 1))()
        lens = (C.c_int64 * max(n, 1))()
        keep = []
        for i, (nm, s) in enumerate(segments):
            a = np.frombuffer(s, dtype=np.uint8) if isinstance(s, (bytes, bytearray)) else s
            b = nm.encode() if isinstance(nm, str) else nm
            keep.append((a, b))
            names[i], seqs[i], lens[i] = b, a.ctypes.data if a.size else None, len(a)
        self._p = lib().ok_index_build(k, h_frac, max_matches or -1, n, names, seqs, lens, threads)
        self.k, self.h_frac = k, h_frac

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.ok_index_free(self._p)
            self._p = None

    def count(self, h: int) -> int:
        return lib().ok_index_count(self._p, h)

    def hits(self, h: int):
        n = self.count(h)
        buf = (C.c_int32 * (4 * max(n, 1)))()
        lib().ok_index_hits(self._p, h, buf, n)
        return [tuple(buf[4 * i:4 * i + 4]) for i in range(n)]

    def stats(self) -> dict:
        L = lib()
        return {"indexed_kmers": L.ok_index_indexed_kmers(self._p), "indexed_hits": L.ok_index_indexed_hits(self._p),
                "blacklisted_kmers": L.ok_index_blacklisted_kmers(self._p)}

    def map_reads(self, prm: OkParams, reads, threads: int = 1) -> np.ndarray:
        """reads: list of bytes, or (bases uint8, offsets int64)."""
        return self.map_reads_timed(prm, reads, threads)[0]

    def map_reads_timed(self, prm: OkParams, reads, threads: int = 1):
        """(records, seconds spent inside ok_map_reads)."""
        if isinstance(reads, tuple):
            bases, off = reads
            n = len(off) - 1
            ptrs = (C.c_void_p * max(n, 1))(*[bases.ctypes.data + int(off[i]) for i in range(n)])
            lens = (C.c_int32 * max(n, 1))(*[int(off[i + 1] - off[i]) for i in range(n)])
            keep = bases
        else:
            n = len(reads)
            keep = [np.frombuffer(s, dtype=np.uint8) if isinstance(s, (bytes, bytearray)) else s for s in reads]
            ptrs = (C.c_void_p * max(n, 1))(*[a.ctypes.data if a.size else None for a in keep])
            lens = (C.c_int32 * max(n, 1))(*[len(a) for a in keep])
        out = np.zeros(max(n, 1), dtype=RESULT_DTYPE)
        ph = C.c_int64(0)
        t0 = time.perf_counter()
        lib().ok_map_reads_counted(self._p, C.byref(prm), n, ptrs, lens, out.ctypes.data, threads, C.byref(ph))
        dt = time.perf_counter() - t0
        self.last_prune_hits = ph.value  # hit records the prune pass read
        del keep
        return out[:n], dt

    def format_paf(self, prm: OkParams, qid: str, rec, map_time: float = 0.0) -> str:
        r = OkResult(*[rec[f].item() for f, _ in OkResult._fields_])
        buf = C.create_string_buffer(4096)
        lib().ok_format_paf(self._p, C.byref(prm), qid.encode(), C.byref(r), map_time, buf, 4096)
        return buf.value.decode()
