"""ctypes binding of the gfx950 mapping library (``include/shmap_hip.h``).

The classes mirror shmap-rs's library surface for the mapping path:
``FracMinHash`` (src/sketch.rs), ``SketchIndex`` (src/index.rs) and the mapper
of ``create_mapper`` / ``Mapper::map_reads`` (src/mapper.rs,
src/shmap/mod.rs).  Every call goes to ``lib/libshmap_hip.so``; when that
library is missing the import fails loudly -- there is no CPU fallback.

``synth_reference`` / ``synth_reads`` bind ``lib/libshmap_synth.so``, the
seeded generator of the synthetic workloads.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
HIP_LIB = os.path.join(LIB_DIR, "libshmap_hip.so")
SYNTH_LIB = os.path.join(LIB_DIR, "libshmap_synth.so")

CONTAINMENT, JACCARD, BUCKET_SH, BUCKET_LCS = 0, 1, 2, 3
METRICS = {"Containment": CONTAINMENT, "Jaccard": JACCARD, "bucket_SH": BUCKET_SH,
           "bucket_LCS": BUCKET_LCS}


class Kmer(C.Structure):
    _fields_ = [("h", C.c_uint64), ("r", C.c_int32), ("strand", C.c_uint8), ("pad", C.c_uint8 * 3)]


class Hit(C.Structure):
    _fields_ = [("r", C.c_int32), ("tpos", C.c_int32), ("strand", C.c_int32), ("segm_id", C.c_int32)]


class Segment(C.Structure):
    _fields_ = [("name", C.c_char_p), ("seq", C.c_void_p), ("len", C.c_int64)]


class Read(C.Structure):
    _fields_ = [("seq", C.c_void_p), ("len", C.c_int32)]


class MapParams(C.Structure):
    _fields_ = [("k", C.c_int32), ("h_frac", C.c_double), ("theta", C.c_double),
                ("min_diff", C.c_double), ("max_overlap", C.c_double), ("metric", C.c_int32),
                ("no_bucket_pruning", C.c_int32), ("one_sweep", C.c_int32), ("abs_pos", C.c_int32)]


_RESULT_INT_FIELDS = [
    "read_len", "m", "u", "S", "mapped", "unmappable_short", "possible_matches", "seed_matches",
    "total_matches", "max_seed_matches", "seeded_buckets", "refined_buckets", "refine_memo_hits",
    "final_buckets", "p_start", "p_end", "strand", "segm_id", "t_l", "t_r", "mapq", "s_sz",
    "intersection", "intersection2", "same_strand", "b_segm", "b_b", "b2_segm", "b2_b", "status",
    "has_second"]
_RESULT_F64_FIELDS = ["J", "J2", "sh", "sigmas", "match_inefficiency"]


class ReadResult(C.Structure):
    _fields_ = [(n, C.c_int32) for n in _RESULT_INT_FIELDS] + [(n, C.c_double) for n in _RESULT_F64_FIELDS]


RESULT_DTYPE = np.dtype([(n, np.int32) for n in _RESULT_INT_FIELDS] + [(n, np.float64) for n in _RESULT_F64_FIELDS],
                        align=True)
assert RESULT_DTYPE.itemsize == C.sizeof(ReadResult)


class ServerConfig(C.Structure):
    _fields_ = [("workers", C.c_int32), ("max_batch_reads", C.c_int32), ("max_batch_bases", C.c_int64),
                ("max_read_len", C.c_int32), ("flags", C.c_int32)]

SERVER_NO_SPECULATION, SERVER_NO_GRAPHS = 1, 2  # ServerConfig.flags


class Served(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("submit_s", C.c_double), ("start_s", C.c_double), ("done_s", C.c_double),
                ("batch_reads", C.c_int32), ("worker", C.c_int32), ("r", ReadResult)]


SERVED_DTYPE = np.dtype([("tag", np.uint64), ("submit_s", np.float64), ("start_s", np.float64),
                         ("done_s", np.float64), ("batch_reads", np.int32), ("worker", np.int32),
                         ("r", RESULT_DTYPE)], align=True)
assert SERVED_DTYPE.itemsize == C.sizeof(Served)


class ShmError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded libshmap_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(HIP_LIB):
            raise ShmError(f"{HIP_LIB} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
        L = C.CDLL(HIP_LIB)
        L.shm_last_error.restype = C.c_char_p
        L.shm_mapper_stage_name.restype = C.c_char_p
        L.shm_sketcher_create.argtypes = [C.c_int32, C.c_double, C.POINTER(C.c_void_p)]
        L.shm_sketcher_free.argtypes = [C.c_void_p]
        L.shm_sketch.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int64,
                                 C.POINTER(C.c_int64)]
        L.shm_sketch_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_int64,
                                         C.POINTER(C.c_int64), C.c_void_p]
        L.shm_sketch_noalloc_bytes.argtypes = [C.c_void_p, C.c_int64, C.c_int64]
        L.shm_sketch_noalloc_bytes.restype = C.c_int64
        L.shm_index_build.argtypes = [C.c_void_p, C.POINTER(Segment), C.c_int32, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_void_p)]
        L.shm_index_free.argtypes = [C.c_void_p]
        L.shm_index_count.argtypes = [C.c_void_p, C.c_uint64, C.POINTER(C.c_int32)]
        L.shm_index_hits.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]
        L.shm_index_n_segments.argtypes = [C.c_void_p, C.POINTER(C.c_int32)]
        L.shm_index_segment.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_char_p), C.POINTER(C.c_int32),
                                        C.POINTER(C.c_int64)]
        L.shm_index_stats.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 5
        L.shm_index_params.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int32)]
        L.shm_index_save.argtypes = [C.c_void_p, C.c_char_p]
        L.shm_index_load.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]
        L.shm_mapper_create.argtypes = [C.c_void_p, C.POINTER(MapParams), C.POINTER(C.c_void_p)]
        L.shm_mapper_free.argtypes = [C.c_void_p]
        L.shm_mapper_set_timing.argtypes = [C.c_void_p, C.c_int32]
        L.shm_mapper_set_speculation.argtypes = [C.c_void_p, C.c_int32]
        L.shm_mapper_set_graphs.argtypes = [C.c_void_p, C.c_int32]
        L.shm_device_compute_units.argtypes = [C.c_int32, C.POINTER(C.c_int32)]
        L.shm_mapper_set_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.shm_mapper_get_tuning.argtypes = [C.c_void_p, C.POINTER(Tuning)]
        L.shm_mapper_autotune.argtypes = [C.c_void_p, C.POINTER(Read), C.c_int32, C.POINTER(Tuning)]
        L.shm_map_batch.argtypes = [C.c_void_p, C.POINTER(Read), C.c_int32, C.c_void_p]
        L.shm_mapper_set_read_limit.argtypes = [C.c_void_p, C.c_int32]
        L.shm_pack_reads.argtypes = [C.POINTER(Read), C.c_int32, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                     C.c_int64, C.POINTER(C.c_int64), C.c_int32]
        L.shm_map_batch_packed.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_int64,
                                           C.c_void_p]
        L.shm_map_batch_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]
        L.shm_format_paf.argtypes = [C.c_void_p, C.POINTER(MapParams), C.c_char_p, C.POINTER(ReadResult),
                                     C.c_double, C.c_char_p, C.c_size_t]
        L.shm_format_unmapped.argtypes = [C.c_char_p, C.c_int32, C.c_char_p, C.c_size_t]
        L.shm_mapper_last_timing.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                             C.c_int32, C.POINTER(C.c_int32)]
        L.shm_mapper_stage_name.argtypes = [C.c_int32]
        L.shm_mapper_last_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32)]
        L.shm_host_alloc.argtypes = [C.c_size_t, C.POINTER(C.c_void_p)]
        L.shm_server_create.argtypes = [C.c_void_p, C.POINTER(MapParams), C.POINTER(ServerConfig),
                                        C.POINTER(C.c_void_p)]
        L.shm_server_submit.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p, C.c_int32]
        L.shm_server_submit_many.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.shm_server_poll.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_int32)]
        L.shm_server_counts.argtypes = [C.c_void_p] + [C.POINTER(C.c_int64)] * 3
        L.shm_server_batch_stats.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.c_int32]
        L.shm_server_now.argtypes = [C.c_void_p]
        L.shm_server_now.restype = C.c_double
        L.shm_server_error.argtypes = [C.c_void_p]
        L.shm_server_error.restype = C.c_char_p
        L.shm_server_free.argtypes = [C.c_void_p]
        L.shm_host_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def device_cus(device: int = 0) -> int:
    """Compute units of the device (shm_device_compute_units)."""
    n = C.c_int32()
    _check(lib().shm_device_compute_units(device, C.byref(n)))
    return n.value


class Tuning(C.Structure):
    """shm_tuning: a mapper's batching choices (0: the built-in rule)."""
    _fields_ = [("contrib_wave_hits", C.c_int32), ("contrib_chunk_hits", C.c_int32),
                ("prune_thread_seeds", C.c_int32), ("contrib_wg_hits", C.c_int32), ("seed_key_bits", C.c_int32)]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


def _check(rc: int) -> None:
    if rc != 0:
        raise ShmError(f"shmap_hip error {rc}: {lib().shm_last_error().decode()}")


def _buf(data: bytes | np.ndarray):
    a = np.frombuffer(data, dtype=np.uint8) if isinstance(data, (bytes, bytearray)) else data
    return a, a.ctypes.data if a.size else None


class FracMinHash:
    """src/sketch.rs:34-219 -- the sketcher; sketches run on the GPU."""

    def __init__(self, k: int, h_frac: float):
        self.k, self.h_frac = k, h_frac
        p = C.c_void_p()
        _check(lib().shm_sketcher_create(k, h_frac, C.byref(p)))
        self._p = p

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.shm_sketcher_free(self._p)
            self._p = None

    def sketch_slice_into(self, seq: bytes, offset: int = 0) -> np.ndarray:
        """Returns a structured array (r, h, strand) in position order."""
        a, ptr = _buf(seq)
        n = C.c_int64()
        cap = max(len(a), 1)
        out = (Kmer * cap)()
        _check(lib().shm_sketch(self._p, ptr, len(a), offset, out, cap, C.byref(n)))
        arr = np.frombuffer(out, dtype=np.dtype([("h", "<u8"), ("r", "<i4"), ("strand", "u1"), ("pad", "u1", 3)]),
                            count=n.value)
        res = np.empty(n.value, dtype=[("r", "<i4"), ("h", "<u8"), ("strand", "u1")])
        res["r"], res["h"], res["strand"] = arr["r"], arr["h"], arr["strand"]
        return res

    def sketch(self, seq: bytes) -> np.ndarray:
        return self.sketch_slice_into(seq, 0)


class SketchIndex:
    """src/index.rs:112-701 -- the reference index, built and kept on the GPU
    (or read back from a file written by save(), with load())."""

    def __init__(self, sketcher: FracMinHash, segments, max_matches: int | None = None, device: int = 0):
        self.sketcher = sketcher
        segs = (Segment * max(len(segments), 1))()
        self._keep = []
        for i, (name, seq) in enumerate(segments):
            a, ptr = _buf(seq)
            nm = name.encode() if isinstance(name, str) else name
            self._keep.append((a, nm))
            segs[i].name, segs[i].seq, segs[i].len = nm, ptr, len(a)
        p = C.c_void_p()
        _check(lib().shm_index_build(sketcher._p, segs, len(segments), max_matches or 0, device, C.byref(p)))
        self._p = p
        self._keep = None

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.shm_index_free(self._p)
            self._p = None

    def count(self, h: int) -> int:
        c = C.c_int32()
        _check(lib().shm_index_count(self._p, h, C.byref(c)))
        return c.value

    def hits(self, h: int) -> list[tuple[int, int, int, int]]:
        """(r, tpos, strand, segm_id) of every hit, in (segm_id, r) order."""
        n = C.c_int32()
        cnt = self.count(h)
        out = (Hit * max(cnt, 1))()
        _check(lib().shm_index_hits(self._p, h, out, max(cnt, 1), C.byref(n)))
        return [(out[i].r, out[i].tpos, out[i].strand, out[i].segm_id) for i in range(n.value)]

    def segments(self):
        n = C.c_int32()
        _check(lib().shm_index_n_segments(self._p, C.byref(n)))
        res = []
        for s in range(n.value):
            nm, sz, nk = C.c_char_p(), C.c_int32(), C.c_int64()
            _check(lib().shm_index_segment(self._p, s, C.byref(nm), C.byref(sz), C.byref(nk)))
            res.append((nm.value.decode(), sz.value, nk.value))
        return res

    def save(self, path) -> None:
        """Writes the index to a file (shm_index_save)."""
        _check(lib().shm_index_save(self._p, os.fsencode(path)))

    @classmethod
    def load(cls, path, device: int = 0) -> "SketchIndex":
        """An index read from a file written by save() (shm_index_load)."""
        p = C.c_void_p()
        _check(lib().shm_index_load(os.fsencode(path), device, C.byref(p)))
        self = cls.__new__(cls)
        self._p, self._keep = p, None
        k, r, mm = self.params()
        self.sketcher = FracMinHash(k, r)
        return self

    def params(self) -> tuple[int, float, int]:
        """(k, h_frac, max_matches) the index was built with (0: -M unset)."""
        k, r, mm = C.c_int32(), C.c_double(), C.c_int32()
        _check(lib().shm_index_params(self._p, C.byref(k), C.byref(r), C.byref(mm)))
        return k.value, r.value, mm.value

    def stats(self) -> dict:
        v = [C.c_int64() for _ in range(5)]
        _check(lib().shm_index_stats(self._p, *[C.byref(x) for x in v]))
        keys = ["indexed_kmers", "indexed_hits", "indexed_highest_freq_kmer", "blacklisted_kmers",
                "blacklisted_hits"]
        return dict(zip(keys, [x.value for x in v]))


def make_params(k=15, h_frac=0.05, theta=0.9, min_diff=0.02, max_overlap=0.5, metric="Containment",
                no_bucket_pruning=False, one_sweep=False, abs_pos=False) -> MapParams:
    """Defaults are the reference CLI's (src/params.rs:28-110)."""
    m = METRICS[metric] if isinstance(metric, str) else int(metric)
    return MapParams(k, h_frac, theta, min_diff, max_overlap, m, int(no_bucket_pruning), int(one_sweep),
                     int(abs_pos))


def _reads(reads):
    """(shm_read array, count, objects to keep alive) for a list of bytes or
    a (bases uint8 array, offsets int64 array) pair."""
    if isinstance(reads, tuple):
        bases, off = reads
        n = len(off) - 1
        arr = (Read * max(n, 1))()
        base_ptr = bases.ctypes.data
        for i in range(n):
            arr[i].seq, arr[i].len = base_ptr + int(off[i]), int(off[i + 1] - off[i])
        return arr, n, bases
    n = len(reads)
    arr = (Read * max(n, 1))()
    keep = []
    for i, s in enumerate(reads):
        a, ptr = _buf(s)
        keep.append(a)
        arr[i].seq, arr[i].len = ptr, len(a)
    return arr, n, keep


class Mapper:
    """create_mapper + Mapper::map_reads (src/mapper.rs:15-39) over a GPU index."""

    def __init__(self, index: SketchIndex, params: MapParams):
        self.index, self.params = index, params
        p = C.c_void_p()
        _check(lib().shm_mapper_create(index._p, C.byref(params), C.byref(p)))
        self._p = p

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.shm_mapper_free(self._p)
            self._p = None

    def set_timing(self, on: bool) -> None:
        _check(lib().shm_mapper_set_timing(self._p, int(on)))

    def set_speculation(self, on: bool) -> None:
        """Speculative batches (one host synchronisation each) on or off."""
        _check(lib().shm_mapper_set_speculation(self._p, int(on)))

    def set_graphs(self, on: bool) -> None:
        """Graph capture of speculative batches, per read class, on or off."""
        _check(lib().shm_mapper_set_graphs(self._p, int(on)))

    def set_tuning(self, **fields) -> None:
        """shm_mapper_set_tuning: contrib_wave_hits, contrib_chunk_hits,
        prune_thread_seeds, contrib_wg_hits, seed_key_bits (0: the built-in
        rule); unnamed fields keep their values."""
        t = Tuning()
        _check(lib().shm_mapper_get_tuning(self._p, C.byref(t)))
        for k, v in fields.items():
            setattr(t, k, int(v))
        _check(lib().shm_mapper_set_tuning(self._p, C.byref(t)))

    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().shm_mapper_get_tuning(self._p, C.byref(t)))
        return t.as_dict()

    def autotune(self, reads) -> dict:
        """shm_mapper_autotune on a sample batch (as map_batch takes it);
        returns the tuning kept."""
        arr, n, keep = _reads(reads)
        t = Tuning()
        _check(lib().shm_mapper_autotune(self._p, arr, n, C.byref(t)))
        del keep
        return t.as_dict()

    def set_read_limit(self, max_read_len: int) -> None:
        """shm_mapper_set_read_limit: longer reads fail alone (0: no limit)."""
        _check(lib().shm_mapper_set_read_limit(self._p, int(max_read_len)))

    def map_batch(self, reads, allow_failed_reads: bool = False) -> np.ndarray:
        """reads: list of bytes (or (bases uint8 array, offsets int64 array)).
        Returns RESULT_DTYPE records in input order.  A read that fails alone
        has a nonzero `status`; with allow_failed_reads the records are
        returned then (the others are complete), otherwise it raises."""
        arr, n, keep = _reads(reads)
        out = np.zeros(max(n, 1), dtype=RESULT_DTYPE)
        rc = lib().shm_map_batch(self._p, arr, n, out.ctypes.data)
        del keep
        if rc and not (allow_failed_reads and out[:n]["status"].any()):
            _check(rc)
        return out[:n]

    def map_packed(self, batch: "PackedReads", out: np.ndarray | None = None) -> np.ndarray:
        """shm_map_batch_packed over a PackedReads batch."""
        n = batch.n
        if out is None:
            out = np.zeros(max(n, 1), dtype=RESULT_DTYPE)
        elif out.dtype != RESULT_DTYPE or len(out) < n or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous RESULT_DTYPE array of at least n records")
        _check(lib().shm_map_batch_packed(self._p, batch.packed.ctypes.data, batch.off.ctypes.data, n,
                                          batch.exc.ctypes.data, batch.n_exc, out.ctypes.data))
        return out[:n]

    def map_device(self, d_bases_ptr: int, d_off_ptr: int, n: int, out: np.ndarray | None = None) -> np.ndarray:
        """Maps n reads whose bases and offsets are already in device memory.
        `out` (RESULT_DTYPE, at least n records; page-locked from host_array for
        full copy speed) receives the results; a new array otherwise."""
        if out is None:
            out = np.zeros(max(n, 1), dtype=RESULT_DTYPE)
        elif out.dtype != RESULT_DTYPE or len(out) < n or not out.flags.c_contiguous:
            raise ValueError("out must be a contiguous RESULT_DTYPE array of at least n records")
        _check(lib().shm_map_batch_device(self._p, d_bases_ptr, d_off_ptr, n, out.ctypes.data))
        return out[:n]

    def last_timing(self) -> tuple[float, dict]:
        tot = C.c_double()
        arr = (C.c_double * 32)()
        nw = C.c_int32()
        _check(lib().shm_mapper_last_timing(self._p, C.byref(tot), arr, 32, C.byref(nw)))
        return tot.value, {lib().shm_mapper_stage_name(i).decode(): arr[i] for i in range(nw.value)}

    def last_counts(self) -> dict:
        """Work counts of the last batch (k-mers, seeds, buckets, refined buckets, groups)."""
        arr = (C.c_int64 * len(COUNT_NAMES))()
        nw = C.c_int32()
        _check(lib().shm_mapper_last_counts(self._p, arr, len(COUNT_NAMES), C.byref(nw)))
        return {COUNT_NAMES[i]: int(arr[i]) for i in range(min(nw.value, len(COUNT_NAMES)))}


class Server:
    """Serving loop over a resident index (shm_server_*): reads are submitted as
    they arrive and mapped in batches by `workers` concurrent mappers."""

    def __init__(self, index: SketchIndex, params: MapParams, workers: int = 3, max_batch_reads: int = 0,
                 max_batch_bases: int = 0, max_read_len: int = 0, flags: int = 0):
        self._index = index
        self._params = params
        cfg = ServerConfig(workers, max_batch_reads, max_batch_bases, max_read_len, flags)
        self._p = C.c_void_p()
        _check(lib().shm_server_create(index._p, C.byref(params), C.byref(cfg), C.byref(self._p)))

    def now(self) -> float:
        return lib().shm_server_now(self._p)

    def submit(self, tag: int, seq: bytes) -> None:
        _check(lib().shm_server_submit(self._p, tag, seq, len(seq)))

    def submit_many(self, tags, seqs) -> None:
        tags = np.ascontiguousarray(tags, dtype=np.uint64)
        keep = [bytes(x) for x in seqs]
        arr = (Read * max(len(keep), 1))()
        for i, x in enumerate(keep):
            arr[i].seq = C.cast(C.c_char_p(x), C.c_void_p)
            arr[i].len = len(x)
        _check(lib().shm_server_submit_many(self._p, tags.ctypes.data, arr, len(keep)))

    def poll(self, cap: int = 65536, min_results: int = 1, timeout_ms: int = -1) -> np.ndarray:
        out = np.zeros(max(cap, 1), dtype=SERVED_DTYPE)
        n = C.c_int32()
        _check(lib().shm_server_poll(self._p, out.ctypes.data, cap, min_results, timeout_ms, C.byref(n)))
        return out[:n.value]

    def counts(self) -> tuple[int, int, int]:
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().shm_server_counts(self._p, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def batch_stats(self) -> dict:
        """Batches so far: all, speculative ones mapped again, ones that
        allocated memory, graph captures."""
        a = (C.c_int64 * 4)()
        _check(lib().shm_server_batch_stats(self._p, a, 4))
        return dict(zip(("batches", "redone", "allocating", "captures"), (int(x) for x in a)))

    def error(self) -> str:
        return lib().shm_server_error(self._p).decode()

    def close(self) -> None:
        if getattr(self, "_p", None) and _lib is not None:
            _lib.shm_server_free(self._p)
            self._p = None

    def __del__(self):
        self.close()


class _HostBlock:
    """Owner of one page-locked allocation; freed when the last view goes."""

    def __init__(self, nbytes: int):
        self.ptr = C.c_void_p()
        _check(lib().shm_host_alloc(max(nbytes, 1), C.byref(self.ptr)))

    def __del__(self):
        if self.ptr:
            lib().shm_host_free(self.ptr)
            self.ptr = C.c_void_p()


def host_array(n: int, dtype) -> np.ndarray:
    """A numpy array of n items in page-locked host memory (shm_host_alloc)."""
    dt = np.dtype(dtype)
    blk = _HostBlock(n * dt.itemsize)
    raw = (C.c_uint8 * max(n * dt.itemsize, 1)).from_address(blk.ptr.value)
    raw._owner = blk  # keep the block alive as long as any view of it
    return np.frombuffer(raw, dtype=dt, count=n)


COUNT_NAMES = ("kmers", "seeds", "buckets", "refined", "groups", "first_pass_candidates", "second_pass_band",
               "examined_hits", "global_table_reads", "host_syncs", "redone", "allocations", "graph", "captured")


class PackedReads:
    """A batch packed by shm_pack_reads: 2-bit bases, base offsets and the
    runs of non-ACGT bytes, in page-locked memory when pinned (full link speed)."""

    def __init__(self, reads, threads: int = 8, pinned: bool = True):
        arr, n, keep = _reads(reads)
        total = sum(arr[i].len for i in range(n))
        nbytes = (total + 3) // 4
        alloc = (lambda m, dt: host_array(m, dt)) if pinned else (lambda m, dt: np.zeros(m, dtype=dt))
        self.packed = alloc(max(nbytes, 1) + 4, np.uint8)
        self.off = alloc(n + 1, np.int64)
        cap = 1024
        while True:
            self.exc = alloc(2 * cap, np.int64)
            ne = C.c_int64()
            rc = lib().shm_pack_reads(arr, n, self.packed.ctypes.data, len(self.packed), self.off.ctypes.data,
                                      self.exc.ctypes.data, cap, C.byref(ne), threads)
            if rc == 4 and ne.value > cap:
                cap = ne.value
                continue
            _check(rc)
            break
        del keep
        self.n, self.n_exc, self.bases = n, ne.value, total


def format_paf(index: SketchIndex, params: MapParams, qid: str, rec, map_time: float = 0.0) -> str:
    r = ReadResult(*[int(rec[n]) for n in _RESULT_INT_FIELDS], *[float(rec[n]) for n in _RESULT_F64_FIELDS])
    buf = C.create_string_buffer(4096)
    n = lib().shm_format_paf(index._p, C.byref(params), qid.encode(), C.byref(r), map_time, buf, 4096)
    if n < 0:
        raise ShmError(lib().shm_last_error().decode())
    return buf.value.decode()


def format_unmapped(qid: str, read_len: int) -> str:
    buf = C.create_string_buffer(4096)
    lib().shm_format_unmapped(qid.encode(), read_len, buf, 4096)
    return buf.value.decode()


# ---------------------------------------------------------------------------
# synthetic workloads
# ---------------------------------------------------------------------------

_synth = None


def synth_lib():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_LIB):
            raise ShmError(f"{SYNTH_LIB} is missing: build it first")
        L = C.CDLL(SYNTH_LIB)
        L.synth_reference.argtypes = [C.c_uint64, C.c_int32, C.c_void_p, C.c_double, C.c_int32, C.c_double,
                                      C.c_double, C.c_int32, C.c_int64, C.c_void_p, C.c_int32]
        L.synth_reads.restype = C.c_int64
        L.synth_reads.argtypes = [C.c_uint64, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                  C.c_double, C.c_double, C.c_int32, C.c_int32, C.c_void_p, C.c_int32,
                                  C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        _synth = L
    return _synth


def synth_reference(seed: int, lens, repeat_frac=0.0, unit_len=171, div_min=0.02, div_max=0.02, n_units=16,
                    array_len=100_000, threads: int | None = None) -> np.ndarray:
    lens = np.asarray(lens, dtype=np.int64)
    out = np.empty(int(lens.sum()), dtype=np.uint8)
    synth_lib().synth_reference(seed, len(lens), lens.ctypes.data, repeat_frac, unit_len, div_min, div_max,
                                n_units, array_len, out.ctypes.data, threads or min(os.cpu_count() or 1, 16))
    return out


def synth_reads(seed: int, ref: np.ndarray, lens, n_reads: int, length=(10_000, 0.0), lognormal=False,
                clip=(1, 1 << 30), errors=((0.0, 1.0, 0.0, 0.0),), both_strands=True,
                threads: int | None = None):
    """Returns (bases uint8, offsets int64[n+1], truth int32[n,4] = contig, start, end, strand)."""
    lens = np.asarray(lens, dtype=np.int64)
    coff = np.zeros(len(lens) + 1, dtype=np.int64)
    coff[1:] = np.cumsum(lens)
    err = np.asarray(errors, dtype=np.float64).reshape(-1, 4)
    off = np.zeros(n_reads + 1, dtype=np.int64)
    th = threads or min(os.cpu_count() or 1, 16)
    L = synth_lib()
    args = (seed, ref.ctypes.data, coff.ctypes.data, len(lens), n_reads, 1 if lognormal else 0,
            float(length[0]), float(length[1]), int(clip[0]), int(clip[1]), err.ctypes.data, len(err),
            int(both_strands))
    total = L.synth_reads(*args, None, off.ctypes.data, None, th)
    bases = np.empty(max(total, 1), dtype=np.uint8)
    truth = np.zeros((n_reads, 4), dtype=np.int32)
    L.synth_reads(*args, bases.ctypes.data, off.ctypes.data, truth.ctypes.data, th)
    return bases[:total], off, truth
