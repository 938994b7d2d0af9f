"""The C-ABI library loads and exports every entry point include/shmap_hip.h
declares; the CLI keeps the reference's flag validation.  No GPU needed."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import fastx

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "shmap_hip.h")
LIB = os.path.join(ROOT, "shmap-rs_amd", "lib", "libshmap_hip.so")
CLI = os.path.join(ROOT, "shmap-rs_amd", "lib", "shmap")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(shm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_mapping_api():
    names = declared_functions()
    for required in ("shm_sketcher_create", "shm_sketch", "shm_index_build", "shm_index_count", "shm_index_hits",
                     "shm_mapper_create", "shm_map_batch", "shm_map_batch_device", "shm_format_paf",
                     "shm_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_reports_argument_errors_without_a_device():
    lib = C.CDLL(LIB)
    lib.shm_last_error.restype = C.c_char_p
    out = C.c_void_p()
    assert lib.shm_sketcher_create(0, C.c_double(0.1), C.byref(out)) != 0
    assert b"K-mer length" in lib.shm_last_error()
    assert lib.shm_sketcher_create(15, C.c_double(1.5), C.byref(out)) != 0
    assert b"hash ratio" in lib.shm_last_error()


def _index_header(version=1, n_kmers=0, nseg=0, name_bytes=0, bits=6):
    """An index file header (shm_index_save's layout: 192 bytes)."""
    import struct
    base = [i << bits for i in range(8)]
    return (b"SHMIDX01" + struct.pack("<IIiid", version, nseg, 15, 0, 0.1) +
            struct.pack("<6q", n_kmers, 0, n_kmers, 0, 0, 0) + struct.pack("<8Q", *base) +
            struct.pack("<8I", *([bits] * 8)) + struct.pack("<QQ", 8 << bits, name_bytes))


def test_index_load_checks_the_file_before_the_device(tmp_path):
    """shm_index_load refuses what is not a complete index file of this
    version with a message, before it touches the device."""
    lib = C.CDLL(LIB)
    lib.shm_last_error.restype = C.c_char_p
    lib.shm_index_load.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]
    out = C.c_void_p()

    def load(data):
        f = tmp_path / "x.idx"
        f.write_bytes(data)
        rc = lib.shm_index_load(str(f).encode(), 0, C.byref(out))
        return rc, lib.shm_last_error().decode()

    assert len(_index_header()) == 192
    rc = lib.shm_index_load(str(tmp_path / "none.idx").encode(), 0, C.byref(out))
    assert rc == 5 and "cannot open" in lib.shm_last_error().decode()
    rc, msg = load(b">r1\nACGT\n")
    assert rc == 1 and "not a shmap index file" in msg
    rc, msg = load(_index_header(version=2))
    assert rc == 1 and "version 2" in msg
    rc, msg = load(_index_header(bits=3))
    assert rc == 1 and "invalid header" in msg
    # an empty index needs 4 + 8 + 8 bytes of tables, strand prefix and slots, and the end marker
    full = _index_header() + (0).to_bytes(8, "little") * 2 + (0).to_bytes(4, "little") + bytes(16 * (8 << 6))
    rc, msg = load(full)
    assert rc == 1 and "truncated" in msg
    rc, msg = load(full + b"SHMIDXXX")
    assert rc == 1 and "end marker" in msg
    rc, msg = load(full + b"SHMIDXEN" + b"extra")
    assert rc == 1 and "longer than" in msg


def _index_file(k=15, h_frac=0.1, bits=6, t_r=(24, 30, 50), t_s=(0, 1, 0), hit_g=(0, 2, 1),
                hit_seg=(0, 0, 0), strand_pre=(0, 0, 0, 1), kstart=(0, 3), slots=None, fill_shard=None):
    """A well-framed index file of one 100-base segment with three k-mers
    (hashes 5, 6, 5) in shm_index_save's layout; the keyword arguments
    replace parts of it."""
    import struct
    n = len(t_r)
    nslots = 8 << bits
    tab = [(0, 0, 0, 0)] * nslots
    for at, slot in (slots or {(5 << bits): (5, 0, 0, 2), (6 << bits): (6, 0, 2, 1)}).items():
        tab[at] = slot
    if fill_shard is not None:  # no empty slot left in that shard
        for i in range(fill_shard << bits, (fill_shard + 1) << bits):
            tab[i] = (100 + i, 0, 0, 1)
    hit_r = [((t_r[g] | (t_s[g] << 31)) & 0xffffffff) if g < n else 0 for g in hit_g]
    base = [i << bits for i in range(8)]
    hdr = (b"SHMIDX01" + struct.pack("<IIiid", 1, 1, k, 0, h_frac) + struct.pack("<6q", n, 2, n, 2, 0, 0) +
           struct.pack("<8Q", *base) + struct.pack("<8I", *([bits] * 8)) + struct.pack("<QQ", nslots, 4 + 1))
    body = (struct.pack("<I", 1) + b"s" + struct.pack("<i", 100) + struct.pack("<2q", *kstart) +
            struct.pack("<2q", 0, 100) + struct.pack(f"<{n}Q", 5, 6, 5) + struct.pack(f"<{n}i", *t_r) +
            bytes(t_s) + struct.pack(f"<{n}I", *hit_g) + struct.pack(f"<{n}I", *hit_r) +
            struct.pack(f"<{n}I", *hit_seg) + struct.pack(f"<{n + 1}I", *strand_pre) +
            b"".join(struct.pack("<4I", *t) for t in tab))
    return hdr + body + b"SHMIDXEN"


def test_index_load_checks_the_contents_before_the_device(tmp_path):
    """A well-framed file whose arrays would send a device lookup out of
    bounds or into an endless probe is refused on the host, before any device
    work: slots past the hits, a shard without an empty slot, hits outside
    their segment, a wrong strand prefix, unordered sketch positions, header
    parameters the sketcher refuses, shards beyond 2^31 slots."""
    lib = C.CDLL(LIB)
    lib.shm_last_error.restype = C.c_char_p
    lib.shm_index_load.argtypes = [C.c_char_p, C.c_int32, C.POINTER(C.c_void_p)]
    lib.shm_index_free.argtypes = [C.c_void_p]

    def load(data):
        f = tmp_path / "x.idx"
        f.write_bytes(data)
        out = C.c_void_p()
        rc = lib.shm_index_load(str(f).encode(), 0, C.byref(out))
        if rc == 0:
            lib.shm_index_free(out)
        return rc, lib.shm_last_error().decode()

    # the well-formed file passes every host check: it loads where a device
    # exists, and fails only at the device (SHM_ERR_HIP) where none does
    rc, msg = load(_index_file())
    assert rc in (0, 2), msg
    bad = {
        "slot past the hits": _index_file(slots={5 << 6: (5, 0, 2, 2)}),
        "runs longer than the hits": _index_file(slots={5 << 6: (5, 0, 0, 3), 6 << 6: (6, 0, 0, 3)}),
        "no empty slot": _index_file(fill_shard=3),
        "segment out of range": _index_file(hit_seg=(0, 1, 0)),
        "hit past the k-mers": _index_file(hit_g=(0, 3, 1), kstart=(0, 3)),
        "strand prefix": _index_file(strand_pre=(0, 1, 1, 1)),
        "positions out of order": _index_file(t_r=(24, 24, 50)),
        "position past the segment": _index_file(t_r=(24, 30, 100)),
        "strand not 0/1": _index_file(t_s=(0, 2, 0)),
    }
    for what, data in bad.items():
        rc, msg = load(data)
        assert rc == 1 and "inconsistent" in msg, (what, rc, msg)
    for kw in (dict(k=0), dict(k=5000), dict(h_frac=0.0), dict(h_frac=1.5)):
        rc, msg = load(_index_file(**kw))
        assert rc == 1 and "invalid index parameters" in msg, (kw, msg)
    rc, msg = load(_index_header(bits=32))
    assert rc == 1 and "invalid header" in msg


@pytest.mark.parametrize("args,needle", [
    (["-k", "0"], "K-mer length (-k) should be positive. You provided 0."),
    (["-r", "0"], "Given hash ratio (-r) should be between 0 and 1. You provided 0."),
    (["-t", "1.5"], "The threshold (-t) should be between 0 and 1. You provided 1.5."),
    (["-d", "-1"], "The minimum difference (-d) should be >= 0. You provided -1."),
    (["-o", "2"], "The maximum overlap (-o) should be between 0 and 1. You provided 2."),
    (["-M", "0"], "The number of seed matches (-M) should be positive. You provided 0."),
    (["-S", "-3"], "The number of seeds (-S) should be positive. You provided -3."),
    (["-@", "0"], "The number of threads (-@) should be positive. You provided 0."),
    (["-v", "3"], "--verbose (-v) should be 0, 1, or 2. You provided 3."),
])
def test_cli_validation_messages(args, needle):
    """src/params.rs:191-263; src/main.rs:18-21 prints `ERROR: ...` and exits 1."""
    r = subprocess.run([CLI, "-s", "ref.fa", "-p", "reads.fa", *args], capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stderr.strip() == "ERROR: " + needle


def test_cli_rejects_an_unknown_metric():
    r = subprocess.run([CLI, "-s", "ref.fa", "-p", "reads.fa", "-m", "Hamming"], capture_output=True, text=True)
    assert r.returncode == 2 and "possible values: Containment, Jaccard, bucket_SH, bucket_LCS" in r.stderr


def _dump(path, threads=1):
    r = subprocess.run([CLI, "-p", str(path), "--dump-reads", "-@", str(threads)], capture_output=True)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.split(b"\n")
    return [(lines[i][1:].decode(), lines[i + 1]) for i in range(0, len(lines) - 1, 2)]


def test_cli_reader_matches_fasta_semantics(tmp_path):
    """The CLI's block reader against tests/fastx.py (needletail's rules:
    id cut at the first space or tab, line breaks and '\\r' removed): CRLF,
    wrapped and blank lines, a record with no bases, a line longer than the
    read block, no final newline."""
    import random
    rng = random.Random(5)
    long_seq = bytes(rng.choice(b"ACGTNacgt") for _ in range(20_000_000))
    body = (b">r1 first read\r\nACGTACGT\r\nacgtN\r\n\r\n>r2\tsecond\nAC\nGT\n\n\nTT\n>empty\n>r4 x y\n"
            + long_seq + b"\n>last\nGATTACA")
    f = tmp_path / "reads.fa"
    f.write_bytes(body)
    want = fastx.read_fasta(f)
    got = _dump(f)
    assert [n for n, _ in got] == [n for n, _ in want]
    assert [s for _, s in got] == [s for _, s in want]


def test_cli_parallel_reader_matches_fasta_semantics(tmp_path):
    """With -@ N plain FASTA is read in parallel byte ranges (windows of 1 MiB
    and up, cut at record starts, one range per thread): the same records as
    the sequential reader and tests/fastx.py, for records that straddle range
    and window edges, CRLF, blank lines, '>' and '@' inside sequence lines of
    the wrong column, a header-only record, text before the first record and
    no final newline."""
    import random
    rng = random.Random(11)
    parts = [b"junk before the first record\n\n"]
    for i in range(700):
        n = rng.choice([0, 1, 59, 60, 61, 700, 5_000, 40_000])
        seq = bytes(rng.choice(b"ACGTNacgtRY") for _ in range(n))
        width = rng.choice([60, 80, 1000])
        eol = b"\r\n" if i % 7 == 0 else b"\n"
        lines = [seq[j:j + width] for j in range(0, len(seq), width)]
        if i % 11 == 0:
            lines.insert(len(lines) // 2, b"")
        if i % 13 == 0 and lines:
            lines[-1] = lines[-1] + b">notheader@"
        parts.append(b">read%d %s\t%s" % (i, b"desc" * (i % 3), eol) + eol.join(lines) + eol)
    parts.append(b">last\nGATTACA")
    body = b"".join(parts)
    assert len(body) > 3_000_000
    f = tmp_path / "reads.fa"
    f.write_bytes(body)
    want = fastx.read_fasta(f)
    one = _dump(f, 1)
    for threads in (2, 8, 16):
        got = _dump(f, threads)
        assert [n for n, _ in got] == [n for n, _ in want] == [n for n, _ in one]
        assert [s for _, s in got] == [s for _, s in want] == [s for _, s in one]


def test_cli_reader_fastq(tmp_path):
    f = tmp_path / "reads.fq"
    f.write_bytes(b"@q1 desc\nACGTN\n+\nIIIII\n@q2\nGGCC\r\n+q2\r\n!!!!\r\n")
    assert _dump(f) == [("q1", b"ACGTN"), ("q2", b"GGCC")]


def test_cli_reader_gzip(tmp_path):
    """gzip input (needletail's compressed FASTA/FASTQ, src/io/mod.rs:1, 357)
    reads as the plain file does: one member, several members (concatenated
    gzip streams), FASTQ, and a file whose blocks straddle the member ends."""
    import gzip
    import random
    rng = random.Random(11)
    recs = [(f"r{i} d", bytes(rng.choice(b"ACGTN") for _ in range(rng.randint(0, 5000)))) for i in range(300)]
    body = b"".join(b">" + n.encode() + b"\n" + sq[:3000] + b"\r\n" + sq[3000:] + b"\n" for n, sq in recs)
    plain = tmp_path / "reads.fa"
    plain.write_bytes(body)
    want = _dump(plain)
    assert [n for n, _ in want] == [n.split()[0] for n, _ in recs]
    one = tmp_path / "one.fa.gz"
    one.write_bytes(gzip.compress(body))
    assert _dump(one) == want
    cut = [0, 1, len(body) // 3, len(body) // 3 + 7, len(body)]
    multi = tmp_path / "multi.fa.gz"
    multi.write_bytes(b"".join(gzip.compress(body[a:b]) for a, b in zip(cut, cut[1:])))
    assert _dump(multi) == want
    fq = tmp_path / "reads.fq.gz"
    fq.write_bytes(gzip.compress(b"@q1 desc\nACGTN\n+\nIIIII\n@q2\nGGCC\r\n+q2\r\n!!!!\r\n"))
    assert _dump(fq) == [("q1", b"ACGTN"), ("q2", b"GGCC")]


def test_cli_reader_refuses_other_compression_and_corrupt_gzip(tmp_path):
    import bz2
    import gzip
    import lzma
    body = b">r1\nACGT\n"
    for name, data, fmt in (("r.fa.bz2", bz2.compress(body), "bzip2"), ("r.fa.xz", lzma.compress(body), "xz"),
                            ("r.fa.zst", bytes([0x28, 0xb5, 0x2f, 0xfd]) + b"\0" * 8, "zstd")):
        f = tmp_path / name
        f.write_bytes(data)
        r = subprocess.run([CLI, "-p", str(f), "--dump-reads"], capture_output=True)
        assert r.returncode == 1 and fmt.encode() + b"-compressed" in r.stderr, r.stderr
    bad = bytearray(gzip.compress(body * 1000))
    bad[len(bad) // 2] ^= 0xff
    f = tmp_path / "bad.fa.gz"
    f.write_bytes(bytes(bad))
    r = subprocess.run([CLI, "-p", str(f), "--dump-reads"], capture_output=True)
    assert r.returncode == 1 and b"decompress" in r.stderr, r.stderr


def test_pack_reads_round_trip():
    """shm_pack_reads (host only): 2-bit codes of A C G T in either case, base i
    at bits 2*(i%4) of byte i/4, reads back to back, and every other byte in
    sorted [start, end) runs that end at read boundaries -- the same whatever
    the thread count (bytes shared by two threads' reads)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "shmap-rs_amd"))
    import shmap_amd as G
    rng = np.random.default_rng(1)
    reads = []
    for i in range(300):
        n = int(rng.integers(0, 2000)) if i % 50 else 0
        alpha = b"ACGTacgtNRYn-" if i % 3 == 0 else b"ACGTacgt"
        reads.append(np.frombuffer(alpha, dtype=np.uint8)[rng.integers(0, len(alpha), n)].tobytes())
    reads[7] = b"N" * 33  # a whole read of other bytes
    allb = np.frombuffer(b"".join(reads), dtype=np.uint8)
    low = allb | np.uint8(0x20)
    acgt = np.isin(low, np.frombuffer(b"acgt", dtype=np.uint8))
    code = np.zeros(len(allb), dtype=np.uint8)
    for j, ch in enumerate(b"acgt"):
        code[low == ch] = j
    want_bytes = np.zeros((len(allb) + 3) // 4, dtype=np.uint8)
    for j in range(4):
        want_bytes |= np.pad(code[j::4], (0, len(want_bytes) - len(code[j::4]))) << np.uint8(2 * j)
    ends = np.cumsum([len(r) for r in reads])
    for th in (1, 4, 16):
        pr = G.PackedReads(reads, threads=th, pinned=False)
        assert pr.n == len(reads) and pr.bases == len(allb)
        assert np.array_equal(pr.off[1:], ends)
        assert np.array_equal(pr.packed[:len(want_bytes)], want_bytes)
        runs = pr.exc[:2 * pr.n_exc].reshape(-1, 2)
        mask = np.zeros(len(allb), dtype=bool)
        for s, e in runs:
            assert s < e and not mask[s:e].any()
            mask[s:e] = True
        assert np.array_equal(mask, ~acgt)
        assert np.all(np.diff(runs[:, 0]) > 0)
        assert np.all(np.isin(runs[:, 1], ends) | acgt[np.minimum(runs[:, 1], len(allb) - 1)])
