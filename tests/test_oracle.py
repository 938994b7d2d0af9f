"""Pins the CPU oracle to the reference's own tests, golden vectors and to what
the mathematics fixes (direct definitions, brute force on small inputs).

Each test names the shmap-rs test or source lines it restates.  None of these
retypes the oracle's own formula: they compare it with an independent
computation (direct ntHash, brute-force windows and buckets, closed forms).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import fastx
import oracle_ffi as O

A, Cc, G, T = 0x3C8BFBB395C60474, 0x3193C18562A02B4C, 0x20323ED082572324, 0x295549F54BE24456
FW = {ord("A"): A, ord("C"): Cc, ord("G"): G, ord("T"): T}
for _k in list(FW):
    FW[_k + 32] = FW[_k]
RC = {ord("A"): T, ord("C"): G, ord("G"): Cc, ord("T"): A}
for _k in list(RC):
    RC[_k + 32] = RC[_k]
M64 = (1 << 64) - 1


def rotl(x, n):
    n %= 64
    return ((x << n) | (x >> (64 - n))) & M64 if n else x


def xorshift_dna(seed, n):
    """tests/nthash_equivalence.rs:14-25 / tests/multithreaded_parity.rs:14-25."""
    s, out = seed, bytearray()
    for _ in range(n):
        s ^= (s << 13) & M64
        s ^= s >> 7
        s ^= (s << 17) & M64
        out.append(b"ACGT"[s % 4])
    return bytes(out)


def direct_hashes(seq, k):
    """The per-window definition of tests/nthash_equivalence.rs:31-61."""
    res = []
    for i in range(len(seq) - k + 1):
        hf = hr = 0
        for j in range(k):
            hf ^= rotl(FW.get(seq[i + j], 0), k - 1 - j)
            hr ^= rotl(RC.get(seq[i + j], 0), j)
        res.append((hf, hr))
    return res


ORACLE_CLI = O.CLI


@pytest.fixture(scope="module", autouse=True)
def _built():
    if not os.path.exists(ORACLE_CLI) or not os.path.exists(O.LIB):
        O.build()


# ---- sketch -----------------------------------------------------------------

@pytest.mark.parametrize("k", [15, 21, 25, 31])
def test_rolling_sketch_equals_direct_definition(k):
    """tests/nthash_equivalence.rs:98-112: h = fw ^ rc, strand = fw > rc."""
    seq = xorshift_dna(12, 1200)
    sk = O.sketch(k, 1.0, seq)
    direct = direct_hashes(seq, k)
    assert len(sk) == len(direct)
    for i, (hf, hr) in enumerate(direct):
        assert int(sk["h"][i]) == hf ^ hr
        assert bool(sk["strand"][i]) == (hf > hr)
        assert int(sk["r"][i]) == i + k - 1


def test_canonical_form_is_xor_not_min():
    """tests/nthash_equivalence.rs:66-93: xor and min agree on < 1% of windows."""
    seq = xorshift_dna(11, 3000)
    direct = direct_hashes(seq, 25)
    agree = sum(1 for hf, hr in direct if min(hf, hr) == hf ^ hr)
    assert agree * 100 < len(direct)


def test_non_acgt_bytes_contribute_zero():
    seq = b"ACGTNNACGTRYacgt"
    k = 5
    sk = O.sketch(k, 1.0, seq)
    direct = direct_hashes(seq, k)
    assert [int(h) for h in sk["h"]] == [hf ^ hr for hf, hr in direct]


def test_chunked_sketching_concatenates_to_the_whole_sequence_sketch():
    """src/sketch.rs:236-270 (LCG-generated sequence, k=7, r=0.5)."""
    k, rng, seq = 7, 0x123456789ABCDEF0, bytearray()
    for _ in range(5000):
        rng = (rng * 6364136223846793005 + 1) & M64
        seq.append(b"ACGT"[(rng >> 60) % 4])
    seq = bytes(seq)
    whole = O.sketch(k, 0.5, seq)
    nw = len(seq) - k + 1
    for chunk in (1, 2, 13, 500, nw - 1, nw):
        parts, w0 = [], 0
        while w0 < nw:
            w1 = min(w0 + chunk, nw)
            parts.append(O.sketch(k, 0.5, seq[w0:w1 + k - 1], w0))
            w0 = w1
        joined = np.concatenate(parts)
        assert np.array_equal(joined, whole), chunk


def test_sketching_is_symmetric_under_reverse_complement():
    """src/sketch.rs:273-296."""
    k = 4
    s, s_rc = O.sketch(k, 1.0, b"ACGGT"), O.sketch(k, 1.0, b"ACCGT")[::-1]
    assert len(s) == len(s_rc)
    for i in range(len(s)):
        assert s[i]["r"] == i + k - 1
        assert s[i]["r"] == len(s) - s_rc[i]["r"] + k + 1
        assert s[i]["h"] == s_rc[i]["h"]


def test_sketching_a_sequence_shorter_than_k_is_empty():
    """src/sketch.rs:225-230."""
    assert len(O.sketch(4, 1.0, b"ACC")) == 0


@pytest.mark.parametrize("r,want", [(0.01, 0x28F5C28F5C28F60), (0.1, 0x1999999999999A00),
                                    (0.05, 0xCCCCCCCCCCCCD00), (1.0, M64)])
def test_threshold_constants(r, want):
    """src/sketch.rs:126-132: (h_frac * u64::MAX as f64) as u64."""
    assert O.lib().ok_threshold(r) == want
    if r < 1.0:  # the same value from exact rational arithmetic on the f64 product
        from fractions import Fraction
        prod = Fraction(r) * (1 << 64)
        assert want == int(float(prod))


# ---- index ------------------------------------------------------------------

def test_single_segment_indexing_matches_expected_hit_counts():
    """src/index.rs:716-741."""
    idx = O.OracleIndex(4, 1.0, [("ref", b"ACCAGTACCA")])
    t = O.sketch(4, 1.0, b"ACCAGTACCA")
    assert len(t) == 7
    assert [idx.count(int(h)) for h in t["h"]] == [2, 1, 1, 1, 1, 1, 2]
    st = idx.stats()
    assert (st["indexed_hits"], st["indexed_kmers"]) == (7, 6)


def test_two_segments_share_kmer_counts_across_both():
    """src/index.rs:743-769."""
    idx = O.OracleIndex(4, 1.0, [("segm1", b"ACCAGTACCA"), ("segm2", b"GGACCA")])
    t = O.sketch(4, 1.0, b"ACCAGTACCA")
    assert [idx.count(int(h)) for h in t["h"]] == [3, 1, 1, 1, 1, 1, 3]
    st = idx.stats()
    assert (st["indexed_hits"], st["indexed_kmers"]) == (10, 8)


def _brute_index(segments, k, r, max_matches):
    """insert_hit / finalize_shard / erase_frequent_kmers of src/index.rs:264-376,
    run literally: first hit to h2single, the next ones to h2multi while it holds
    fewer than M+1, single migrated in, lists sorted by (segm, r), lists longer
    than M erased."""
    single, multi = {}, {}
    for sid, (_, seq) in enumerate(segments):
        for tpos, km in enumerate(O.sketch(k, r, seq)):
            h = int(km["h"])
            hit = (int(km["r"]), tpos, int(km["strand"]), sid)
            if h not in single:
                single[h] = hit
            else:
                lst = multi.setdefault(h, [])
                if max_matches is None or len(lst) < max_matches + 1:
                    lst.append(hit)
    for h, lst in multi.items():
        if h in single:
            lst.append(single.pop(h))
        lst.sort(key=lambda x: (x[3], x[0]))
    if max_matches is not None:
        multi = {h: l for h, l in multi.items() if len(l) <= max_matches}
    return single, multi


@pytest.mark.parametrize("max_matches", [None, 1, 2, 3, 7])
def test_max_matches_closed_form_against_literal_insertion(max_matches):
    rep = b"ACGTACGTACGTACGTACGT"
    segs = [(f"segm{i}", rep + b"TTTTGGGGCCCCAAAA" + str(i).encode() + xorshift_dna(100 + i, 60))
            for i in range(10)]
    single, multi = _brute_index(segs, 6, 1.0, max_matches)
    idx = O.OracleIndex(6, 1.0, segs, max_matches)
    hashes = set(single) | set(multi)
    for _, s in segs:
        hashes |= {int(h) for h in O.sketch(6, 1.0, s)["h"]}
    for h in hashes:
        want = [single[h]] if h in single else multi.get(h, [])
        assert idx.hits(h) == want, h
    if max_matches == 3:
        assert idx.stats()["blacklisted_kmers"] > 0


def test_multithreaded_indexing_matches_single_threaded_indexing():
    """src/index.rs:782-855."""
    rep = "ACGTACGTACGTACGTACGT"
    segs = [(f"segm{i}", f"{rep}TTTTGGGGCCCCAAAA{i}".encode()) for i in range(10)]
    i1 = O.OracleIndex(6, 1.0, segs, 3, threads=1)
    i8 = O.OracleIndex(6, 1.0, segs, 3, threads=8)
    assert i1.stats() == i8.stats() and i1.stats()["blacklisted_kmers"] > 0
    for _, s in segs:
        for h in O.sketch(6, 1.0, s)["h"]:
            assert i1.hits(int(h)) == i8.hits(int(h))


# ---- seeding, pruning, scoring, mapping helpers -----------------------------

def test_unique_elements_groups_by_hash_and_collects_pmatches():
    """src/shmap/seeding.rs:222-245."""
    ks = [(60, 0x111111, 0), (70, 0x222222, 0), (10, 0x111111, 1), (20, 0x222222, 1),
          (30, 0x111111, 1), (40, 0x444444, 0), (50, 0x555555, 0)]
    arr = (O.OkKmer * len(ks))(*[O.OkKmer(r, h, s) for r, h, s in ks])
    hs = (C.c_uint64 * 7)()
    occs = (C.c_int32 * 7)()
    pm = (C.c_int32 * 7)()
    u = O.lib().ok_unique_pmatches(arr, len(ks), hs, occs, pm)
    groups, at = [], 0
    for i in range(u):
        groups.append(tuple(pm[at:at + occs[i]]))
        at += occs[i]
    assert sorted(groups) == sorted([(60, 30, 10), (70, 20), (40,), (50,)])
    assert list(hs[:u]) == sorted(hs[:u])


def test_hseed_worked_example():
    """docs/sections/09_pruning.tex worked example: m=100, thr=0.9."""
    h = O.lib().ok_hseed
    assert h(100, 40, 38) >= 0.9
    seeds, matches = 40, 38
    for miss in range(1, 10):
        seeds += 1
        sh = h(100, seeds, matches)
        assert (sh < 0.9) == (miss == 9), (miss, sh)
    assert h(100, 48, 38) == 1.0 - 10 / 100


def test_lcs_over_a_bucket_matches_expected_length():
    """src/shmap/scoring.rs:499-535: bucket (0,1) at halflen 2 spans t[2..6) =
    0x111111@30, 0x444444@40, 0x555555@50, 0x111111@60; p_ht maps 0x111111 to
    pmatches [5, 1] and 0x444444 to [3], so the LCS input is [5,1,3,5,1]."""
    seq = [5, 1, 3, 5, 1]
    arr = (C.c_int32 * 5)(*seq)
    rev = (C.c_int32 * 5)(*seq[::-1])
    L = O.lib().ok_lis_len
    assert max(L(arr, 5), L(rev, 5)) == 3
    # brute force: longest strictly increasing subsequence
    import itertools
    best = 0
    for n in range(1, 6):
        for comb in itertools.combinations(range(5), n):
            v = [seq[i] for i in comb]
            if all(a < b for a, b in zip(v, v[1:])):
                best = max(best, n)
    assert L(arr, 5) == best


def test_overlap_fixtures():
    """src/mapping.rs:469-486."""
    ov = O.lib().ok_overlap
    x = ov(0, 0, 100, 1, 0, 100)
    assert x == 0.0 and np.signbit(x)
    assert abs(ov(0, 0, 100, 0, 50, 150) - 50 / 150) < 1e-12


def test_calc_mapq_fixtures():
    """src/mapping.rs:488-495 and the rules of :341-354."""
    j2 = C.c_double()
    q = O.lib().ok_calc_mapq(0.9, -1.0, 3, 5, 0.5, 0.02, C.byref(j2))
    assert j2.value == 0.5  # J2 below theta2 is raised to it (and printed so)
    assert q == 60  # |3| < 5/2 == 2 is false
    assert O.lib().ok_calc_mapq(0.9, -1.0, 1, 5, 0.5, 0.02, C.byref(j2)) == 5
    assert O.lib().ok_calc_mapq(0.9, 0.89, 5, 5, 0.5, 0.02, C.byref(j2)) == 0


# ---- end to end -------------------------------------------------------------

def _run_cli(*args):
    return subprocess.run([ORACLE_CLI, *args], capture_output=True, text=True)


def test_paf_output_matches_the_checked_in_golden_file(golden_dir):
    """tests/golden_paf.rs:26-53."""
    r = _run_cli("-s", os.path.join(golden_dir, "tiny_ref.fa"), "-p", os.path.join(golden_dir, "tiny_reads.fa"),
                 "-k", "8", "-r", "1.0", "-t", "0.1")
    assert r.returncode == 0
    want = fastx.strip_time(open(os.path.join(golden_dir, "tiny.golden.paf")).read())
    assert fastx.strip_time(r.stdout) == want


def test_maps_reads_against_the_tiny_fixture(golden_dir):
    """tests/integration_paf.rs:8-42."""
    r = _run_cli("-s", os.path.join(golden_dir, "tiny_ref.fa"), "-p", os.path.join(golden_dir, "tiny_reads.fa"),
                 "-k", "8", "-r", "1.0", "-t", "0.1")
    lines = r.stdout.strip().split("\n")
    assert len(lines) == 2
    for line in lines:
        f = line.split("\t")
        assert len(f) >= 12 and f[5] == "ref" and f[4] == "+"
    assert lines[0].startswith("read1\t") and lines[1].startswith("read2\t")


def test_rejects_an_invalid_parameter_with_a_clear_error(golden_dir):
    """tests/integration_paf.rs:44-57."""
    r = _run_cli("-s", os.path.join(golden_dir, "tiny_ref.fa"), "-p", os.path.join(golden_dir, "tiny_reads.fa"),
                 "-k", "0")
    assert r.returncode != 0 and "K-mer length" in r.stderr


def _floor_fixture(tmp_path, read_lens, K=8):
    x, seq = 0x2545F4914F6CDD1D, []
    for _ in range(4000):
        x ^= (x << 13) & M64
        x ^= x >> 7
        x ^= (x << 17) & M64
        seq.append("ACGT"[x % 4])
    ref = "".join(seq)
    (tmp_path / "ref.fa").write_text(f">ref\n{ref}\n")
    reads = "".join(f">len{l}_{i}\n{ref[100 + i * 7:100 + i * 7 + l]}\n" for i, l in enumerate(read_lens))
    (tmp_path / "reads.fa").write_text(reads)
    r = _run_cli("-s", str(tmp_path / "ref.fa"), "-p", str(tmp_path / "reads.fa"), "-k", str(K), "-r", "1.0",
                 "-t", "0.4", "-m", "Containment")
    assert r.returncode == 0
    lines = [l for l in r.stdout.split("\n") if l.strip()]
    for l in lines:
        c = l.split("\t")
        assert 0 <= int(c[2]) < int(c[3]) <= int(c[1]), l
    return lines


def test_short_read_floor(tmp_path):
    """tests/short_read_sketch_floor.rs:96-132."""
    K = 8
    assert _floor_fixture(tmp_path, [K, K + 1, K + 2, K + 3]) == []
    assert len(_floor_fixture(tmp_path, [K + 4, K + 12, K + 40])) == 3
    assert len(_floor_fixture(tmp_path, [K + 20, K, K + 20, K + 1, K + 20])) == 3


def test_threaded_output_matches_single_threaded_output(tmp_path):
    """tests/multithreaded_parity.rs:60-101 (+ the per-read-stats file)."""
    reference = xorshift_dna(42, 20_000)
    (tmp_path / "ref.fa").write_bytes(b">chr1\n" + reference + b"\n")
    out = []
    for i in range(60):
        start, ln = (i * 137) % 19_000, 200 + (i % 5) * 50
        out.append(b">mapped_%d\n%s\n" % (i, reference[start:start + ln]))
        if i % 4 == 0:
            out.append(b">unmapped_%d\n%s\n" % (i, xorshift_dna(9000 + i, 150)))
    (tmp_path / "reads.fa").write_bytes(b"".join(out))
    res = {}
    for th in ("1", "3", "8"):
        st = tmp_path / f"st{th}.tsv"
        r = _run_cli("-s", str(tmp_path / "ref.fa"), "-p", str(tmp_path / "reads.fa"), "-k", "12", "-r", "0.5",
                     "-t", "0.2", "-@", th, "--per-read-stats", str(st))
        rows = ["\t".join(l.split("\t")[:-1]) for l in st.read_text().split("\n") if l]
        res[th] = (fastx.strip_time(r.stdout), rows)
    assert res["1"][0]
    assert res["1"] == res["3"] == res["8"]


def _brute_best_window(t_h, t_r, t_s, seeds, b, hl, m, metric):
    """All windows of best_fixed_length (src/shmap/scoring.rs:189-291) scored
    independently: I(l) = sum over seeds of min(count in window, occs)."""
    end = min(len(t_h), (b + 2) * hl)
    best = None
    for l in range(max(0, b * hl), end):
        r = min(l + m, end)
        cnt, same = {}, 0
        for q in range(l, r):
            s = seeds.get(int(t_h[q]))
            if s is not None:
                cnt[int(t_h[q])] = cnt.get(int(t_h[q]), 0) + 1
                same += 1 if s[1] == int(t_s[q]) else -1
        inter = sum(min(c, seeds[h][0]) for h, c in cnt.items())
        score = inter / m if metric == "Containment" else inter / (m + (r - l) - inter)
        if best is None or score > best[0]:
            best = (score, inter, same, int(t_r[l]), int(t_r[r - 1]))
    return best


@pytest.mark.parametrize("metric", ["Containment", "Jaccard"])
def test_best_fixed_length_against_brute_force_windows(metric):
    ref = xorshift_dna(7, 40_000)
    segs = [("chr1", ref[:25_000]), ("chr2", ref[25_000:])]
    k, hf = 15, 0.1
    rng = np.random.default_rng(5)
    reads = []
    for i in range(25):
        s = int(rng.integers(0, 38_000))
        r = bytearray(ref[s:s + 1500])
        for j in rng.integers(0, len(r), 8):
            r[j] = b"ACGT"[(b"ACGT".index(r[j]) + 1) % 4]
        reads.append(bytes(r))
    idx = O.OracleIndex(k, hf, segs)
    p = O.params(k=k, h_frac=hf, theta=0.3, metric=metric)
    res = idx.map_reads(p, reads)
    assert res["mapped"].sum() >= 20
    tsk = [O.sketch(k, hf, s) for _, s in segs]
    for read, rec in zip(reads, res):
        if not rec["mapped"]:
            continue
        rs = O.sketch(k, hf, read)
        m = len(rs)
        seeds = {}
        order = np.lexsort((-rs["r"].astype(np.int64), rs["h"]))
        for j in order:  # the seed keeps the strand of its smallest-r occurrence
            h = int(rs["h"][j])
            seeds[h] = (seeds.get(h, (0, 0))[0] + 1, int(rs["strand"][j]))
        t = tsk[rec["b_segm"]]
        got = _brute_best_window(t["h"], t["r"], t["strand"], seeds, int(rec["b_b"]), m, m, metric)
        assert got[0] == rec["J"]
        assert got[1:] == (rec["intersection"], rec["same_strand"], rec["t_l"], rec["t_r"])
        assert rec["s_sz"] == rec["t_r"] - rec["t_l"] + 1


def test_seeded_buckets_and_seed_counters_against_brute_force():
    """match_seeds + get_sorted_buckets (src/shmap/seeding.rs:98-212,
    src/buckets.rs:512-518): the buckets a read touches are {b, b-1} of every
    hit of its first S occurrences' seeds, rarest first."""
    unit = xorshift_dna(3, 171)
    ref = xorshift_dna(8, 30_000) + unit * 40 + xorshift_dna(9, 30_000)
    k, hf = 13, 0.2
    idx = O.OracleIndex(k, hf, [("chr1", ref)])
    rng = np.random.default_rng(2)
    reads = [ref[s:s + 2000] for s in rng.integers(0, len(ref) - 2000, 30)]
    p = O.params(k=k, h_frac=hf, theta=0.4, min_diff=0.075)
    res = idx.map_reads(p, reads)
    for read, rec in zip(reads, res):
        rs = O.sketch(k, hf, read)
        m = len(rs)
        occ = {}
        for h in rs["h"]:
            occ[int(h)] = occ.get(int(h), 0) + 1
        seeds = sorted(occ, key=lambda h: (idx.count(h), h))
        S = int((1.0 - (0.4 - 0.075)) * m) + 1
        consumed, seen = [], 0
        for h in seeds:
            if seen >= S:
                break
            consumed.append(h)
            seen += occ[h]
        buckets = set()
        for h in consumed:
            for (r, tpos, strand, sid) in idx.hits(h):
                b = tpos // m
                buckets.add((sid, b))
                if b > 0:
                    buckets.add((sid, b - 1))
        assert rec["S"] == S
        assert rec["seeded_buckets"] == len(buckets)
        assert rec["seed_matches"] == sum(idx.count(h) for h in consumed)
        assert rec["max_seed_matches"] == max([idx.count(h) for h in consumed] + [0])


# ---- match_rest's decisions against brute force (tests/brute_map.py) ---------

def _dup_reference():
    """Two copies of a 12 kb block (0.2% substitutions apart), a tandem array
    of a 171 bp unit, and unique sequence: reads from the block have a
    competing second placement, reads from the array many equal buckets."""
    rng = np.random.default_rng(123)
    block = bytearray(xorshift_dna(31, 12_000))
    copy = bytearray(block)
    for j in rng.integers(0, len(copy), 24):
        copy[j] = b"ACGT"[(b"ACGT".index(copy[j]) + 1 + int(rng.integers(0, 3))) % 4]
    unit = xorshift_dna(32, 171)
    array = bytearray(unit * 60)
    for j in rng.integers(0, len(array), 150):
        array[j] = b"ACGT"[(b"ACGT".index(array[j]) + 1) % 4]
    chr1 = xorshift_dna(33, 9_000) + bytes(block) + xorshift_dna(34, 7_000) + bytes(array)
    chr2 = xorshift_dna(35, 5_000) + bytes(copy) + xorshift_dna(36, 6_000)
    return [("chr1", chr1), ("chr2", chr2)]


def _dup_reads(segs, n=36, seed=9):
    rng = np.random.default_rng(seed)
    reads = []
    for i in range(n):
        sid = int(rng.integers(0, 2))
        s = segs[sid][1]
        ln = int(rng.integers(800, 3200))
        at = int(rng.integers(0, len(s) - ln))
        r = bytearray(s[at:at + ln])
        for j in rng.integers(0, len(r), int(ln * 0.004)):
            r[j] = b"ACGT"[(b"ACGT".index(r[j]) + 1) % 4]
        if i % 3 == 0:  # reverse complement
            r = bytearray(r[::-1].translate(bytes.maketrans(b"ACGT", b"TGCA")))
        reads.append(bytes(r))
    return reads


_BRUTE_FIELDS = ["b_segm", "b_b", "J", "intersection", "same_strand", "t_l", "t_r", "sh", "b2_segm", "b2_b",
                 "intersection2", "J2", "sigmas", "mapq"]


@pytest.mark.parametrize("metric", ["Containment", "Jaccard", "bucket_SH", "bucket_LCS"])
@pytest.mark.parametrize("variant", ["default", "P", "B"])
def test_match_rest_decisions_against_brute_force(metric, variant):
    """The best and second-best buckets, their scores, sigma, MAPQ, the
    printed J2 and sh, final_buckets, refined_buckets and refine_memo_hits of
    every read, against brute_map's closed forms (src/shmap/scoring.rs:367-489,
    src/shmap/mod.rs:809-861, src/mapping.rs:331-361)."""
    import brute_map as B
    segs = _dup_reference()
    reads = _dup_reads(segs)
    k, hf = 15, 0.1
    prm = dict(theta=0.3, min_diff=0.075, max_overlap=0.3)
    bi = B.BruteIndex(k, hf, segs)
    oi = O.OracleIndex(k, hf, segs)
    p = O.params(k=k, h_frac=hf, metric=metric, no_bucket_pruning=variant == "P", one_sweep=variant == "B", **prm)
    res = oi.map_reads(p, reads)
    n_second = n_mapq0 = 0
    for i, (read, rec) in enumerate(zip(reads, res)):
        want = B.brute_map(bi, k, hf, read, metric=metric, nbp=variant == "P", one_sweep=variant == "B", **prm)
        assert rec["mapped"] == want["mapped"], i
        for f in ("final_buckets", "refined_buckets", "refine_memo_hits", "seeded_buckets", "S"):
            assert rec[f] == want[f], (i, f, rec[f], want[f])
        if not want["mapped"]:
            continue
        for f in _BRUTE_FIELDS:
            assert rec[f] == want[f], (i, f, rec[f], want[f])
        n_second += want["b2_segm"] >= 0
        n_mapq0 += want["mapq"] == 0
    assert res["mapped"].sum() >= 20
    if metric != "bucket_SH" or variant != "P":  # under -P every bucket_SH score is 1.0
        assert n_second >= 5, n_second  # the second pass is exercised
    assert n_mapq0 >= 3, n_mapq0


@pytest.mark.parametrize("metric", ["Containment", "Jaccard", "bucket_LCS"])
def test_pruning_is_sound(metric):
    """-P scores every seeded bucket; the default prunes by the seed
    heuristic.  Pruning only drops buckets that cannot win, so both give the
    same best and second-best mappings (src/shmap/pruning.rs:89-119), and the
    heuristic bounds every score: sh >= J."""
    segs = _dup_reference()
    reads = _dup_reads(segs, n=60, seed=17)
    k, hf = 15, 0.1
    oi = O.OracleIndex(k, hf, segs)
    prm = dict(k=k, h_frac=hf, theta=0.3, min_diff=0.075, max_overlap=0.3, metric=metric)
    a = oi.map_reads(O.params(**prm), reads)
    b = oi.map_reads(O.params(no_bucket_pruning=True, **prm), reads)
    assert np.array_equal(a["mapped"], b["mapped"]) and a["mapped"].sum() >= 30
    mk = a["mapped"] == 1
    fields = ["b_segm", "b_b", "J", "b2_segm", "b2_b", "J2", "mapq"]
    if metric != "bucket_LCS":  # LCS reports its bucket's content, which -P leaves unextended
        fields += ["intersection", "intersection2", "t_l", "t_r", "same_strand", "sigmas"]
    for f in fields:
        assert np.array_equal(a[f][mk], b[f][mk]), f
    # sh = 1 - (m - matches)/m and J = I/m with I <= matches: sh >= J exactly,
    # up to the last bit of the two roundings when I == matches
    assert np.all(a["sh"][mk] >= a["J"][mk] - 1e-12)
    assert np.any(a["sh"][mk] > a["J"][mk] + 1e-3)
    assert np.all(b["sh"][mk] == 1.0)
    assert np.all(a["refined_buckets"] <= b["refined_buckets"])


def test_prune_hit_count_is_a_per_read_measurement():
    """The hit records the prune pass reads (the P term of the path's
    algorithmic bytes) are counted per read: none under -P, which skips the
    pass; the same on 1 and 4 threads; additive over disjoint read sets."""
    segs = _dup_reference()
    reads = _dup_reads(segs, n=40, seed=23)
    k, hf = 15, 0.1
    oi = O.OracleIndex(k, hf, segs)
    prm = dict(k=k, h_frac=hf, theta=0.3, min_diff=0.075, max_overlap=0.3)
    oi.map_reads(O.params(no_bucket_pruning=True, **prm), reads)
    assert oi.last_prune_hits == 0
    p = O.params(**prm)
    oi.map_reads(p, reads, threads=1)
    total = oi.last_prune_hits
    assert total > 0
    oi.map_reads(p, reads, threads=4)
    assert oi.last_prune_hits == total
    oi.map_reads(p, reads[:17])
    a = oi.last_prune_hits
    oi.map_reads(p, reads[17:])
    assert a + oi.last_prune_hits == total


def test_sigma_and_mapq_at_hand_computed_points():
    """src/mapping.rs:331-361: sigma = |I2 - I| / sqrt(I2 + I) with -1 read
    as 0; J2 below theta2 is raised to theta2 and printed so; MAPQ is 60 (5
    when |same_strand| < I/2, integer division) only when 1 - J2/J > d
    strictly."""
    q = O.lib().ok_calc_mapq
    j2 = C.c_double()
    # frac == d exactly: 1 - 0.5/1.0 == 0.5 in binary, not > 0.5
    assert q(1.0, 0.5, 10, 10, 0.1, 0.5, C.byref(j2)) == 0 and j2.value == 0.5
    assert q(1.0, 0.5, 10, 10, 0.1, 0.4999999, C.byref(j2)) == 60
    # J2 below theta2: printed J2 is theta2, and frac uses it
    assert q(0.8, 0.1, 10, 10, 0.4, 0.075, C.byref(j2)) == 60 and j2.value == 0.4
    assert q(0.42, 0.1, 10, 10, 0.4, 0.075, C.byref(j2)) == 0 and j2.value == 0.4  # 1 - 0.4/0.42 < 0.075
    # strand agreement: |same| < I/2 with I/2 truncated
    assert q(0.9, -1.0, 2, 5, 0.3, 0.075, C.byref(j2)) == 60   # 2 < 5/2 == 2 is false
    assert q(0.9, -1.0, -1, 5, 0.3, 0.075, C.byref(j2)) == 5
    assert q(0.9, -1.0, -3, 7, 0.3, 0.075, C.byref(j2)) == 60
    # sigma through whole reads: equal intersections give 0, a missing second best -1
    import brute_map as B
    segs = _dup_reference()
    reads = _dup_reads(segs)
    oi = O.OracleIndex(15, 0.1, segs)
    res = oi.map_reads(O.params(k=15, h_frac=0.1, theta=0.3, min_diff=0.075, max_overlap=0.3), reads)
    for rec in res[res["mapped"] == 1]:
        I, I2 = int(rec["intersection"]), int(rec["intersection2"])
        if rec["b2_segm"] < 0:
            assert rec["sigmas"] == -1.0 and I2 == -1
        else:
            assert rec["sigmas"] == abs(I2 - I) / np.sqrt(I2 + I)
