"""Shared helpers for parity tests: run the CPU oracle and the HIP library on
the same inputs and diff every field of every read."""
import numpy as np

import oracle_ffi as O
import shmap_amd as G

COUNTER_FIELDS = ["mapped", "unmappable_short", "read_len", "m", "u", "S", "possible_matches",
                  "seed_matches", "total_matches", "max_seed_matches", "seeded_buckets",
                  "refined_buckets", "refine_memo_hits", "final_buckets"]
MAPPED_FIELDS = ["p_start", "p_end", "strand", "segm_id", "t_l", "t_r", "mapq", "s_sz",
                 "intersection", "intersection2", "same_strand", "b_segm", "b_b", "b2_segm", "b2_b",
                 "J", "J2", "sh", "sigmas", "match_inefficiency"]


def _eq(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Element-wise equality; floats compared bit for bit, any NaN equal to any NaN."""
    if a.dtype.kind == "f" or b.dtype.kind == "f":
        a, b = a.astype(np.float64), b.astype(np.float64)
        both_nan = np.isnan(a) & np.isnan(b)
        return both_nan | (a.view(np.int64) == b.view(np.int64))
    return a.astype(np.int64) == b.astype(np.int64)


def diff_results(ores, gres, limit=20):
    """[(read index, field, oracle value, gpu value)] for every mismatch (first `limit`)."""
    assert len(ores) == len(gres)
    bad = []
    both = (ores["mapped"] == 1) & (gres["mapped"] == 1)
    for f in COUNTER_FIELDS + MAPPED_FIELDS:
        ok = _eq(ores[f], gres[f])
        if f in MAPPED_FIELDS:
            ok |= ~both
        for i in np.nonzero(~ok)[0][:limit]:
            bad.append((int(i), f, ores[f][i].item(), gres[f][i].item()))
    bad.sort()
    return bad[:limit]


def oracle_map(segments, reads, k, h_frac, max_matches=None, threads=8, **prm):
    idx = O.OracleIndex(k, h_frac, segments, max_matches)
    p = O.params(k=k, h_frac=h_frac, max_matches=max_matches, **prm)
    return idx, p, idx.map_reads(p, reads, threads=threads)


def gpu_map(segments, reads, k, h_frac, max_matches=None, **prm):
    sk = G.FracMinHash(k, h_frac)
    idx = G.SketchIndex(sk, segments, max_matches)
    p = G.make_params(k=k, h_frac=h_frac, **prm)
    mp = G.Mapper(idx, p)
    return idx, p, mp, mp.map_batch(reads)


def split_reads(bases, off):
    return [bases[off[i]:off[i + 1]].tobytes() for i in range(len(off) - 1)]


def contigs(ref, lens, prefix="chr"):
    out, at = [], 0
    for i, n in enumerate(lens):
        out.append((f"{prefix}{i + 1}", ref[at:at + n]))
        at += n
    return out
