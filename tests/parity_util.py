"""Shared helpers for parity tests: run the CPU oracle and the HIP library on
the same inputs and diff every field of every read."""
import math

import numpy as np

import oracle_ffi as O
import shmap_amd as G

COUNTER_FIELDS = ["mapped", "unmappable_short", "read_len", "m", "u", "S", "possible_matches",
                  "seed_matches", "total_matches", "max_seed_matches", "seeded_buckets",
                  "refined_buckets", "refine_memo_hits", "final_buckets"]
MAPPED_FIELDS = ["p_start", "p_end", "strand", "segm_id", "t_l", "t_r", "mapq", "s_sz",
                 "intersection", "intersection2", "same_strand", "b_segm", "b_b", "b2_segm", "b2_b",
                 "J", "J2", "sh", "sigmas", "match_inefficiency"]


def same_value(a, b) -> bool:
    a, b = a.item() if hasattr(a, "item") else a, b.item() if hasattr(b, "item") else b
    if isinstance(a, float) or isinstance(b, float):
        a, b = float(a), float(b)
        if math.isnan(a) and math.isnan(b):
            return True
        return a == b and math.copysign(1.0, a) == math.copysign(1.0, b)
    return int(a) == int(b)


def diff_results(ores, gres, limit=20):
    bad = []
    for i in range(len(ores)):
        o, g = ores[i], gres[i]
        fields = COUNTER_FIELDS + (MAPPED_FIELDS if o["mapped"] and g["mapped"] else [])
        for f in fields:
            if not same_value(o[f], g[f]):
                bad.append((i, f, o[f].item(), g[f].item()))
        if len(bad) >= limit:
            break
    return bad


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
