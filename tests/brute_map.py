"""A brute-force statement of what map_read decides, for checking the oracle.

It shares no code with the oracle's mapping path.  From the read's and the
segments' sketches (whose computation is pinned separately against the
direct ntHash definition) it builds its own hash -> hits table, then states
each decision of src/shmap/mod.rs:653-971 in the closed form the mathematics
gives it rather than by the reference's sequential sweep:

* every seeded bucket is scored on its own: Containment/Jaccard by
  enumerating every window of best_fixed_length (src/shmap/scoring.rs:189-291),
  bucket_SH from the bucket's fully extended content, bucket_LCS by an O(n^2)
  longest-increasing-subsequence table (not patience sorting);
* a bucket's fully extended seed heuristic is matches/m over all seeds
  (src/shmap/pruning.rs:10-119), so pruning is sound -- no bucket scores above
  it -- and match_rest's first pass (src/shmap/scoring.rs:367-489) returns the
  FIRST bucket in sorted order among those with the greatest score above theta
  (the threshold ratchets to each accepted score, acceptance is strict);
* the second pass returns the first bucket with the greatest score above
  J(1-d) among those overlapping the best by less than max_overlap;
* final_buckets counts, per pass, the buckets whose score beats every earlier
  accepted score and the pass's starting threshold (the "records");
* refined_buckets / refine_memo_hits come from the two passes' survivor sets,
  a bucket surviving a pass iff its full heuristic is at least the threshold
  standing when the sweep reaches it (the heuristic only falls as seeds are
  added);
* sigma, MAPQ and the printed J2 from src/mapping.rs:331-361.
"""
from __future__ import annotations

import math

import numpy as np

import oracle_ffi as O

METRICS = ("Containment", "Jaccard", "bucket_SH", "bucket_LCS")
MIN_HALFLEN = 5  # src/buckets.rs:15


class BruteIndex:
    """hash -> [(segm_id, tpos, r, strand)] in (segm_id, r) order, from the
    segments' sketches."""

    def __init__(self, k, h_frac, segments):
        self.t = [O.sketch(k, h_frac, s) for _, s in segments]
        hits = {}
        for sid, t in enumerate(self.t):
            for tpos in range(len(t)):
                hits.setdefault(int(t["h"][tpos]), []).append((sid, tpos, int(t["r"][tpos]), int(t["strand"][tpos])))
        self.hits = hits

    def count(self, h):
        return len(self.hits.get(h, ()))


def _lis_len(vals):
    """Longest strictly increasing subsequence by the O(n^2) table."""
    a = np.asarray(vals, dtype=np.int64)
    n = len(a)
    if n == 0:
        return 0
    best = np.ones(n, dtype=np.int64)
    for i in range(1, n):
        smaller = a[:i] < a[i]
        if smaller.any():
            best[i] = best[:i][smaller].max() + 1
    return int(best.max())


def _overlap(a, b):
    """Fraction of two mappings' [t_l, t_r] spans that overlap (src/mapping.rs:411-420)."""
    if a["segm"] != b["segm"]:
        return -0.0
    cap = max(0, min(a["t_r"], b["t_r"]) - max(a["t_l"], b["t_l"]))
    cup = max(a["t_r"], b["t_r"]) - min(a["t_l"], b["t_l"])
    return cap / cup if cup else math.nan


def brute_map(idx: BruteIndex, k, h_frac, read, theta, min_diff, max_overlap, metric, nbp=False, one_sweep=False):
    rs = O.sketch(k, h_frac, read)
    m = len(rs)
    out = {"m": m, "mapped": 0}
    if m < MIN_HALFLEN:
        return out
    occs, strand, ppos = {}, {}, {}
    for j in np.argsort(rs["r"], kind="stable"):
        h = int(rs["h"][j])
        occs[h] = occs.get(h, 0) + 1
        strand.setdefault(h, int(rs["strand"][j]))      # the smallest-r occurrence's strand
        ppos.setdefault(h, []).insert(0, int(rs["r"][j]))  # positions, descending
    hashes = sorted(occs, key=lambda h: (idx.count(h), h))  # rarest first, ties by hash
    u = len(hashes)
    theta2 = theta - min_diff
    S = u if one_sweep else int((1.0 - theta2) * m) + 1
    consumed, seen = [], 0
    for h in hashes:
        if seen >= S:
            break
        consumed.append(h)
        seen += occs[h]
    hl = m

    def content(hs, sid, b):
        """(matches, codirection, r_min, r_max) of bucket (sid, b) over seeds hs."""
        lo, hi = b * hl, (b + 2) * hl
        mt = cd = 0
        rmin, rmax = 2**31 - 1, -(2**31)
        for h in hs:
            inside = [x for x in idx.hits.get(h, ()) if x[0] == sid and lo <= x[1] < hi]
            mt += min(len(inside), occs[h])
            for x in inside:
                cd += 1 if x[3] == strand[h] else -1
                rmin, rmax = min(rmin, x[2]), max(rmax, x[2])
        return mt, cd, rmin, rmax

    buckets = set()
    for h in consumed:
        for (sid, tpos, _, _) in idx.hits.get(h, ()):
            buckets.add((sid, tpos // hl))
            if tpos // hl > 0:
                buckets.add((sid, tpos // hl - 1))
    first = {bk: content(consumed, *bk) for bk in buckets}
    order = sorted(buckets, key=lambda bk: (-first[bk][0], bk[0], bk[1]))
    out["seeded_buckets"] = len(order)
    out["S"] = S

    cands = []
    for bk in order:
        sid, b = bk
        full = content(hashes, sid, b)
        sh_full = 1.0 - (m - full[0]) / m
        cont = first[bk] if nbp else full        # -P never extends a bucket
        sh = 1.0 if nbp else sh_full
        t = idx.t[sid]
        lo, end = max(0, b * hl), min(len(t), (b + 2) * hl)
        c = {"bk": bk, "segm": sid, "sh_full": sh_full, "sh": sh}
        if metric in ("Containment", "Jaccard"):
            c.update(_best_window(t, occs, strand, lo, end, m, metric))
        else:
            if metric == "bucket_SH":
                score = sh
            else:
                pp = [p for q in range(lo, end) for p in ppos.get(int(t["h"][q]), ())]
                score = max(_lis_len(pp), _lis_len(pp[::-1])) / m
            c.update(score=score, I=cont[0], same=cont[1], t_l=cont[2], t_r=cont[3])
        cands.append(c)

    def records(start, allowed):
        """Buckets whose score beats the pass's start and every earlier
        allowed score (the pass's accepted maxima)."""
        n, top = 0, start
        for c in cands:
            if c["score"] > top:
                n += 1
            if allowed(c) and c["score"] > top:
                top = c["score"]
        return n

    def survivors(start, allowed):
        s, top = set(), start
        for i, c in enumerate(cands):
            if nbp or c["sh_full"] >= top:
                s.add(i)
            if allowed(c) and c["score"] > top:
                top = c["score"]
        return s

    def first_best(start, allowed):
        ok = [c for c in cands if allowed(c) and c["score"] > start]
        if not ok:
            return None
        top = max(c["score"] for c in ok)
        return next(c for c in ok if c["score"] == top)

    anything = lambda c: True
    best = first_best(theta, anything)
    final = records(theta, anything)
    s1 = survivors(theta, anything)
    s2 = set()
    best2 = None
    if best is not None:
        thr2 = best["score"] * (1.0 - min_diff)
        apart = lambda c: _overlap(c, best) < max_overlap
        best2 = first_best(thr2, apart)
        final += records(thr2, apart)
        s2 = survivors(thr2, apart)
    memo = metric in ("Containment", "Jaccard")
    out["final_buckets"] = final
    out["refined_buckets"] = len(s1) + (len(s2 - s1) if memo else len(s2))
    out["refine_memo_hits"] = len(s2 & s1) if memo else 0
    if best is None:
        return out
    I = best["I"]
    I2 = best2["I"] if best2 is not None else -1
    J2 = best2["score"] if best2 is not None else -1.0
    if J2 < theta2:
        J2 = theta2
    frac = 1.0 - J2 / best["score"]
    strand5 = 5 if abs(best["same"]) < int(I / 2) else 60
    sig = -1.0
    if best2 is not None:
        x, y = (0 if I2 == -1 else I2), (0 if I == -1 else I)
        sig = abs(x - y) / math.sqrt(x + y)
    out.update(mapped=1, b_segm=best["segm"], b_b=best["bk"][1], J=best["score"], intersection=I,
               same_strand=best["same"], t_l=best["t_l"], t_r=best["t_r"], sh=best["sh"],
               b2_segm=best2["segm"] if best2 is not None else -1, b2_b=best2["bk"][1] if best2 is not None else -1,
               intersection2=I2, J2=J2, sigmas=sig, mapq=min(60, strand5) if frac > min_diff else 0)
    return out


def _best_window(t, occs, strand, lo, end, m, metric):
    """The first window [l, min(l+m, end)) of greatest score over l in [lo, end)."""
    th = t["h"]
    pos = {}
    for q in range(lo, end):
        h = int(th[q])
        if h in occs:
            pos.setdefault(h, []).append(q)
    ls = np.arange(lo, end, dtype=np.int64)
    rs = np.minimum(ls + m, end)
    inter = np.zeros(len(ls), dtype=np.int64)
    sign = np.zeros(end - lo + 1, dtype=np.int64)
    for h, ps in pos.items():
        ps = np.asarray(ps, dtype=np.int64)
        cnt = np.searchsorted(ps, rs) - np.searchsorted(ps, ls)
        inter += np.minimum(cnt, occs[h])
        for q in ps:
            sign[q - lo + 1] = 1 if int(t["strand"][q]) == strand[h] else -1
    cs = np.cumsum(sign)
    same = cs[rs - lo] - cs[ls - lo]
    best = None
    for i in range(len(ls)):
        I = int(inter[i])
        n = int(rs[i] - ls[i])
        score = I / m if metric == "Containment" else I / (m + n - I)
        if best is None or score > best[0]:
            best = (score, I, int(same[i]), int(t["r"][ls[i]]), int(t["r"][rs[i] - 1]))
    return dict(zip(("score", "I", "same", "t_l", "t_r"), best))
