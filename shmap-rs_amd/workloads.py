"""The five synthetic mapping workloads, seeded and deterministic.

Each is a reference (contig lengths, satellite content) plus a read set
(count, length distribution, error profile) drawn by the seeded generator of
csrc/synth.c; the benchmark times them and the GPU parity tests map them
against the oracle.

  plain_small     10 Mbp uniform ACGT in one contig; 2,000 reads of 10 kb with
                  0.5% substitutions; k=25 r=0.01 t=0.4 Containment
  batched_uneven  100 Mbp, 20% tandem arrays (171 bp unit, 2% divergence);
                  20,000 HiFi-like reads, lognormal mean 20 kb / sd 8 kb
                  clipped to 1-80 kb, 0.1% errors (60/20/20 sub/ins/del)
  genome          3.1 Gbp of 24 contigs (50-250 Mbp), 8% satellite arrays at
                  1-3% copy divergence; 92,000 ONT-like reads, lognormal mean
                  24 kb / sd 12 kb, 7% errors (40/30/30), both strands
  dense           the genome reference sketched at r=0.10; 125,000 HiFi-like
                  reads, lognormal mean 23 kb, 0.1% errors
  serving         the genome reference at r=0.01; reads lognormal 1-100 kb
                  (mean 15 kb), an even mix of the 0.1% and 7% error
                  profiles, arriving as a Poisson stream
"""
from __future__ import annotations

import time

import numpy as np

import shmap_amd as G

HIFI = (0.001, 0.6, 0.2, 0.2)   # error rate, then the sub/ins/del split
ONT = (0.07, 0.4, 0.3, 0.3)

PARAMS = dict(k=25, theta=0.4, min_diff=0.075, max_overlap=0.3)


def genome_contigs(total=3_100_000_000, n=24, lo=50_000_000, hi=250_000_000) -> np.ndarray:
    """n contig lengths from lo to hi (a power curve) summing to total."""
    f = lambda p: np.array([lo + (hi - lo) * (i / (n - 1)) ** p for i in range(n)])
    a, b = 0.3, 5.0
    for _ in range(80):
        mid = (a + b) / 2
        if f(mid).sum() > total:
            a = mid
        else:
            b = mid
    lens = np.round(f(mid)).astype(np.int64)
    lens[n // 2] += total - lens.sum()
    return lens


WORKLOADS = {
    "plain_small": dict(lens=[10_000_000], repeat=0.0, n_reads=2000, length=(10_000, 0), lognormal=False,
                        clip=(1, 10_000), errors=((0.005, 1.0, 0.0, 0.0),), r=0.01),
    "batched_uneven": dict(lens=[100_000_000], repeat=0.20, div=(0.02, 0.02), n_units=16, n_reads=20_000,
                           length=(20_000, 8_000), lognormal=True, clip=(1_000, 80_000), errors=(HIFI,), r=0.01),
    "genome": dict(lens=None, repeat=0.08, div=(0.01, 0.03), n_units=64, n_reads=92_000,
                   length=(24_000, 12_000), lognormal=True, clip=(1_000, 200_000), errors=(ONT,), r=0.01),
    "dense": dict(lens=None, repeat=0.08, div=(0.01, 0.03), n_units=64, n_reads=125_000,
                  length=(23_000, 10_000), lognormal=True, clip=(1_000, 200_000), errors=(HIFI,), r=0.10),
}

# the serving workload's reads (its reference is the genome's)
SERVING = dict(length=(15_000, 10_000), lognormal=True, clip=(1_000, 100_000), errors=(HIFI, ONT), seed=5)


def reference(name: str, seed: int = 1):
    """(contig lengths, bases) of a workload's reference."""
    w = WORKLOADS[name]
    lens = genome_contigs() if w["lens"] is None else np.asarray(w["lens"], dtype=np.int64)
    div = w.get("div", (0.02, 0.02))
    ref = G.synth_reference(seed, lens, repeat_frac=w["repeat"], unit_len=171, div_min=div[0], div_max=div[1],
                            n_units=w.get("n_units", 16), array_len=200_000)
    return lens, ref


def reads(name: str, lens, ref, seed: int = 2, n_reads: int | None = None):
    """(bases, offsets, truth) of a workload's reads; the first n of the full
    set when n_reads is given (reads are drawn independently per index)."""
    w = WORKLOADS[name]
    return G.synth_reads(seed, ref, lens, n_reads or w["n_reads"], w["length"], lognormal=w["lognormal"],
                         clip=w["clip"], errors=w["errors"], both_strands=True)


def serving_reads(lens, ref, n_reads: int):
    s = SERVING
    return G.synth_reads(s["seed"], ref, lens, n_reads, s["length"], lognormal=s["lognormal"], clip=s["clip"],
                         errors=s["errors"], both_strands=True)


def segments(lens, ref, prefix="chr"):
    """[(name, bases view)] of the contigs."""
    out, at = [], 0
    for i, n in enumerate(lens):
        out.append((f"{prefix}{i + 1}", ref[at:at + int(n)]))
        at += int(n)
    return out


def make(name: str, seed: int = 1, n_reads: int | None = None, log=None):
    """(lens, ref, bases, off, truth, config) of a workload."""
    t0 = time.time()
    lens, ref = reference(name, seed)
    bases, off, truth = reads(name, lens, ref, seed + 1, n_reads)
    if log:
        log(f"workload {name}: ref {lens.sum() / 1e9:.3f} Gbp in {len(lens)} contigs, {len(off) - 1} reads "
            f"({off[-1] / 1e9:.3f} Gbp), generated in {time.time() - t0:.1f}s")
    return lens, ref, bases, off, truth, WORKLOADS[name]
