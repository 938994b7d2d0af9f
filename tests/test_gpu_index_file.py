"""The persistent index: an index written by shm_index_save and read back by
shm_index_load equals the built one -- statistics, segments, hit lists of
sampled k-mers and every mapping record, bit for bit -- and the CLI maps the
golden fixture the same from a saved index as from the FASTA."""
import os
import subprocess

import numpy as np
import pytest

import fastx
import parity_util as P
import shmap_amd as G

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "shmap-rs_amd", "lib", "shmap")


@pytest.mark.parametrize("max_matches", [None, 3])
def test_saved_index_maps_as_the_built_one(tmp_path, max_matches):
    lens = [300_000, 120_000, 5]
    ref = G.synth_reference(81, lens, repeat_frac=0.15, unit_len=171, div_min=0.01, div_max=0.03, n_units=4,
                            array_len=20_000)
    bases, off, _ = G.synth_reads(82, ref, lens[:2], 120, (3000, 1200), lognormal=True, clip=(100, 12_000),
                                  errors=((0.002, 0.6, 0.2, 0.2), (0.06, 0.4, 0.3, 0.3)))
    segs = P.contigs(ref, lens)
    reads = P.split_reads(bases, off)
    built = G.SketchIndex(G.FracMinHash(15, 0.1), segs, max_matches)
    path = tmp_path / "ref.idx"
    built.save(path)
    loaded = G.SketchIndex.load(path)
    assert loaded.params() == (15, 0.1, max_matches or 0)
    assert loaded.stats() == built.stats()
    assert loaded.segments() == built.segments()
    sk = G.FracMinHash(15, 0.1).sketch_slice_into(bytes(segs[0][1][:50_000]))
    for h in np.unique(sk["h"])[:200]:
        assert loaded.hits(int(h)) == built.hits(int(h))
    prm = G.make_params(k=15, h_frac=0.1, theta=0.4, min_diff=0.075, max_overlap=0.3)
    want = G.Mapper(built, prm).map_batch(reads)
    got = G.Mapper(loaded, prm).map_batch(reads)
    bad = P.diff_results(want, got)
    assert not bad, bad
    assert want["mapped"].sum() > 0


def _run(args, cwd):
    return subprocess.run([CLI, *args], capture_output=True, cwd=cwd, timeout=120)


def test_cli_maps_from_a_saved_index(golden_dir, tmp_path):
    ref = os.path.join(golden_dir, "tiny_ref.fa")
    reads = os.path.join(golden_dir, "tiny_reads.fa")
    opts = ["-k", "8", "-r", "1.0", "-t", "0.1"]
    r = _run(["-s", ref, "-p", reads, *opts, "--save-index", "tiny.idx"], tmp_path)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    want = fastx.strip_time(open(os.path.join(golden_dir, "tiny.golden.paf")).read())
    assert fastx.strip_time(r.stdout.decode()) == want
    r = _run(["-s", "tiny.idx", "-p", reads, *opts], tmp_path)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    assert fastx.strip_time(r.stdout.decode()) == want
    r = _run(["-s", "tiny.idx", "-p", reads, "-k", "9", "-r", "1.0", "-t", "0.1"], tmp_path)
    assert r.returncode == 1 and b"was built with -k 8" in r.stderr
