"""The shmap command line end to end on the GPU: the reference's golden
fixture (tests/golden_paf.rs:26-53), plain and gzip-compressed, and a
multi-batch run whose PAF lines, unmapped list and per-read rows equal the
library's records."""
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


def _run(args, cwd):
    r = subprocess.run([CLI, *args], capture_output=True, cwd=cwd, timeout=120)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    return r.stdout.decode()


def test_cli_golden_paf(golden_dir, tmp_path):
    out = _run(["-s", os.path.join(golden_dir, "tiny_ref.fa"), "-p", os.path.join(golden_dir, "tiny_reads.fa"),
                "-k", "8", "-r", "1.0", "-t", "0.1"], tmp_path)
    want = fastx.strip_time(open(os.path.join(golden_dir, "tiny.golden.paf")).read())
    assert fastx.strip_time(out) == want


def test_cli_golden_paf_from_gzip(golden_dir, tmp_path):
    """The same fixture with both inputs gzip-compressed."""
    import gzip
    for nm in ("tiny_ref.fa", "tiny_reads.fa"):
        (tmp_path / (nm + ".gz")).write_bytes(gzip.compress(open(os.path.join(golden_dir, nm), "rb").read()))
    out = _run(["-s", "tiny_ref.fa.gz", "-p", "tiny_reads.fa.gz", "-k", "8", "-r", "1.0", "-t", "0.1"], tmp_path)
    want = fastx.strip_time(open(os.path.join(golden_dir, "tiny.golden.paf")).read())
    assert fastx.strip_time(out) == want


def _write_fasta(path, names, seqs, width):
    with open(path, "wb") as f:
        for n, s in zip(names, seqs):
            s = bytes(s)
            f.write(b">" + n.encode() + b" synthetic\n")
            for i in range(0, len(s), width):
                f.write(s[i:i + width] + b"\n")


def test_cli_batches_match_the_library(tmp_path):
    lens = [220_000, 180_000]
    ref = G.synth_reference(71, lens, repeat_frac=0.1, unit_len=171, div_min=0.01, div_max=0.03, n_units=4,
                            array_len=20_000)
    bases, off, _ = G.synth_reads(72, ref, lens, 150, (3000, 1200), lognormal=True, clip=(100, 12_000),
                                  errors=((0.002, 0.6, 0.2, 0.2), (0.08, 0.4, 0.3, 0.3)))
    segs = P.contigs(ref, lens)
    reads = P.split_reads(bases, off)
    names = [f"read{i}" for i in range(len(reads))]
    _write_fasta(tmp_path / "ref.fa", [n for n, _ in segs], [s for _, s in segs], 70)
    _write_fasta(tmp_path / "reads.fa", names, reads, 61)
    out = _run(["-s", "ref.fa", "-p", "reads.fa", "-k", "15", "-r", "0.1", "-t", "0.4", "-d", "0.075", "-o", "0.3",
                "--batch", "37", "-z", "run", "--per-read-stats", "stats.tsv"], tmp_path)
    idx = G.SketchIndex(G.FracMinHash(15, 0.1), segs, None)
    prm = G.make_params(k=15, h_frac=0.1, theta=0.4, min_diff=0.075, max_overlap=0.3)
    res = G.Mapper(idx, prm).map_batch(reads)
    want = [G.format_paf(idx, prm, n, r) for n, r in zip(names, res) if r["mapped"]]
    assert int(res["mapped"].sum()) > 0
    assert fastx.strip_time(out) == fastx.strip_time("\n".join(want))
    unmapped = [l.split("\t")[0] for l in open(tmp_path / "run.unmapped.paf").read().splitlines()]
    assert unmapped == [n for n, r in zip(names, res) if not r["mapped"]]
    rows = open(tmp_path / "stats.tsv").read().splitlines()
    assert len(rows) == len(reads) + 1
    kmers = np.array([int(r.split("\t")[2]) for r in rows[1:]])
    assert np.array_equal(kmers, res["m"])


ORACLE_CLI = os.path.join(ROOT, "oracle", "lib", "shmap_oracle")


@pytest.mark.parametrize("kind", ["uneven", "noisy"])
def test_side_outputs_match_the_oracle_cli(tmp_path, kind):
    """--per-read-stats rows (every column but query_mapping_s,
    src/shmap/mod.rs:280-327) and <-z>.unmapped.paf (src/shmap/mod.rs:388-411,
    src/mapping.rs:82-84) of the GPU CLI against the oracle's CLI, with the
    benchmark parameters (k=25 r=0.01 t=0.4 -d 0.075 -o 0.3): HiFi-like reads
    of uneven length on a reference with 20% tandem arrays, and ONT-like reads
    with 7% errors on a satellite-rich reference, where most reads are
    unmapped and the side outputs carry the evidence."""
    if kind == "uneven":
        lens = [12_000_000, 8_000_000]
        ref = G.synth_reference(91, lens, repeat_frac=0.2, unit_len=171, div_min=0.02, div_max=0.02, n_units=16,
                                array_len=200_000)
        bases, off, _ = G.synth_reads(92, ref, lens, 2000, (20_000, 8_000), lognormal=True, clip=(1_000, 80_000),
                                      errors=((0.001, 0.6, 0.2, 0.2),))
    else:
        lens = [16_000_000, 9_000_000, 5_000_000]
        ref = G.synth_reference(93, lens, repeat_frac=0.08, unit_len=171, div_min=0.01, div_max=0.03, n_units=64,
                                array_len=200_000)
        bases, off, _ = G.synth_reads(94, ref, lens, 3000, (24_000, 12_000), lognormal=True, clip=(1_000, 200_000),
                                      errors=((0.07, 0.4, 0.3, 0.3),))
    segs = P.contigs(ref, lens)
    reads = P.split_reads(bases, off)
    names = [f"r{i}" for i in range(len(reads))]
    _write_fasta(tmp_path / "ref.fa", [n for n, _ in segs], [s for _, s in segs], 80)
    _write_fasta(tmp_path / "reads.fa", names, reads, 100)
    opts = ["-s", "ref.fa", "-p", "reads.fa", "-k", "25", "-r", "0.01", "-t", "0.4", "-d", "0.075", "-o", "0.3"]
    outs = {}
    for who, exe in (("gpu", CLI), ("oracle", ORACLE_CLI)):
        for sample in ("1", "10"):
            tag = f"{who}{sample}"
            r = subprocess.run([exe, *opts, "-z", tag, "--per-read-stats", f"{tag}.tsv", "--per-read-stats-sample",
                                sample, "-@", "4"], capture_output=True, cwd=tmp_path, timeout=300)
            assert r.returncode == 0, r.stderr.decode()[-2000:]
            rows = [l.rsplit("\t", 1)[0] for l in (tmp_path / f"{tag}.tsv").read_text().splitlines()]
            outs[tag] = (fastx.strip_time(r.stdout.decode()), (tmp_path / f"{tag}.unmapped.paf").read_text(), rows)
    for sample in ("1", "10"):
        g, o = outs["gpu" + sample], outs["oracle" + sample]
        assert g[0] == o[0]  # PAF
        assert g[1] == o[1]  # unmapped list
        assert g[2] == o[2]  # per-read rows
    paf, unmapped, rows = outs["gpu1"]
    assert len(rows) == len(reads) + 1 and len(outs["gpu10"][2]) == (len(reads) + 9) // 10 + 1
    n_mapped = len([l for l in paf.splitlines() if l])
    assert n_mapped + len(unmapped.splitlines()) == len(reads)
    assert (n_mapped > 1500) if kind == "uneven" else (len(unmapped.splitlines()) > 2500)
