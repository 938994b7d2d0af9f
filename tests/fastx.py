"""Minimal FASTA reader for tests: records as needletail delivers them to the
reference (src/io/mod.rs:383-427) -- id cut at the first space or tab, line
breaks and trailing '\\r' removed."""


def read_fasta(path):
    recs = []
    name, chunks = None, []
    with open(path, "rb") as f:
        for line in f:
            line = line.rstrip(b"\n").rstrip(b"\r")
            if line.startswith(b">"):
                if name is not None:
                    recs.append((name, b"".join(chunks)))
                hdr = line[1:]
                for sep in (b" ", b"\t"):
                    hdr = hdr.split(sep)[0]
                name, chunks = hdr.decode(), []
            elif name is not None:
                chunks.append(line)
    if name is not None:
        recs.append((name, b"".join(chunks)))
    return recs


def strip_time(paf: str) -> str:
    """tests/golden_paf.rs:13-23 -- drop the wall-clock t:f: field."""
    return "\n".join("\t".join(f for f in l.split("\t") if not f.startswith("t:f:"))
                     for l in paf.strip().split("\n") if l)
