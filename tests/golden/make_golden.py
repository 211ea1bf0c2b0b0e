"""Writes the golden fixtures under tests/golden/<case>/ by running FAME's own
CPU path (built from its sources into oracle/_ref by oracle/Makefile).

    python tests/golden/make_golden.py [case ...]

Each case directory holds
    index.bin.gz, index.bin_strands.gz  the reference-built --load_index files
    reads.npz                           flat read bytes + offsets + lengths
    golden.npz                          per-read valid / MATCH from the reference,
                                        CpG counts parsed from its TSV
    cpg.tsv.gz                          the reference's <out>_cpg.tsv, verbatim
    case.json                           configuration, counters, strand tallies

Inputs are synthetic (fame_amd.synth) with fixed seeds, plus hand-placed
special reads (N, short, unknown characters, random, chromosome edges,
longer than READLEN).  The reference is run here, in the container that holds
/root/reference; the fixtures travel with the repository.
"""
import gzip
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from fame_amd import synth  # noqa: E402

CASES = {
    # name: genome spec, reference config, read spec
    "se_basic": dict(chroms=[60000, 45000], repeats=0, htab=1 << 16, readlen=100, chunk=300000,
                     n=5000, seed=11, unord=False, flip=False),
    "se_repeats": dict(chroms=[90000, 70000, 40000], repeats=24, htab=1 << 14, readlen=100,
                       chunk=300000, n=6000, seed=12, unord=False, flip=False),
    "se_unord": dict(chroms=[90000, 70000, 40000], repeats=24, htab=1 << 14, readlen=100,
                     chunk=300000, n=3000, seed=13, unord=True, flip=False, share="se_repeats"),
    "se_ga_chunked": dict(chroms=[60000, 45000], repeats=6, htab=1 << 15, readlen=100, chunk=700,
                          n=3000, seed=14, unord=False, flip=True),
    "se_rl150": dict(chroms=[80000, 50000], repeats=8, htab=1 << 15, readlen=150, chunk=300000,
                     n=4000, seed=15, unord=False, flip=False, len_range=(90, 150)),
    # reads of 120..255 letters against a READLEN 100 index: every verify pool
    # width (two to four 64-bit words) and long DP tracebacks
    "se_long": dict(chroms=[60000, 45000], repeats=0, htab=1 << 16, readlen=100, chunk=300000,
                    n=3000, seed=16, unord=False, flip=False, len_range=(120, 255), gen_len=255,
                    share="se_basic"),
}


def ref_binary(cfg):
    h = cfg["htab"]
    name = f"h{h}_r{cfg['readlen']}_c8_n{len(cfg['chroms'])}_p50-550_k{cfg['chunk']}"
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"HTAB={h}",
                    f"READLEN={cfg['readlen']}", "CORENUM=8", f"CHROMNUM={len(cfg['chroms'])}",
                    f"CHUNK={cfg['chunk']}"], check=True)
    return os.path.join(ROOT, "oracle", "_ref", f"fame_ref_{name}")


def make_genome(cfg):
    rng = np.random.default_rng(cfg["seed"] + 1000)
    chroms = synth.make_genome(cfg["chroms"], seed=cfg["seed"])
    # near-copies of a few segments (5-50 bp differences) so that reads have
    # several candidate windows and some matches are not unique
    for r in range(cfg["repeats"]):
        src = rng.integers(0, len(chroms))
        L = int(rng.integers(800, 3000))
        s = int(rng.integers(0, chroms[src].size - L))
        seg = chroms[src][s:s + L].copy()
        for _ in range(int(rng.integers(1, 4))):
            dst = rng.integers(0, len(chroms))
            d = int(rng.integers(0, chroms[dst].size - L))
            cp = seg.copy()
            k = int(rng.integers(0, max(1, L // 200)))
            idx = rng.integers(0, L, size=k)
            cp[idx] = synth._BASES[rng.integers(0, 4, size=k)]
            chroms[dst][d:d + L] = cp
    return chroms


def make_reads(cfg, chroms):
    rng = np.random.default_rng(cfg["seed"] + 2000)
    RL = cfg["readlen"]
    rs = synth.make_reads(chroms, cfg["n"], cfg.get("gen_len", RL), seed=cfg["seed"] + 3,
                          len_range=cfg.get("len_range"))
    seqs = [bytes(rs.seqs[i, :rs.lens[i]]) for i in range(len(rs))]
    n = len(seqs)
    pick = rng.permutation(n)
    k = max(4, n // 100)
    # reads with an N
    for i in pick[:k]:
        s = bytearray(seqs[i])
        s[rng.integers(0, len(s))] = ord("N")
        seqs[i] = bytes(s)
    # too short
    for i in pick[k:2 * k]:
        seqs[i] = seqs[i][:int(rng.integers(40, RL - 20))]
    # unknown characters (kept by the reference, which only reports them)
    for i in pick[2 * k:2 * k + k // 2]:
        s = bytearray(seqs[i])
        s[rng.integers(0, len(s))] = ord(rng.choice([b"n", b"X", b"a"]))
        seqs[i] = bytes(s)
    # random sequence (no origin)
    for i in pick[2 * k + k // 2:3 * k]:
        seqs[i] = bytes(synth._BASES[rng.integers(0, 4, size=RL)])
    # reads near chromosome starts and ends (>= 12 bp away, so the DP of a
    # matched read never reads outside its chromosome)
    for j, i in enumerate(pick[3 * k:4 * k]):
        c = chroms[j % len(chroms)]
        if j % 2:
            p = int(rng.integers(12, 400))
        else:
            p = c.size - RL - int(rng.integers(12, 400))
        s = c[p:p + RL].copy()
        s[s == ord("C")] = np.where(rng.random((s == ord("C")).sum()) < 0.9, ord("T"), ord("C"))
        seqs[i] = bytes(s)
    # longer than READLEN
    for i in pick[4 * k:4 * k + k // 2]:
        c = chroms[0]
        p = int(rng.integers(100, c.size - RL - 100))
        s = c[p:p + RL + 10].copy()
        s[s == ord("C")] = ord("T")
        seqs[i] = bytes(s)
    if cfg["flip"]:
        # G->A libraries: read 1 is the reverse complement of the converted strand
        comp = bytes.maketrans(b"ACGTN", b"TGCAN")
        seqs = [s.translate(comp)[::-1] for s in seqs]
    return seqs


def write_fastq(path, seqs):
    with open(path, "wb") as f:
        for i, s in enumerate(seqs):
            f.write(b"@r%d\n%s\n+\n%s\n" % (i, s, b"I" * len(s)))


def gz(src, dst):
    with open(src, "rb") as a, gzip.open(dst, "wb", compresslevel=9) as b:
        shutil.copyfileobj(a, b)


def parse_tsv(path):
    rows = [ln.split(b"\t") for ln in open(path, "rb").read().split(b"\n") if ln]
    return np.array([[int(x) for x in r[2:6]] for r in rows], dtype=np.uint16).reshape(-1, 4)


def run_case(name):
    cfg = CASES[name]
    out = os.path.join(HERE, name)
    os.makedirs(out, exist_ok=True)
    binp = ref_binary(cfg)
    with tempfile.TemporaryDirectory() as td:
        share = cfg.get("share")
        if share:
            sd = os.path.join(HERE, share)
            for f in ("index.bin", "index.bin_strands"):
                with gzip.open(os.path.join(sd, f + ".gz"), "rb") as a, open(os.path.join(td, f), "wb") as b:
                    shutil.copyfileobj(a, b)
            chroms = make_genome(CASES[share])
        else:
            chroms = make_genome(cfg)
            synth.write_fasta(os.path.join(td, "g.fa"), chroms)
            subprocess.run([binp, "build", "g.fa", "index.bin"], cwd=td, check=True,
                           stdout=subprocess.DEVNULL)
        seqs = make_reads(cfg, chroms)
        write_fastq(os.path.join(td, "r.fq"), seqs)
        args = [binp, "match", "index.bin", "r.fq", "out"] + (["--unord_reads"] if cfg["unord"] else [])
        subprocess.run(args, cwd=td, check=True, stdout=subprocess.DEVNULL)
        raw = open(os.path.join(td, "out_reads.bin"), "rb").read()
        assert raw[:8] == b"FAMERES1"
        n = int(np.frombuffer(raw[8:16], np.uint64)[0])
        valid = np.frombuffer(raw[16:16 + n], np.uint8).copy()
        mat = np.frombuffer(raw[16 + n:16 + 9 * n], np.uint64).copy()
        meth = parse_tsv(os.path.join(td, "out_cpg.tsv"))
        stats = json.load(open(os.path.join(td, "out_stats.json")))
        if not share:
            gz(os.path.join(td, "index.bin"), os.path.join(out, "index.bin.gz"))
            gz(os.path.join(td, "index.bin_strands"), os.path.join(out, "index.bin_strands.gz"))
        gz(os.path.join(td, "out_cpg.tsv"), os.path.join(out, "cpg.tsv.gz"))
    lens = np.array([len(s) for s in seqs], np.uint32)
    offs = np.zeros(len(seqs), np.uint64)
    offs[1:] = np.cumsum(lens.astype(np.uint64))[:-1]
    flat = np.frombuffer(b"".join(seqs), np.uint8)
    np.savez_compressed(os.path.join(out, "reads.npz"), seqs=flat, offs=offs, lens=lens)
    np.savez_compressed(os.path.join(out, "golden.npz"), valid=valid, mat=mat, meth=meth)
    meta = {k: v for k, v in cfg.items() if k != "len_range"}
    meta["len_range"] = cfg.get("len_range")
    meta["reference"] = {k: stats[k] for k in ("reads", "succ", "nonunique", "unsucc", "r1_fwd",
                                                "r1_rev", "match_r1_fwd", "htabsize", "readlen",
                                                "chunksize")}
    meta["index_dir"] = share or name
    json.dump(meta, open(os.path.join(out, "case.json"), "w"), indent=1)
    print(name, meta["reference"])


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES)
    for nm in names:
        run_case(nm)
