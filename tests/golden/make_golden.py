"""Writes the golden fixtures under tests/golden/<case>/ by running FAME's own
CPU path (built from its sources into oracle/_ref by oracle/Makefile).

    python tests/golden/make_golden.py [case ...]

Each single-end case directory holds
    index.bin.gz, index.bin_strands.gz  the reference-built --load_index files
    reads.npz                           flat read bytes + offsets + lengths
    golden.npz                          per-read valid / MATCH from the reference,
                                        CpG counts parsed from its TSV
    cpg.tsv.gz                          the reference's <out>_cpg.tsv, verbatim
    case.json                           configuration, counters, strand tallies

Paired-end cases (pe_*) hold reads.npz / golden.npz with per-mate arrays
(seqs1/offs1/lens1, seqs2/..., valid1/valid2/mat1/mat2) and the five paired
counters in case.json.

Single-cell cases (sc_*) hold per-cell FASTQs cut from another fixture's
reads, the cells.txt list FAME's single-cell driver reads, and the count
matrix it wrote (sc_out.txt.gz).

The fastq/ directory holds FASTQ inputs with unusual layouts (CRLF, no
final newline, truncated records, blank lines, gzip, chunk-size multiples)
and, per input, what FAME's parseChunk / parseChunkGZ (CHUNKSIZE 700) read
from it: <name>.expect.npz with lens, seqs and the (procReads, more) of each
call.

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
    # reads of 256..1200 letters (beyond the device's short-read kernels) mixed
    # with the usual special reads of READLEN letters, and reads of 2 100 to
    # 5 000 letters, longer than any window text plus the errors allowed
    "se_xlong": dict(chroms=[60000, 45000], repeats=0, htab=1 << 16, readlen=100, chunk=300000,
                     n=2500, seed=19, unord=False, flip=False, len_range=(256, 1200), gen_len=1200,
                     share="se_basic", xlong=True),
    # an index with a blacklist: KMERCUTOFF 24 and planted poly-A / poly-T runs
    # and motif families (16 A then a 16-letter tail), so that filterHashTable
    # (RefGenome.cpp:2267-2352) removes k-mers and filteredKmers holds codes
    # below 2^32, the only ones the paired path's u32 kSeq can find
    # (ReadQueue.cpp:2527-2551, 2643, 2974); 20 % of the reads cover a motif
    "se_filt": dict(chroms=[70000, 50000], repeats=4, htab=1 << 14, readlen=100, chunk=300000,
                    n=4000, seed=17, unord=False, flip=False, cut=24, motifs=True),
    # the same genome indexed with --no_loss (no blacklist: the motif buckets
    # keep every entry)
    "se_noloss": dict(chroms=[70000, 50000], repeats=4, htab=1 << 14, readlen=100, chunk=300000,
                      n=2000, seed=18, unord=False, flip=False, cut=24, motifs=True, noloss=True,
                      genome_of="se_filt"),
}


# Paired-end cases.  Each shares the genome and index of a single-end case
# (MINPDIST / MAXPDIST do not enter the index); `gap` is the distance between
# the mates' inner ends, `flip` swaps the mates (read 1 from the G->A strand).
PAIRED = {
    "pe_basic": dict(share="se_basic", n=3000, seed=21, chunk=300000, minpd=50, maxpd=550,
                     gap=(20, 400), unord=False, flip=False),
    "pe_repeats": dict(share="se_repeats", n=4000, seed=22, chunk=300000, minpd=50, maxpd=550,
                       gap=(20, 450), unord=False, flip=False),
    "pe_unord": dict(share="se_repeats", n=2000, seed=23, chunk=300000, minpd=50, maxpd=550,
                     gap=(20, 400), unord=True, flip=False),
    "pe_rl150": dict(share="se_rl150", n=3000, seed=24, chunk=300000, minpd=20, maxpd=400,
                     gap=(20, 400), unord=False, flip=False, len_range=(100, 150)),
    "pe_ga_chunked": dict(share="se_ga_chunked", n=2500, seed=25, chunk=700, minpd=50, maxpd=550,
                          gap=(20, 400), unord=False, flip=True),
    # pairs over the blacklisted index: the filteredKmers lookup of both
    # mates' k-mers; a fifth of the fragments cover a planted motif
    "pe_filt": dict(share="se_filt", n=3000, seed=26, chunk=300000, minpd=50, maxpd=550,
                    gap=(20, 400), unord=False, flip=False),
    "pe_noloss": dict(share="se_noloss", n=1500, seed=27, chunk=300000, minpd=50, maxpd=550,
                      gap=(20, 400), unord=False, flip=False),
}


def ref_binary(cfg, minpd=50, maxpd=550):
    h = cfg["htab"]
    cut = cfg.get("cut", 1500)
    name = f"h{h}_r{cfg['readlen']}_c8_n{len(cfg['chroms'])}_p{minpd}-{maxpd}_k{cfg['chunk']}"
    if cut != 1500:
        name += f"_x{cut}"
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"HTAB={h}",
                    f"READLEN={cfg['readlen']}", "CORENUM=8", f"CHROMNUM={len(cfg['chroms'])}",
                    f"CHUNK={cfg['chunk']}", f"MINPD={minpd}", f"MAXPD={maxpd}", f"CUT={cut}"], check=True)
    return os.path.join(ROOT, "oracle", "_ref", f"fame_ref_{name}")


# motif families planted by plant_motifs: 16 A then these tails.  The paired
# path's kSeq is the unspaced code of a k-mer's last 16 letters, a filtered
# code the spaced code of all 32 (SEED position 30 is a gap), so a tail
# whose letter 14 is A can be found (and blacklists every read k-mer ending
# in it), one whose letter 14 is G never
MOTIF_TAILS = (b"GTTCAGTCGATCAAAG", b"CATGGTTACGTCTTGA", b"TTGACCAGTAGCAGAT")


def plant_motifs(cfg, chroms):
    """Plants poly-A and poly-T runs and copies of the motif families in
    place; returns the sites as (chromosome, start, length)."""
    rng = np.random.default_rng(cfg["seed"] + 3000)
    sites = []
    # (AG)^8 is a tail too: its code is blacklisted through the copies of
    # A^16 (AG)^8, and runs of (AG)^24 elsewhere give reads whose k-mers
    # straddling a run's start end in it every second letter
    runs = ([b"A" * 48] * 30 + [b"T" * 48] * 30 + [b"A" * 16 + t for t in MOTIF_TAILS for _ in range(80)]
            + [b"A" * 16 + b"AG" * 8] * 80 + [b"A" * 16 + b"GA" * 8] * 80 + [b"AG" * 24] * 30)
    for r in runs:
        c = int(rng.integers(0, len(chroms)))
        p = int(rng.integers(200, chroms[c].size - len(r) - 200))
        chroms[c][p:p + len(r)] = np.frombuffer(r, np.uint8)
        sites.append((c, p, len(r)))
    return sites


def make_genome(cfg, sites=None):
    if cfg.get("genome_of"):
        return make_genome(CASES[cfg["genome_of"]], sites)
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
    if cfg.get("motifs"):
        st = plant_motifs(cfg, chroms)
        if sites is not None:
            sites.extend(st)
    return chroms


def motif_read(rng, chroms, sites, L):
    """A read of L letters over a planted site from either strand (the
    reverse complement taken before the conversion: a directional library),
    C->T converted at 90 %."""
    c, p, n = sites[int(rng.integers(0, len(sites)))]
    s0 = p + n - int(rng.integers(n // 2, L))
    s0 = min(max(s0, 0), chroms[c].size - L)
    s = chroms[c][s0:s0 + L].copy()
    if rng.random() < 0.5:
        s = synth.revcomp(s)
    s[s == ord("C")] = np.where(rng.random((s == ord("C")).sum()) < 0.9, ord("T"), ord("C"))
    return s


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
    if cfg.get("xlong"):
        # longer than a window text can hold (WINLEN 2048 + errors): never matched
        for j, i in enumerate(pick[5 * k:5 * k + 12]):
            c = chroms[j % len(chroms)]
            L = int(rng.integers(2100, 5000))
            p = int(rng.integers(0, c.size - L))
            s = c[p:p + L].copy()
            s[s == ord("C")] = ord("T")
            seqs[i] = bytes(s)
    if cfg.get("motifs"):
        sites = []
        make_genome(cfg, sites)
        for i in pick[5 * k:5 * k + n // 5]:
            seqs[i] = bytes(motif_read(rng, chroms, sites, RL))
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
            subprocess.run([binp, "build", "g.fa", "index.bin"] + (["--no_loss"] if cfg.get("noloss") else []),
                           cwd=td, check=True, stdout=subprocess.DEVNULL)
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


def _ct(s):
    s = s.copy()
    s[s == ord("C")] = ord("T")
    return s


def make_pair_reads(pc, gcfg, chroms):
    """Mates from synth.make_pairs plus hand-placed special pairs."""
    rng = np.random.default_rng(pc["seed"] + 2000)
    RL = gcfg["readlen"]
    a, b = synth.make_pairs(chroms, pc["n"], RL, gap=pc["gap"], seed=pc["seed"] + 3,
                            len_range=pc.get("len_range"))
    s1 = [bytes(a.seqs[i, :a.lens[i]]) for i in range(len(a))]
    s2 = [bytes(b.seqs[i, :b.lens[i]]) for i in range(len(b))]
    n = len(s1)
    pick = iter(rng.permutation(n).tolist())
    k = max(4, n // 100)
    short = int(np.ceil(RL * 0.75))

    def take(m):
        return [next(pick) for _ in range(m)]

    def put_n(s):
        s = bytearray(s)
        s[rng.integers(0, len(s))] = ord("N")
        return bytes(s)

    def put_unknown(s):
        s = bytearray(s)
        s[rng.integers(0, len(s))] = ord(rng.choice([b"n", b"X", b"a"]))
        return bytes(s)

    def rand_read():
        return bytes(synth._BASES[rng.integers(0, 4, size=RL)])

    for i in take(k):                      # an N in mate 1
        s1[i] = put_n(s1[i])
    for i in take(k):                      # an N in mate 2
        s2[i] = put_n(s2[i])
    for i in take(k // 2):                 # N in both
        s1[i], s2[i] = put_n(s1[i]), put_n(s2[i])
    for i in take(k):                      # mate 1 too short (and just long enough)
        s1[i] = s1[i][:int(rng.integers(30, short))]
    for i in take(k // 2):
        s1[i] = s1[i][:short]
    for i in take(k // 2):                 # mate 2 too short
        s2[i] = s2[i][:int(rng.integers(30, short))]
    for i in take(k // 2):                 # both too short
        s1[i], s2[i] = s1[i][:40], s2[i][:int(rng.integers(30, short))]
    for i in take(k // 2):                 # too short plus an N
        s1[i], s2[i] = s1[i][:50], put_n(s2[i])
    for i in take(k // 2):                 # unknown characters
        s1[i] = put_unknown(s1[i])
    for i in take(k // 2):
        s2[i] = put_unknown(s2[i])
    for i in take(k):                      # mate 2 without origin
        s2[i] = rand_read()
    for i in take(k // 2):                 # mate 1 without origin
        s1[i] = rand_read()
    for i in take(k // 2):                 # neither has an origin
        s1[i], s2[i] = rand_read(), rand_read()
    for j, i in enumerate(take(k)):        # mates far apart or on different chromosomes
        c = chroms[(j + 1) % len(chroms)]
        p = int(rng.integers(20, c.size - RL - 20))
        s2[i] = bytes(synth.revcomp(_ct(c[p:p + RL])))
    for j, i in enumerate(take(k)):        # fragments at chromosome ends (>= 12 bp in)
        c = chroms[j % len(chroms)]
        g = int(rng.integers(pc["gap"][0], pc["gap"][1] + 1))
        F = 2 * RL + g
        p = int(rng.integers(12, 300)) if j % 2 else c.size - F - int(rng.integers(12, 300))
        fr = _ct(c[p:p + F])
        s1[i], s2[i] = bytes(fr[:RL]), bytes(synth.revcomp(fr[F - RL:]))
    for i in take(k // 2):                 # mates longer than READLEN
        c = chroms[0]
        L1, L2 = RL + int(rng.integers(1, 40)), RL + int(rng.integers(1, 40))
        g = int(rng.integers(pc["gap"][0], pc["gap"][1] + 1))
        F = L1 + g + L2
        p = int(rng.integers(100, c.size - F - 100))
        fr = _ct(c[p:p + F])
        s1[i], s2[i] = bytes(fr[:L1]), bytes(synth.revcomp(fr[F - L2:]))
    for i in take(k // 2):                 # mate distance just inside / outside the limit
        c = chroms[1 % len(chroms)]
        g = pc["maxpd"] - int(rng.integers(-3, 4))
        F = 2 * RL + g
        p = int(rng.integers(100, c.size - F - 100))
        fr = _ct(c[p:p + F])
        s1[i], s2[i] = bytes(fr[:RL]), bytes(synth.revcomp(fr[F - RL:]))
    if gcfg.get("motifs"):
        # fragments over a planted site (either mate may hold it)
        sites = []
        make_genome(gcfg, sites)
        for i in take(n // 5):
            c, p, w = sites[int(rng.integers(0, len(sites)))]
            g = int(rng.integers(pc["gap"][0], pc["gap"][1] + 1))
            F = 2 * RL + g
            s0 = p + w - int(rng.integers(w // 2, RL)) - (F - RL if rng.random() < 0.5 else 0)
            s0 = min(max(s0, 0), chroms[c].size - F)
            fr = chroms[c][s0:s0 + F].copy()
            if rng.random() < 0.5:  # the bottom strand, converted after the complement
                fr = synth.revcomp(fr)
            fr[fr == ord("C")] = np.where(rng.random((fr == ord("C")).sum()) < 0.9, ord("T"), ord("C"))
            s1[i], s2[i] = bytes(fr[:RL]), bytes(synth.revcomp(fr[F - RL:]))
        # mate 1 starting 16 letters before an (AG)^24 run: its first k-mer
        # ends in (AG)^8, whose code is blacklisted, and k-mer 0 is never
        # looked up (ReadQueue.cpp:2527-2555)
        ag = [x for x in sites if x[2] == 48 and bytes(chroms[x[0]][x[1]:x[1] + 2]) == b"AG"]
        for i in take(n // 20):
            c, p, w = ag[int(rng.integers(0, len(ag)))]
            g = int(rng.integers(pc["gap"][0], pc["gap"][1] + 1))
            L1 = int(rng.integers(short, RL + 1))
            F = L1 + g + RL
            s0 = p - 16
            if s0 + F > chroms[c].size:
                continue
            fr = chroms[c][s0:s0 + F].copy()
            fr[fr == ord("C")] = ord("T")
            s1[i], s2[i] = bytes(fr[:L1]), bytes(synth.revcomp(fr[F - RL:]))
    if pc["flip"]:
        s1, s2 = s2, s1
    return s1, s2


def _flat(seqs):
    lens = np.array([len(s) for s in seqs], np.uint32)
    offs = np.zeros(len(seqs), np.uint64)
    offs[1:] = np.cumsum(lens.astype(np.uint64))[:-1]
    return np.frombuffer(b"".join(seqs), np.uint8), offs, lens


def gxx_version():
    r = subprocess.run(["g++", "--version"], capture_output=True, text=True, check=True)
    return r.stdout.splitlines()[0]


def run_pair_case(name):
    pc = PAIRED[name]
    share = pc["share"]
    gcfg = dict(CASES[share])
    gcfg["chunk"] = pc["chunk"]
    out = os.path.join(HERE, name)
    os.makedirs(out, exist_ok=True)
    binp = ref_binary(gcfg, pc["minpd"], pc["maxpd"])
    with tempfile.TemporaryDirectory() as td:
        sd = os.path.join(HERE, CASES[share].get("share") or share)
        for f in ("index.bin", "index.bin_strands"):
            with gzip.open(os.path.join(sd, f + ".gz"), "rb") as a, open(os.path.join(td, f), "wb") as b:
                shutil.copyfileobj(a, b)
        chroms = make_genome(CASES[CASES[share].get("share") or share])
        s1, s2 = make_pair_reads(pc, gcfg, chroms)
        write_fastq(os.path.join(td, "r1.fq"), s1)
        write_fastq(os.path.join(td, "r2.fq"), s2)
        args = [binp, "match_pe", "index.bin", "r1.fq", "r2.fq", "out"]
        args += ["--unord_reads"] if pc["unord"] else []
        subprocess.run(args, cwd=td, check=True, stdout=subprocess.DEVNULL)
        raw = open(os.path.join(td, "out_pairs.bin"), "rb").read()
        assert raw[:8] == b"FAMEPE01"
        n = int(np.frombuffer(raw[8:16], np.uint64)[0])
        v1 = np.frombuffer(raw[16:16 + n], np.uint8).copy()
        v2 = np.frombuffer(raw[16 + n:16 + 2 * n], np.uint8).copy()
        m1 = np.frombuffer(raw[16 + 2 * n:16 + 10 * n], np.uint64).copy()
        m2 = np.frombuffer(raw[16 + 10 * n:16 + 18 * n], np.uint64).copy()
        meth = parse_tsv(os.path.join(td, "out_cpg.tsv"))
        stats = json.load(open(os.path.join(td, "out_stats.json")))
        gz(os.path.join(td, "out_cpg.tsv"), os.path.join(out, "cpg.tsv.gz"))
    f1, o1, l1 = _flat(s1)
    f2, o2, l2 = _flat(s2)
    np.savez_compressed(os.path.join(out, "reads.npz"), seqs1=f1, offs1=o1, lens1=l1,
                        seqs2=f2, offs2=o2, lens2=l2)
    np.savez_compressed(os.path.join(out, "golden.npz"), valid1=v1, valid2=v2, mat1=m1, mat2=m2,
                        meth=meth)
    meta = {k: v for k, v in pc.items()}
    meta["mode"] = "pe"
    meta["readlen"] = gcfg["readlen"]
    meta["index_dir"] = CASES[share].get("share") or share
    meta["reference"] = {k: stats[k] for k in ("reads", "succ", "nonunique", "unsucc", "succ_paired",
                                                "too_short", "r1_fwd", "r1_rev", "match_r1_fwd",
                                                "htabsize", "readlen", "chunksize")}
    # std::sort's order of equal counts decides ties in this path
    meta["libstdcxx"] = gxx_version()
    json.dump(meta, open(os.path.join(out, "case.json"), "w"), indent=1)
    print(name, meta["reference"])


SC_CASES = {
    # single-cell runs: cells cut from an existing fixture's reads, matched
    # against its index by FAME's single-cell driver
    "sc_se": dict(reads="se_ga_chunked", cells=[("cell_A", 0, 1000), ("cell_B", 1000, 1900),
                                                ("cell_C", 1900, 3000)], pe=False, unord=False),
    "sc_pe": dict(reads="pe_basic", cells=[("p1", 0, 1000), ("p2", 1000, 2200), ("p3", 2200, 3000)], pe=True,
                  unord=False),
}


def run_sc_case(name):
    sc = SC_CASES[name]
    src = json.load(open(os.path.join(HERE, sc["reads"], "case.json")))
    idx_dir = src["index_dir"]
    if sc["pe"]:
        gcfg = dict(CASES[src["share"]])
        gcfg["chunk"] = src["chunk"]
        binp = ref_binary(gcfg, src["minpd"], src["maxpd"])
    else:
        binp = ref_binary(CASES[sc["reads"]])
    out = os.path.join(HERE, name)
    os.makedirs(out, exist_ok=True)
    r = np.load(os.path.join(HERE, sc["reads"], "reads.npz"))
    mates = (("seqs1", "offs1", "lens1"), ("seqs2", "offs2", "lens2")) if sc["pe"] else (("seqs", "offs", "lens"),)
    with tempfile.TemporaryDirectory() as td:
        for f in ("index.bin", "index.bin_strands"):
            with gzip.open(os.path.join(HERE, idx_dir, f + ".gz"), "rb") as a, open(os.path.join(td, f), "wb") as b:
                shutil.copyfileobj(a, b)
        lines = []
        for cid, c0, c1 in sc["cells"]:
            paths = []
            for k, (sn, on, ln) in enumerate(mates):
                flat, offs, lens = r[sn].tobytes(), r[on], r[ln]
                seqs = [flat[int(o):int(o) + int(n)] for o, n in zip(offs[c0:c1], lens[c0:c1])]
                fn = f"{cid}_{k + 1}.fq.gz"
                write_fastq(os.path.join(td, "tmp.fq"), seqs)
                gz(os.path.join(td, "tmp.fq"), os.path.join(out, fn))
                shutil.copy(os.path.join(out, fn), os.path.join(td, fn))
                paths.append(fn)
            lines.append(" ".join([cid] + paths))
        with open(os.path.join(out, "cells.txt"), "w") as f:
            f.write("\n".join(lines) + "\n")
        shutil.copy(os.path.join(out, "cells.txt"), os.path.join(td, "cells.txt"))
        args = [binp, "sc", "index.bin", "cells.txt", "sc_out.txt", "--gz"]
        args += (["--pe"] if sc["pe"] else []) + (["--unord_reads"] if sc["unord"] else [])
        subprocess.run(args, cwd=td, check=True, stdout=subprocess.DEVNULL)
        gz(os.path.join(td, "sc_out.txt"), os.path.join(out, "sc_out.txt.gz"))
    meta = {"mode": "sc", "pe": sc["pe"], "unord": sc["unord"], "index_dir": idx_dir, "reads": sc["reads"],
            "chunk": src["chunk"], "cells": [c[0] for c in sc["cells"]]}
    if sc["pe"]:
        meta["minpd"], meta["maxpd"] = src["minpd"], src["maxpd"]
    json.dump(meta, open(os.path.join(out, "case.json"), "w"), indent=1)
    print(name, meta)


def _fastq_inputs():
    """name -> (bytes, gz): FASTQ layouts the parser must read as FAME does."""
    r = np.load(os.path.join(HERE, "se_basic", "reads.npz"))
    flat, offs, lens = r["seqs"].tobytes(), r["offs"], r["lens"]
    seqs = [flat[int(o):int(o) + int(n)] for o, n in zip(offs, lens)]

    def fq(ss, nl=b"\n", tail=True):
        recs = [b"@r%d%s%s%s+%s%s" % (i, nl, x, nl, nl, b"I" * len(x)) for i, x in enumerate(ss)]
        return nl.join(recs) + (nl if tail else b"")

    out = {
        "lf": (fq(seqs[:2000]), False),
        "crlf": (fq(seqs[:300], b"\r\n"), False),
        "notrail": (fq(seqs[:50], tail=False), False),
        "exact1400": (fq(seqs[:1400]), False),
        "trunc_seq": (fq(seqs[:20]) + b"@cut\n", False),
        "trunc_nonl": (fq(seqs[:20]) + b"@cut", False),
        "trunc_plus": (fq(seqs[:20]) + b"@cut\nACGT\n+\n", False),
        "blank": (fq(seqs[:10]) + b"\n\n" + fq(seqs[10:30]) + b"\n", False),
        "empty": (b"", False),
        "gz_lf": (fq(seqs[:2000]), True),
        "gz_notrail": (fq(seqs[:707], tail=False), True),
    }
    return out


def run_fastq_cases():
    cfg = CASES["se_ga_chunked"]
    binp = ref_binary(cfg)
    out = os.path.join(HERE, "fastq")
    os.makedirs(out, exist_ok=True)
    with tempfile.TemporaryDirectory() as td:
        for f in ("index.bin", "index.bin_strands"):
            with gzip.open(os.path.join(HERE, "se_ga_chunked", f + ".gz"), "rb") as a, \
                    open(os.path.join(td, f), "wb") as b:
                shutil.copyfileobj(a, b)
        for name, (data, is_gz) in _fastq_inputs().items():
            src = os.path.join(td, name + ".fq")
            with open(src, "wb") as f:
                f.write(data)
            # the input travels gzipped either way; gz cases are read as such
            stored = os.path.join(out, name + ".fq.gz")
            with gzip.open(stored, "wb", compresslevel=9) as f:
                f.write(data)
            inp = stored if is_gz else src
            subprocess.run([binp, "parse", "index.bin", inp, "out.bin"] + (["--gz"] if is_gz else []), cwd=td,
                           check=True, stdout=subprocess.DEVNULL)
            raw = open(os.path.join(td, "out.bin"), "rb").read()
            assert raw[:8] == b"FAMEFQ01"
            n = int(np.frombuffer(raw[8:16], np.uint64)[0])
            lens = np.frombuffer(raw[16:16 + 4 * n], np.uint32).copy()
            p = 16 + 4 * n
            tot = int(lens.sum(dtype=np.uint64))
            seqs = np.frombuffer(raw[p:p + tot], np.uint8).copy()
            p += tot
            nc = int(np.frombuffer(raw[p:p + 8], np.uint64)[0])
            p += 8
            calls = np.array([(int(np.frombuffer(raw[p + 5 * i:p + 5 * i + 4], np.uint32)[0]), raw[p + 5 * i + 4])
                              for i in range(nc)], np.uint32).reshape(-1, 2)
            np.savez_compressed(os.path.join(out, name + ".expect.npz"), lens=lens, seqs=seqs, calls=calls)
            print("fastq", name, n, "reads", calls.tolist())
        json.dump({"mode": "fastq", "chunksize": cfg["chunk"],
                   "gz": [k for k, v in _fastq_inputs().items() if v[1]]},
                  open(os.path.join(out, "case.json"), "w"), indent=1)


if __name__ == "__main__":
    names = sys.argv[1:] or list(CASES) + list(PAIRED) + ["fastq"] + list(SC_CASES)
    for nm in names:
        if nm == "fastq":
            run_fastq_cases()
        elif nm in SC_CASES:
            run_sc_case(nm)
        else:
            run_pair_case(nm) if nm in PAIRED else run_case(nm)
